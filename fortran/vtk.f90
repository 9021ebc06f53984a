! vtk.f90 -- drop-in for module vtk (src/vtk.F90): VtkOutput writes the
! sampled spline as a text VTK ImageData file, <filename>.vti, with the
! reference's layout (header lines, "0 <I5> 0 <I5> 0 <I5>" extents, one
! F30.10 value per line, x fastest).
module vtk
   use plot, only: PlotParams
   implicit none
   private
   public :: VtkOutput

contains

   subroutine VtkOutput(filename, vals, params)
      character(len=*), intent(in) :: filename
      type(PlotParams), intent(in) :: params
      real(kind=8), intent(in) :: vals(params%resx, params%resy, params%resz)
      integer(kind=4) :: ix, iy, iz, u
      character(len=200) :: line, extent

      open(newunit=u, file=trim(filename)//'.vti', form='formatted', access='sequential', status='replace')
      write(u, '(a)') '<?xml version="1.0"?>'
      write(u, '(a)') '<VTKFile type="ImageData" version="0.1">'
      write(extent, '("0 ",(i5)," 0 ",(i5)," 0 ",(i5))') params%resx - 1, params%resy - 1, params%resz - 1
      write(u, '(a)') '  <ImageData WholeExtent="'//trim(extent)//'" origin="0 0 0" spacing="1 1 1">'
      write(u, '(a)') '    <Piece Extent="'//trim(extent)//'">'
      write(u, '(a)') '      <PointData Scalars="Result">'
      write(u, '(a)') '        <DataArray Name="Result" type="Float32">'
      do iz = 1, params%resz
         do iy = 1, params%resy
            do ix = 1, params%resx
               write(line, '(f30.10)') vals(ix, iy, iz)
               write(u, '(a)') '          '//trim(adjustl(line))
            end do
         end do
      end do
      write(u, '(a)') '        </DataArray>'
      write(u, '(a)') '      </PointData>'
      write(u, '(a)') '    </Piece>'
      write(u, '(a)') '  </ImageData>'
      write(u, '(a)') '</VTKFile>'
      close(u)
   end subroutine VtkOutput

end module vtk
