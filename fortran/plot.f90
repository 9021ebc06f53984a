! plot.f90 -- drop-in for module plot (src/plot.F90): the plot parameter
! type that PrintSolution passes to the writer.  Sampling happens on the GPU
! (ads_eval_grid), so SaveSplinePlot's loops are not repeated here.
module plot
   implicit none
   private
   public :: PlotParams

   ! type PlotParams (src/plot.F90:14-22): ranges and sample counts
   type PlotParams
      real(kind=8) :: startx, endx
      real(kind=8) :: starty, endy
      real(kind=8) :: startz, endz
      integer(kind=4) :: resx, resy, resz
   end type PlotParams
end module plot
