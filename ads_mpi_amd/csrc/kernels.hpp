// kernels.hpp -- launch interface of the gfx950 kernels (kernels.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace ads {

// Device copies of one direction's tables, in the working precision T.
template <typename T>
struct DirTables {
    const int* first_dof;  // O(e)
    const T* tab;          // NN at [((e*q + iq)*P + a)*2 + k]
    const T* jac;          // J(e)
    const T* wts;          // W(iq)
    const T* pts;          // X(iq, e) at [e*q + iq]
    const T* emat;         // element matrices by rows [((e*P + a)*2 + k)*P + b]
    const T* emat_sym;     // packed upper triangles [(e*2 + k)*P(P+1)/2 + t]
    int nelem;
    int dofs;
};

template <typename T>
struct Tables {
    DirTables<T> d[3];
};

// Sweep tables of one direction (see setup.hpp sweep_tables), precision T.
template <typename T>
struct SweepTables {
    const T* fwd;    // [n][p]
    const int* piv;  // [n]
    const T* bwd;    // [n][2p+1]
    int n;
    int pivots;      // nonzero when the factorisation has row interchanges
};

enum Form { kFormHeat = 0, kFormProject = 1 };

// F = sum over elements of the element integrals (Form3DRHS after FormUn).
//   kFormHeat:    F = A(kappa) c, A = M - dt K(kappa)    (c: coefficients, FP64)
//   kFormProject: F = int f0 v_a                         (f0 kind in f0)
// F must be zero on entry (the kernel accumulates).  l2_part receives one
// partial sum per workgroup (reduced by reduce_partials); their count goes
// to *nparts, and a launch that would need more than part_cap is refused
// (hipErrorInvalidValue) before anything runs.
template <typename T>
hipError_t launch_assemble(int p, int form, int f0, const double* c, T* F, const Tables<T>& tb, double dt,
                           const double* kappa, double* l2_part, int part_cap, int* nparts, hipStream_t s);

// Adjoint assembly given mu = M^-T lam:  lam_out = A(kappa) mu (A symmetric)
// and, when dkappa != nullptr, dkappa[e] += -dt c_n,e^T K_e mu_e in the same
// pass (dkappa accumulates in FP64 for both T).  lam_out must be zero on
// entry; l2_part as in launch_assemble.
template <typename T>
hipError_t launch_heat_adjoint(int p, const double* mu, const double* c_n, T* lam_out, const Tables<T>& tb, double dt,
                               const double* kappa, double* dkappa, double* l2_part, int part_cap, int* nparts,
                               hipStream_t s);

// The spline with coefficients c (x fastest, sx*sy*sz) at the tensor grid
// given by per-direction first-index and basis-value tables (p+1 values per
// point; zero past a lower-degree direction's own); out x fastest (rx*ry*rz).
hipError_t launch_eval_grid(int p, const double* c, int sx, int sy, int sz, const int* fx, const double* bx, int rx,
                            const int* fy, const double* by, int ry, const int* fz, const double* bz, int rz,
                            double* out, hipStream_t s);

// MultiStep's heat right-hand side: F = M c_val - dt kappa K c_grad (point
// values from c_val, gradients from c_grad); l2 from c_val.  F zero on entry.
hipError_t launch_heat_dual(int p, const double* c_val, const double* c_grad, double* F, const Tables<double>& tb,
                            double dt, const double* kappa, double* l2_part, int part_cap, int* nparts, hipStream_t s);

// Deterministic sum of n partials into out[0].
hipError_t launch_reduce_partials(const double* part, int n, double* out, hipStream_t s);

// The three 1-D solves along x, y, z of the x-fastest array F (in place):
// F <- (Mz (x) My (x) Mx)^-1 F, or with M^T when trans.
template <typename T>
hipError_t launch_kron_solve(int p, T* F, const SweepTables<T> sw[3], bool trans, hipStream_t s);

// The 1-D solves of nrows contiguous lines of st.n entries each (line r at
// F + r*st.n), in place: F <- A^-1 F (A^-T when trans) for the factors in st.
hipError_t launch_rows_solve(int p, double* F, int nrows, const SweepTables<double>& st, bool trans, hipStream_t s);

// r = b + A c - M x in FP64 with the 1-D banded matrices (constant kappa):
// A = M - dt*kappa*K  where M = Mz(x)My(x)Mx, K = Mz(x)My(x)Kx + Mz(x)Ky(x)Mx + Kz(x)My(x)Mx.
// c, x and b may each be nullptr (zero).  x is read in precision TX; r is
// written in precision TR.  With l2_part != nullptr each workgroup also
// writes its partial sum of c^T M c; their count goes to *nparts (for
// launch_reduce_partials), and a launch that would need more than part_cap
// is refused before anything runs.  With bo != nullptr (x = nullptr only) r
// is also written to bo in FP64.  Without c (x given) only the M products are formed: the mass
// residual r = b - M x (no l2 then).
struct BandSet {
    const double* mass[3];   // row tables: [i][t] = M(i, i-p+t), t = 0..2p, zero outside
    const double* stiff[3];  // [i][t] = K(i, i-p+t)
    int n[3];
};
template <typename TX, typename TR>
hipError_t launch_kron_residual(int p, const double* c, const TX* x, const double* b, TR* r, const BandSet& bs,
                                double dtk, double* l2_part, int part_cap, int* nparts, hipStream_t s,
                                double* bo = nullptr);
// The oil problem's inputs on the device (problems/oil/input_data.F90):
// curve points (cx, cy, cz of cN curves x cL points, consecutive), pumps and
// drains (x, y, z each) and the scalar constants.
struct OilDev {
    const double* curves;
    int cN, cL;
    const double* pumps;
    int npumps;
    const double* drains;
    int ndrains;
    double mi, kq_min, kq_max, radius, pump_strength, drain_strength;
};
// kq at every Gauss point, [kx + Q (ky + Q (kz + Q e))] (CacheKqValues)
hipError_t launch_oil_kq(int p, const Tables<double>& tb, const OilDev& o, double* kq, hipStream_t s);
// the oil right-hand side: t > 0 from c (tzero false) or the projection of
// the initial state; per-workgroup partial sums of l2 and drained
hipError_t launch_oil(int p, bool tzero, const double* c, double* F, const Tables<double>& tb, const OilDev& o,
                      const double* kq, double dt, double* l2_part, double* dr_part, int part_cap, int* nparts,
                      hipStream_t s);
// The FP32 solve of the segmented refinement form: out = (Mz (x) My (x)
// Mx)^-1 in up to the truncation below, direction by direction, each line cut
// into segments solved independently from w entries on either side (see
// k_seg_lines).  Per direction, the factors' rows (2p+1 floats each:
// L(j, j-1..j-p), 1/U(j, j), U(j, j+1..j+p)) at rows[j*(2p+1)], zero padded
// for j in [-kSegPadLo, 0) and [n, n + kSegPadHi); the middle row at mid,
// which rows [lo, hi) equal; w a multiple of 16 up to kSegPadLo.  in is left
// as it is, tmp is scratch of the same size; the factors must have no row
// interchanges.
constexpr int kSegMax = 32;      // outputs per lane at most
constexpr int kSegPadLo = 48;    // largest w
constexpr int kSegPadHi = kSegMax + kSegPadLo;
struct SegLU {
    const float* mid;
    const float* rows;
    int n, lo, hi, w;
    int seg;  // outputs per lane (set at launch)
};
// The result goes to out, or, with fo.x set, to fo.x in FP64 (out then
// holds the x-direction intermediate): kSegSet x = v, kSegAdd x = fo.add + v,
// kSegAccumulate x += v.
enum SegMode { kSegSet = 0, kSegAdd = 1, kSegAccumulate = 2 };
struct SegOut {
    double* x;
    const float* add;
    int mode;
};
// (TI = float: the one instance.)
template <typename TI>
hipError_t launch_seg_solve(int p, const TI* in, float* tmp, float* out, const SegLU sl[3], const SegOut& fo,
                            hipStream_t s);

// upper bound of the workgroup (partial-sum) count of launch_kron_residual
long kron_residual_partials_bound(const BandSet& bs);

// Workgroups of `fn` (threads each, lds dynamic bytes) resident at once on
// the current device: occupancy per CU times the CU count.
int resident_groups(const void* fn, int threads, size_t lds);

// out = float(in)
hipError_t launch_convert_d2f(const double* in, float* out, int64_t n, hipStream_t s);
// out = (a ? a : out) + (b ? b : 0), FP32 operands widened to FP64
hipError_t launch_add_f2d(const float* a, const float* b, double* out, int64_t n, hipStream_t s);

}  // namespace ads
