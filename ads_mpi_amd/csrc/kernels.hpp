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
};

enum Form { kFormHeat = 0, kFormProject = 1 };

// F = sum over elements of the element integrals (Form3DRHS after FormUn).
//   kFormHeat:    F = A(kappa) c, A = M - dt K(kappa)    (c: coefficients, FP64)
//   kFormProject: F = int f0 v_a                         (f0 kind in f0)
// F must be zero on entry (the kernel accumulates).  l2_part receives one
// partial sum per workgroup (reduced by reduce_partials).
template <typename T>
hipError_t launch_assemble(int p, int form, int f0, const double* c, T* F, const Tables<T>& tb, double dt,
                           const double* kappa, double* l2_part, int* nparts, hipStream_t s);

// dL/dc_n and dL/dkappa of one step given mu = M^-T lam:
//   lam_out = A(kappa) mu,  dkappa[e] += -dt sum_q J W grad c_n . grad mu
hipError_t launch_assemble_adjoint(int p, const double* c_n, const double* mu, double* lam_out,
                                   const Tables<double>& tb, double dt, const double* kappa, double* dkappa,
                                   hipStream_t s);

// Deterministic sum of n partials into out[0].
hipError_t launch_reduce_partials(const double* part, int n, double* out, hipStream_t s);

// The three 1-D solves along x, y, z of the x-fastest array F (in place):
// F <- (Mz (x) My (x) Mx)^-1 F, or with M^T when trans.
template <typename T>
hipError_t launch_kron_solve(int p, T* F, const SweepTables<T> sw[3], bool trans, hipStream_t s);

// r = A c - M x in FP64 with the 1-D banded matrices (constant kappa):
// A = M - dt*kappa*K  where M = Mz(x)My(x)Mx, K = Mz(x)My(x)Kx + Mz(x)Ky(x)Mx + Kz(x)My(x)Mx.
// x is read in precision TX (x == nullptr means x = 0); r is written in
// precision TR.
struct BandSet {
    const double* mass[3];   // row tables: [i][t] = M(i, i-p+t), t = 0..2p, zero outside
    const double* stiff[3];  // [i][t] = K(i, i-p+t)
    int n[3];
};
template <typename TX, typename TR>
hipError_t launch_kron_residual(int p, const double* c, const TX* x, TR* r, const BandSet& bs, double dtk,
                                double* work0, double* work1, double* work2, hipStream_t s);

// out[0] = sum_i a_i b_i (deterministic: per-block partials, then one block)
hipError_t launch_dot(const double* a, const double* b, int64_t n, double* part, int part_cap, double* out,
                      hipStream_t s);

// elementwise helpers
hipError_t launch_convert_d2f(const double* in, float* out, int64_t n, hipStream_t s);
hipError_t launch_add_f2d(const float* a, const float* b, double* out, int64_t n, hipStream_t s);
hipError_t launch_check_finite(const double* v, int* flag, hipStream_t s);

}  // namespace ads
