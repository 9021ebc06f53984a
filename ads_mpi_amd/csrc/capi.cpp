// capi.cpp -- the C ABI (include/ads_mi355x.h): context, device state,
// stream ordering and the step orchestration (Step / Sub_Step of
// src/ADS.F90:437-794 on one GPU).
#include "ads_mi355x.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <utility>
#include <vector>

#include "bspline.hpp"
#include "kernels.hpp"
#include "part.hpp"
#include "setup.hpp"

namespace {

thread_local std::string g_init_error;

template <typename T>
struct DevDir {
    int* first_dof = nullptr;
    T* tab = nullptr;
    T* jac = nullptr;
    T* wts = nullptr;
    T* pts = nullptr;
    T* emat = nullptr;
    T* emat_sym = nullptr;
    T* fwd = nullptr;
    int* piv = nullptr;
    T* bwd = nullptr;
};

// Device factor tables of one MultiStep operator mix(1)M + mix(2)K +
// mix(3)B + mix(4)B^T (the same operator in x, y and z).
struct MixSet {
    double mix[4];
    ads::SweepTables<double> sw[3];
};

}  // namespace

struct ads_ctx {
    std::string err;
    int p = 0;               // degree of the kernels: the largest of the three
    ads::Direction dir[3];   // each direction with its own degree (host tables)
    ads::Direction kdir[3];  // the same widened to degree p (the kernels' tables)
    long D = 0;
    long E = 0;
    hipStream_t own = nullptr;
    hipStream_t stream = nullptr;

    double* c = nullptr;      // state coefficients (x fastest)
    double* F = nullptr;      // FP64 right-hand side / solve buffer
    double* kappa = nullptr;  // per-element diffusivity or nullptr (kappa = 1)
    double* l2_part = nullptr;
    int l2_cap = 0;
    double* l2 = nullptr;  // device scalar: last l2norm

    DevDir<double> d64[3];
    DevDir<float> d32[3];
    double* knots[3] = {nullptr, nullptr, nullptr};  // U(0:n+p+1) per direction
    // the partitioned FP64 solve of M (part.hpp), when every direction's
    // lines fit its kernels; solver: ADS_SOLVER_*
    ads::PartDev<double> part[3]{};
    bool part_ok = false;
    int solver = ADS_SOLVER_AUTO;
    // FormUn: a copy of the coefficients kept by ads_form_un(1) (U_n) and the
    // state version it was taken at (version counts the changes of c)
    double* un = nullptr;
    bool un_kept = false;
    long un_version = -1;
    long version = 0;
    void* grid_scratch = nullptr;  // this device's static scratch (ads_eval_grid_noalloc_d)
    double* mass[3] = {nullptr, nullptr, nullptr};
    double* stiff[3] = {nullptr, nullptr, nullptr};
    ads::Tables<double> tb64{};
    ads::Tables<float> tb32{};
    ads::SweepTables<double> sw64[3]{};
    ads::SweepTables<float> sw32[3]{};

    // lazily allocated scratch
    float* f32a = nullptr;
    float* f32b = nullptr;
    double* x64 = nullptr;
    double* b64 = nullptr;
    double* mu = nullptr;

    std::vector<MixSet> mixsets;  // MultiStep operators factored so far

    // segmented form: FP32 factor tables per direction (built on first use)
    ads::SegLU seg[3]{};
    float* f32c = nullptr;
    float* f32d = nullptr;

    // the oil problem (ads_oil_setup)
    struct Oil {
        double* curves = nullptr;
        long curves_cap = 0;
        double* pumps = nullptr;
        long pumps_cap = 0;
        double* drains = nullptr;
        long drains_cap = 0;
        double* kq = nullptr;       // kq at every Gauss point
        double* dr_part = nullptr;  // per-workgroup drained partial sums
        double* drained = nullptr;  // device scalar: the last step's drained
        ads::OilDev dev{};
        bool ready = false;
    } oil;

    std::vector<void*> allocs;
};

// One 1-D matrix of ComputeMatrix (MKBBT on the caller's knots) with its
// banded LU; the device copies of its sweep tables and a right-hand-side
// buffer are made on the first solve.
struct ads_matrix {
    std::string err;
    ads::Direction dir;
    double mix[4] = {0.0, 0.0, 0.0, 0.0};
    std::vector<double> band, lu;
    std::vector<int> piv;
    hipStream_t stream = nullptr;
    double* fwd = nullptr;
    double* bwd = nullptr;
    int* dpiv = nullptr;
    int pivots = 0;
    double* rhs = nullptr;
    size_t rhs_cap = 0;
};

namespace {

int fail(ads_ctx* ctx, int code, const std::string& msg)
{
    if (ctx) ctx->err = msg;
    else g_init_error = msg;
    return code;
}

int hip_fail(ads_ctx* ctx, hipError_t e, const char* what)
{
    return fail(ctx, ADS_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

#define ADS_HIP(ctx, call)                                        \
    do {                                                          \
        hipError_t e_ = (call);                                   \
        if (e_ != hipSuccess) return hip_fail((ctx), e_, #call);  \
    } while (0)

template <typename T>
int dev_alloc(ads_ctx* ctx, T** out, size_t count)
{
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, sizeof(T) * (count ? count : 1));
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return fail(ctx, ADS_ERR_NOMEM, std::string("hipMalloc of ") + std::to_string(sizeof(T) * count) +
                                            " bytes failed: " + hipGetErrorString(e));
    }
    ctx->allocs.push_back(p);
    *out = static_cast<T*>(p);
    return ADS_OK;
}

template <typename T, typename S>
int upload(ads_ctx* ctx, T** out, const std::vector<S>& host)
{
    std::vector<T> tmp(host.begin(), host.end());
    int rc = dev_alloc(ctx, out, tmp.size());
    if (rc) return rc;
    ADS_HIP(ctx, hipMemcpy(*out, tmp.data(), sizeof(T) * tmp.size(), hipMemcpyHostToDevice));
    return ADS_OK;
}

template <typename T>
int upload_dir(ads_ctx* ctx, const ads::Direction& d, DevDir<T>& dd, ads::DirTables<T>& dt, ads::SweepTables<T>& st)
{
    int rc;
    if ((rc = upload(ctx, &dd.first_dof, d.first_dof))) return rc;
    if ((rc = upload(ctx, &dd.tab, d.tab))) return rc;
    if ((rc = upload(ctx, &dd.jac, d.jac))) return rc;
    if ((rc = upload(ctx, &dd.wts, d.wts))) return rc;
    if ((rc = upload(ctx, &dd.pts, d.pts))) return rc;
    if ((rc = upload(ctx, &dd.emat, d.emat))) return rc;
    if ((rc = upload(ctx, &dd.emat_sym, d.emat_sym))) return rc;
    std::vector<double> fwd, bwd;
    std::vector<int> piv;
    ads::sweep_tables(d, fwd, piv, bwd);
    if ((rc = upload(ctx, &dd.fwd, fwd))) return rc;
    if ((rc = upload(ctx, &dd.piv, piv))) return rc;
    if ((rc = upload(ctx, &dd.bwd, bwd))) return rc;
    dt.first_dof = dd.first_dof;
    dt.tab = dd.tab;
    dt.jac = dd.jac;
    dt.wts = dd.wts;
    dt.pts = dd.pts;
    dt.emat = dd.emat;
    dt.emat_sym = dd.emat_sym;
    dt.nelem = d.nelem;
    dt.dofs = d.dofs;
    st.fwd = dd.fwd;
    st.piv = dd.piv;
    st.bwd = dd.bwd;
    st.n = d.dofs;
    st.pivots = 0;
    for (int v : piv) st.pivots |= (v != 0);
    return ADS_OK;
}

int ensure_ir_scratch(ads_ctx* ctx)
{
    int rc;
    if (!ctx->f32a && (rc = dev_alloc(ctx, &ctx->f32a, ctx->D))) return rc;
    if (!ctx->f32b && (rc = dev_alloc(ctx, &ctx->f32b, ctx->D))) return rc;
    if (!ctx->x64 && (rc = dev_alloc(ctx, &ctx->x64, ctx->D))) return rc;
    if (!ctx->b64 && (rc = dev_alloc(ctx, &ctx->b64, ctx->D))) return rc;
    return ADS_OK;
}

// the segmented form's factor tables, built on first use: rows matching the
// middle one to 1e-7 use it, and the warm-up w covers a decay to 1e-8
constexpr double kSegTol = 1e-7, kSegTolW = 1e-8;

int ensure_seg(ads_ctx* ctx)
{
    int rc;
    if (!ctx->f32c && (rc = dev_alloc(ctx, &ctx->f32c, ctx->D))) return rc;
    if (!ctx->f32d && (rc = dev_alloc(ctx, &ctx->f32d, ctx->D))) return rc;
    for (int d = 0; d < 3; ++d) {
        if (ctx->seg[d].mid) continue;
        std::vector<float> tab;
        int w = 0, lo = 0, hi = 0;
        const std::string e =
            ads::seg_lu_tables(ctx->kdir[d], kSegTol, kSegTolW, ads::kSegPadLo, ads::kSegMax, tab, w, lo, hi);
        if (!e.empty()) return fail(ctx, ADS_ERR_UNSUPPORTED, "segmented form, direction " + std::to_string(d) + ": " + e);
        // device layout: the middle row, kSegPadLo zero rows, rows 0..n-1, kSegPadHi zero rows
        const int n = ctx->dir[d].dofs, nt = 2 * ctx->p + 1;
        std::vector<float> dev_tab(static_cast<size_t>(1 + ads::kSegPadLo + n + ads::kSegPadHi) * nt, 0.0f);
        std::copy(tab.begin() + static_cast<long>(n) * nt, tab.end(), dev_tab.begin());
        std::copy(tab.begin(), tab.begin() + static_cast<long>(n) * nt,
                  dev_tab.begin() + static_cast<long>(1 + ads::kSegPadLo) * nt);
        float* dev = nullptr;
        if ((rc = dev_alloc(ctx, &dev, dev_tab.size()))) return rc;
        ADS_HIP(ctx, hipMemcpy(dev, dev_tab.data(), sizeof(float) * dev_tab.size(), hipMemcpyHostToDevice));
        ctx->seg[d] = ads::SegLU{dev, dev + static_cast<long>(1 + ads::kSegPadLo) * nt, n, lo, hi, w, 0};
    }
    return ADS_OK;
}

int seg_solve(ads_ctx* ctx, const float* in, float* tmp, float* out, const ads::SegOut& fo = ads::SegOut{})
{
    ADS_HIP(ctx, ads::launch_seg_solve(ctx->p, in, tmp, out, ctx->seg, fo, ctx->stream));
    return ADS_OK;
}

// The device tables of the partitioned solve for every direction, when the
// lines fit its kernels (at most kPartMaxK partitions of kPartSeg entries, or
// of kPartSegLong for lines longer than kPartMaxK * kPartSeg).
int build_part(ads_ctx* ctx)
{
    ctx->part_ok = false;
    if (!ads::part_planes_fit(ctx->kdir[0].dofs, ctx->kdir[1].dofs, sizeof(double))) return ADS_OK;  // LU sweeps
    for (int d = 0; d < 3; ++d) {
        const int seg = ads::part_seg(ctx->kdir[d].dofs);
        if (seg == 0) return ADS_OK;  // the LU sweeps serve these lines
        ads::PartHost ph;
        const std::string e = ads::part_tables(ctx->kdir[d], seg, ph);
        if (!e.empty()) return fail(ctx, ADS_ERR_SINGULAR, "direction " + std::to_string(d) + ": " + e);
        if (ph.K > ads::kPartMaxK) return ADS_OK;
        double *lr = nullptr, *ur = nullptr, *gs = nullptr, *hs = nullptr;
        int rc;
        if ((rc = upload(ctx, &lr, ph.lrow)) || (rc = upload(ctx, &ur, ph.urow)) || (rc = upload(ctx, &gs, ph.gsp)) ||
            (rc = upload(ctx, &hs, ph.hsp)))
            return rc;
        ctx->part[d] = ads::PartDev<double>{lr, ur, gs, hs, ph.n, ph.K, ph.S};
    }
    ctx->part_ok = true;
    return ADS_OK;
}

// F <- M^-1 F in FP64 (M^-T: the same, M is symmetric; the LU sweeps apply
// dgbtrs 'T'), by the partitioned solve or the LU sweeps (ads_set_solver).
int mass_solve64(ads_ctx* ctx, double* F, bool trans)
{
    const bool part = ctx->solver == ADS_SOLVER_PARTITIONED || (ctx->solver == ADS_SOLVER_AUTO && ctx->part_ok);
    if (part) ADS_HIP(ctx, ads::launch_part_solve<double>(ctx->p, F, ctx->part, ctx->stream));
    else ADS_HIP(ctx, ads::launch_kron_solve<double>(ctx->p, F, ctx->sw64, trans, ctx->stream));
    return ADS_OK;
}

__global__ void k_fill_zero_d(double* p, long n)
{
    const long i = blockIdx.x * static_cast<long>(blockDim.x) + threadIdx.x;
    if (i < n) p[i] = 0.0;
}
__global__ void k_fill_zero_f(float* p, long n)
{
    const long i = blockIdx.x * static_cast<long>(blockDim.x) + threadIdx.x;
    if (i < n) p[i] = 0.0f;
}

int zero(ads_ctx* ctx, double* p, long n)
{
    hipLaunchKernelGGL(k_fill_zero_d, dim3(static_cast<unsigned>((n + 255) / 256)), dim3(256), 0, ctx->stream, p, n);
    ADS_HIP(ctx, hipGetLastError());
    return ADS_OK;
}
int zero(ads_ctx* ctx, float* p, long n)
{
    hipLaunchKernelGGL(k_fill_zero_f, dim3(static_cast<unsigned>((n + 255) / 256)), dim3(256), 0, ctx->stream, p, n);
    ADS_HIP(ctx, hipGetLastError());
    return ADS_OK;
}

// F = A(kappa) c (heat) or F = int f0 v (project), l2 -> l2_d
template <typename T>
int assemble(ads_ctx* ctx, int form, int f0, const double* c, T* F, double dt, const double* kappa, double* l2_d)
{
    int rc = zero(ctx, F, ctx->D);
    if (rc) return rc;
    int nparts = 0;
    const ads::Tables<T>* tb;
    if constexpr (sizeof(T) == 8) tb = &ctx->tb64;
    else tb = &ctx->tb32;
    ADS_HIP(ctx, ads::launch_assemble<T>(ctx->p, form, f0, c, F, *tb, dt, kappa, ctx->l2_part, ctx->l2_cap, &nparts,
                                         ctx->stream));
    ADS_HIP(ctx, ads::launch_reduce_partials(ctx->l2_part, nparts, l2_d ? l2_d : ctx->l2, ctx->stream));
    return ADS_OK;
}

int check_ctx(const ads_ctx* ctx) { return ctx ? ADS_OK : ADS_ERR_INVALID; }

// grid_tables of direction d in the kernels' layout: ctx->p + 1 values per
// point, zero past the direction's own degree
void grid_tables_k(const ads_ctx* ctx, int d, int res, double lo, double hi, std::vector<int>& first,
                   std::vector<double>& vals)
{
    std::vector<double> own;
    ads::grid_tables(ctx->dir[d], res, lo, hi, first, own);
    const int P = ctx->dir[d].p + 1, PK = ctx->p + 1;
    if (P == PK) {
        vals.swap(own);
        return;
    }
    vals.assign(static_cast<size_t>(res) * PK, 0.0);
    for (int i = 0; i < res; ++i)
        for (int a = 0; a < P; ++a) vals[static_cast<size_t>(i) * PK + a] = own[static_cast<size_t>(i) * P + a];
}

int read_l2(ads_ctx* ctx, double* l2norm)
{
    if (!l2norm) return ADS_OK;
    ADS_HIP(ctx, hipMemcpyAsync(l2norm, ctx->l2, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    ADS_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (!std::isfinite(*l2norm)) return fail(ctx, ADS_ERR_NUMERIC, "l2norm is not finite");
    return ADS_OK;
}

ads::BandSet band_set(const ads_ctx* ctx)
{
    ads::BandSet bs{};
    for (int d = 0; d < 3; ++d) {
        bs.mass[d] = ctx->mass[d];
        bs.stiff[d] = ctx->stiff[d];
        bs.n[d] = ctx->dir[d].dofs;
    }
    return bs;
}

// r = b + A c - M x in FP64 (nullptr operands are zero), rounded to TR;
// with l2_d also l2 = c^T M c -> l2_d; with bo also r in FP64 -> bo.
template <typename TX, typename TR = float>
int residual(ads_ctx* ctx, const double* c, const TX* x, const double* b, TR* r, double dtk, double* l2_d,
             double* bo = nullptr)
{
    const ads::BandSet bs = band_set(ctx);
    int nres = 0;
    ADS_HIP(ctx, (ads::launch_kron_residual<TX, TR>(ctx->p, c, x, b, r, bs, dtk, l2_d ? ctx->l2_part : nullptr,
                                                    ctx->l2_cap, &nres, ctx->stream, bo)));
    if (l2_d) ADS_HIP(ctx, ads::launch_reduce_partials(ctx->l2_part, nres, l2_d, ctx->stream));
    return ADS_OK;
}

// x = M^-1 b (M^-T b when trans) by FP32 sweeps corrected `corrections`
// times with FP64 mass residuals r = b - M x (M is symmetric).  x0, when
// given, is the FP32 first estimate; otherwise it is M^-1 float(b).  x must
// not alias b.
int mass_solve_ir(ads_ctx* ctx, const double* b, double* x, const float* x0, int corrections, bool trans)
{
    const float* d1 = x0;
    if (!d1) {
        ADS_HIP(ctx, ads::launch_convert_d2f(b, ctx->f32a, ctx->D, ctx->stream));
        ADS_HIP(ctx, ads::launch_kron_solve<float>(ctx->p, ctx->f32a, ctx->sw32, trans, ctx->stream));
        d1 = ctx->f32a;
    }
    if (corrections == 0) {
        ADS_HIP(ctx, ads::launch_add_f2d(d1, nullptr, x, ctx->D, ctx->stream));
        return ADS_OK;
    }
    // the first correction reads the FP32 estimate directly
    int rc = residual<float>(ctx, nullptr, d1, b, ctx->f32b, 0.0, nullptr);
    if (rc) return rc;
    ADS_HIP(ctx, ads::launch_kron_solve<float>(ctx->p, ctx->f32b, ctx->sw32, trans, ctx->stream));
    ADS_HIP(ctx, ads::launch_add_f2d(d1, ctx->f32b, x, ctx->D, ctx->stream));
    for (int k = 1; k < corrections; ++k) {
        if ((rc = residual<double>(ctx, nullptr, x, b, ctx->f32b, 0.0, nullptr))) return rc;
        ADS_HIP(ctx, ads::launch_kron_solve<float>(ctx->p, ctx->f32b, ctx->sw32, trans, ctx->stream));
        ADS_HIP(ctx, ads::launch_add_f2d(nullptr, ctx->f32b, x, ctx->D, ctx->stream));
    }
    return ADS_OK;
}

// x = M^-1 b (M is symmetric, so also M^-T b) by segmented FP32 solves: the
// first from x0 when given (an FP32 estimate, in neither f32a, f32b nor
// f32d), else a segmented solve of b, then `corrections` more with FP64 mass
// residuals b - M x, the last z solve of each forming x in FP64.  x must not
// alias b.
int mass_solve_seg(ads_ctx* ctx, const double* b, double* x, const float* x0, int corrections)
{
    int rc = ensure_ir_scratch(ctx);
    if (rc || (rc = ensure_seg(ctx))) return rc;
    float* r = ctx->f32a;
    float* tmp = ctx->f32b;
    float* mid = ctx->f32d;
    const float* d1 = x0;
    if (!d1) {
        float* d = ctx->f32c;
        ADS_HIP(ctx, ads::launch_convert_d2f(b, r, ctx->D, ctx->stream));
        if (corrections == 0) return seg_solve(ctx, r, tmp, d, ads::SegOut{x, nullptr, ads::kSegSet});
        if ((rc = seg_solve(ctx, r, tmp, d))) return rc;
        d1 = d;
    } else if (corrections == 0) {
        ADS_HIP(ctx, ads::launch_add_f2d(d1, nullptr, x, ctx->D, ctx->stream));
        return ADS_OK;
    }
    if ((rc = residual<float>(ctx, nullptr, d1, b, r, 0.0, nullptr))) return rc;
    if ((rc = seg_solve(ctx, r, tmp, mid, ads::SegOut{x, d1, ads::kSegAdd}))) return rc;
    for (int k = 1; k < corrections; ++k) {
        if ((rc = residual<double>(ctx, nullptr, x, b, r, 0.0, nullptr))) return rc;
        if ((rc = seg_solve(ctx, r, tmp, mid, ads::SegOut{x, nullptr, ads::kSegAccumulate}))) return rc;
    }
    return ADS_OK;
}

// c_out = M^-1 A(kappa) c_in: FP32 assembly and a segmented FP32 solve, then
// `refine` corrections with FP64 residuals.  Constant kappa: the residual is
// the Kronecker operator A c - M x (it also gives l2).  Per-element kappa: b =
// A(kappa) c comes from the FP64 element assembly (with l2) and the
// corrections use the mass residual b - M x.
int step_ir(ads_ctx* ctx, const double* c_in, double* c_out, const double* kappa_d, double dt, int refine,
            double* l2_d)
{
    int rc = ensure_ir_scratch(ctx);
    if (rc || (rc = ensure_seg(ctx))) return rc;
    float* r = ctx->f32a;
    float* tmp = ctx->f32b;
    float* d1 = ctx->f32c;
    float* mid = ctx->f32d;
    if ((rc = assemble<float>(ctx, ads::kFormHeat, 0, c_in, r, dt, kappa_d, nullptr))) return rc;
    if ((rc = seg_solve(ctx, r, tmp, d1))) return rc;
    if (kappa_d) {
        if ((rc = assemble<double>(ctx, ads::kFormHeat, 0, c_in, ctx->b64, dt, kappa_d, l2_d))) return rc;
        return mass_solve_seg(ctx, ctx->b64, c_out, d1, refine);
    }
    // the residuals read c_in: x accumulates in c_out unless the caller steps in place
    double* x = c_out != c_in ? c_out : ctx->x64;
    if ((rc = residual<float>(ctx, c_in, d1, nullptr, r, dt, l2_d))) return rc;
    if ((rc = seg_solve(ctx, r, tmp, mid, ads::SegOut{x, d1, ads::kSegAdd}))) return rc;
    for (int k = 1; k < refine; ++k) {
        if ((rc = residual<double>(ctx, c_in, x, nullptr, r, dt, nullptr))) return rc;
        if ((rc = seg_solve(ctx, r, tmp, mid, ads::SegOut{x, nullptr, ads::kSegAccumulate}))) return rc;
    }
    if (x != c_out)
        ADS_HIP(ctx, hipMemcpyAsync(c_out, x, sizeof(double) * ctx->D, hipMemcpyDeviceToDevice, ctx->stream));
    return ADS_OK;
}

// c_out = M^-1 A(kappa) c_in from FP64 residuals and FP32 sweeps, `refine`
// sweeps in all.  Constant kappa: the first pass computes the Kronecker
// right-hand side b = A c (the residual at x = 0; also l2 = c^T M c), kept
// in FP64, and the later ones the mass residual b - M x; the first
// correction d1 is FP32, so the second residual reads it as FP32 and
// x = d1 + d2 is formed in FP64 once.  Per-element kappa: b = A(kappa) c
// from the FP64 element assembly, then x = M^-1 b by mass_solve_ir.  c_in is
// not read once b is formed, so x accumulates in c_out even in place.
int step_ir_kron(ads_ctx* ctx, const double* c_in, double* c_out, const double* kappa_d, double dt, int refine,
                 double* l2_d)
{
    int rc = ensure_ir_scratch(ctx);
    if (rc) return rc;
    double* x = c_out;
    if (kappa_d) {
        if ((rc = assemble<double>(ctx, ads::kFormHeat, 0, c_in, ctx->b64, dt, kappa_d, l2_d))) return rc;
        if ((rc = mass_solve_ir(ctx, ctx->b64, x, nullptr, refine - 1, false))) return rc;
    } else {
        float* d1 = ctx->f32a;
        float* d2 = ctx->f32b;
        // the first residual (x = 0) is b = A c: kept in FP64 for the mass
        // residuals b - M x of the later passes
        double* b = refine > 1 ? ctx->b64 : nullptr;
        if ((rc = residual<float>(ctx, c_in, nullptr, nullptr, d1, dt, l2_d, b))) return rc;
        ADS_HIP(ctx, ads::launch_kron_solve<float>(ctx->p, d1, ctx->sw32, false, ctx->stream));
        if (refine == 1) {
            ADS_HIP(ctx, ads::launch_add_f2d(d1, nullptr, x, ctx->D, ctx->stream));
        } else {
            if ((rc = residual<float>(ctx, nullptr, d1, b, d2, 0.0, nullptr))) return rc;
            ADS_HIP(ctx, ads::launch_kron_solve<float>(ctx->p, d2, ctx->sw32, false, ctx->stream));
            ADS_HIP(ctx, ads::launch_add_f2d(d1, d2, x, ctx->D, ctx->stream));
            for (int k = 2; k < refine; ++k) {
                if ((rc = residual<double>(ctx, nullptr, x, b, d2, 0.0, nullptr))) return rc;
                ADS_HIP(ctx, ads::launch_kron_solve<float>(ctx->p, d2, ctx->sw32, false, ctx->stream));
                ADS_HIP(ctx, ads::launch_add_f2d(nullptr, d2, x, ctx->D, ctx->stream));
            }
        }
    }
    return ADS_OK;
}

// The segmented form: as step_ir_kron with every FP32 solve replaced by the
// segmented solve, whose last direction also forms x in FP64.  A per-element
// kappa: b = A(kappa) c from the FP64 element assembly, then x = M^-1 b by
// mass_solve_seg.
int step_ir_seg(ads_ctx* ctx, const double* c_in, double* c_out, const double* kappa_d, double dt, int refine,
                double* l2_d)
{
    int rc = ensure_ir_scratch(ctx);
    if (rc || (rc = ensure_seg(ctx))) return rc;
    if (kappa_d) {
        if ((rc = assemble<double>(ctx, ads::kFormHeat, 0, c_in, ctx->b64, dt, kappa_d, l2_d))) return rc;
        return mass_solve_seg(ctx, ctx->b64, c_out, nullptr, refine - 1);
    }
    double* x = c_out;  // c_in is not read once b is formed
    float* r = ctx->f32a;
    float* tmp = ctx->f32b;
    float* d1 = ctx->f32c;
    float* mid = ctx->f32d;  // the x-direction intermediate of the later solves
    // the first residual (x = 0) is b = A c: kept in FP64 for the mass
    // residuals b - M x of the later passes
    double* b = refine > 1 ? ctx->b64 : nullptr;
    if ((rc = residual<float>(ctx, c_in, nullptr, nullptr, r, dt, l2_d, b))) return rc;
    if (refine == 1) {
        rc = seg_solve(ctx, r, tmp, d1, ads::SegOut{x, nullptr, ads::kSegSet});
    } else {
        if ((rc = seg_solve(ctx, r, tmp, d1))) return rc;
        if ((rc = residual<float>(ctx, nullptr, d1, b, r, 0.0, nullptr))) return rc;
        if ((rc = seg_solve(ctx, r, tmp, mid, ads::SegOut{x, d1, ads::kSegAdd}))) return rc;
        for (int k = 2; k < refine && !rc; ++k) {
            if ((rc = residual<double>(ctx, nullptr, x, b, r, 0.0, nullptr))) return rc;
            rc = seg_solve(ctx, r, tmp, mid, ads::SegOut{x, nullptr, ads::kSegAccumulate});
        }
    }
    return rc;
}

// device copy of 3 n values (three arrays of n, or n triples when b and c
// are null), reallocated when larger than before
int oil_upload(ads_ctx* ctx, double** dst, long* cap, const double* a, const double* b, const double* c, long n)
{
    int rc;
    if (*cap < 3 * n || !*dst) {
        if ((rc = dev_alloc(ctx, dst, static_cast<size_t>(3 * n > 3 ? 3 * n : 3)))) return rc;
        *cap = 3 * n;
    }
    if (n == 0) return ADS_OK;
    if (b) {
        ADS_HIP(ctx, hipMemcpyAsync(*dst, a, sizeof(double) * n, hipMemcpyHostToDevice, ctx->stream));
        ADS_HIP(ctx, hipMemcpyAsync(*dst + n, b, sizeof(double) * n, hipMemcpyHostToDevice, ctx->stream));
        ADS_HIP(ctx, hipMemcpyAsync(*dst + 2 * n, c, sizeof(double) * n, hipMemcpyHostToDevice, ctx->stream));
    } else {
        ADS_HIP(ctx, hipMemcpyAsync(*dst, a, sizeof(double) * 3 * n, hipMemcpyHostToDevice, ctx->stream));
    }
    return ADS_OK;
}

int read_scalar(ads_ctx* ctx, const double* d, double* h, const char* what)
{
    if (!h) return ADS_OK;
    ADS_HIP(ctx, hipMemcpyAsync(h, d, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    ADS_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (!std::isfinite(*h)) return fail(ctx, ADS_ERR_NUMERIC, std::string(what) + " is not finite");
    return ADS_OK;
}

// The right-hand side of a device form (Form3DRHS) of the context's
// coefficients into F; l2 -> l2_d.  Heat: the point values of U_n (kept by
// ads_form_un(1); the current coefficients when none is kept or nothing has
// changed since), the gradients of the current coefficients.
int form_rhs(ads_ctx* ctx, int form, int f0, double dt, double* F, double* l2_d)
{
    const bool un_current = !ctx->un_kept || ctx->un_version == ctx->version;
    int nparts = 0, rc;
    switch (form) {
    case ADS_FORM_HEAT:
        if (!std::isfinite(dt)) return fail(ctx, ADS_ERR_INVALID, "dt is not finite");
        if (un_current) return assemble<double>(ctx, ads::kFormHeat, 0, ctx->c, F, dt, ctx->kappa, l2_d);
        if ((rc = zero(ctx, F, ctx->D))) return rc;
        ADS_HIP(ctx, ads::launch_heat_dual(ctx->p, ctx->un, ctx->c, F, ctx->tb64, dt, ctx->kappa, ctx->l2_part,
                                           ctx->l2_cap, &nparts, ctx->stream));
        ADS_HIP(ctx, ads::launch_reduce_partials(ctx->l2_part, nparts, l2_d, ctx->stream));
        return ADS_OK;
    case ADS_FORM_PROJECT:
        if (f0 < ADS_F0_CONST || f0 > ADS_F0_BUMP3D) return fail(ctx, ADS_ERR_UNSUPPORTED, "unknown f0 kind");
        return assemble<double>(ctx, ads::kFormProject, f0, nullptr, F, 0.0, nullptr, l2_d);
    case ADS_FORM_OIL:
    case ADS_FORM_OIL_PROJECT: {
        const bool tz = form == ADS_FORM_OIL_PROJECT;
        if (!ctx->oil.ready) return fail(ctx, ADS_ERR_INVALID, "oil form: call ads_oil_setup first");
        if (!tz && !std::isfinite(dt)) return fail(ctx, ADS_ERR_INVALID, "dt is not finite");
        if (!tz && !un_current)
            return fail(ctx, ADS_ERR_UNSUPPORTED,
                        "oil form: its device form reads values and gradients of one field, so U_n must be the "
                        "current coefficients");
        if ((rc = zero(ctx, F, ctx->D))) return rc;
        ADS_HIP(ctx, ads::launch_oil(ctx->p, tz, tz ? nullptr : ctx->c, F, ctx->tb64, ctx->oil.dev, ctx->oil.kq,
                                     tz ? 0.0 : dt, ctx->l2_part, tz ? nullptr : ctx->oil.dr_part, ctx->l2_cap,
                                     &nparts, ctx->stream));
        ADS_HIP(ctx, ads::launch_reduce_partials(ctx->l2_part, nparts, l2_d, ctx->stream));
        if (!tz) ADS_HIP(ctx, ads::launch_reduce_partials(ctx->oil.dr_part, nparts, ctx->oil.drained, ctx->stream));
        return ADS_OK;
    }
    default:
        return fail(ctx, ADS_ERR_INVALID, "unknown right-hand-side form");
    }
}

}  // namespace

// ---- the allocation-free grid evaluation (ads_eval_grid_noalloc_d) -------
//
// Its tables live in `scratch`, a static device array: per direction the
// B-spline values [i*(p+1) + a] of up to kGridNoAllocRes grid points at
// degree <= kGridNoAllocDeg, then per direction the first basis index of
// each point.  Each GPU the library runs on holds one copy (540,672 bytes)
// until the process ends, and every call on that GPU shares it: a call
// overwrites the first res[d] entries of each table before reading them, so
// nothing needs clearing, and calls that overlap on one GPU race.  Host code
// reaches the array only through the address hipGetSymbolAddress returns for
// the current device (scratch_address), never through its name.
constexpr int kGridNoAllocRes = 4096;
constexpr int kGridNoAllocDeg = 4;
constexpr size_t kGridValBytes = sizeof(double) * kGridNoAllocRes * (kGridNoAllocDeg + 1);
constexpr size_t kGridIdxBytes = sizeof(int) * kGridNoAllocRes;
constexpr size_t kGridScratchBytes = 3 * (kGridValBytes + kGridIdxBytes);
static_assert(kGridValBytes % 256 == 0, "tables start 256-byte aligned");

static __device__ __align__(256) unsigned char scratch[kGridScratchBytes];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_addr[kMaxDevices] = {};

// The device address of `scratch` on the current device, looked up once per
// device.
int scratch_address(ads_ctx* ctx, void** out)
{
    int dev = -1;
    ADS_HIP(ctx, hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return fail(ctx, ADS_ERR_UNSUPPORTED, "device ordinal beyond 63");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_addr[dev]) {
        void* addr = nullptr;
        ADS_HIP(ctx, hipGetSymbolAddress(&addr, HIP_SYMBOL(scratch)));
        if (!addr) return fail(ctx, ADS_ERR_HIP, "hipGetSymbolAddress returned a null address for the grid scratch");
        g_scratch_addr[dev] = addr;
    }
    *out = g_scratch_addr[dev];
    return ADS_OK;
}

struct GridDir {
    const double* U;  // knots on the device
    int n, p, res;    // p: the direction's own degree
    double lo, hi;
    double* vals;  // res * (kernel degree + 1)
    int* first;    // res
};
struct GridArgs {
    GridDir d[3];
    int p;  // the kernels' degree (the largest)
};

// One thread per (grid point, direction): the point's first basis index
// and B-spline values, rounded exactly as the host tables (grid_point), zero
// past the direction's own degree.
__global__ __launch_bounds__(256) void k_grid_tables(GridArgs g)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const GridDir& a = g.d[blockIdx.y];
    if (i >= a.res) return;
    double* v = a.vals + static_cast<long>(i) * (g.p + 1);
    ads::grid_point(a.n, a.p, a.U, a.res, a.lo, a.hi, i, a.first + i, v);
    for (int k = a.p + 1; k <= g.p; ++k) v[k] = 0.0;
}

}  // namespace

extern "C" {

int ads_init(const int n[3], const int p[3], ads_ctx** out)
{
    if (!n || !p || !out) return fail(nullptr, ADS_ERR_INVALID, "null argument");
    *out = nullptr;
    for (int d = 0; d < 3; ++d)
        if (p[d] < 1 || p[d] > 5)
            return fail(nullptr, ADS_ERR_UNSUPPORTED, "degree must be 1..5 in each direction for the GPU kernels");
    ads_ctx* ctx = new (std::nothrow) ads_ctx();
    if (!ctx) return fail(nullptr, ADS_ERR_NOMEM, "host allocation failed");
    ctx->p = std::max(p[0], std::max(p[1], p[2]));
    for (int d = 0; d < 3; ++d) {
        std::string e = ads::build_direction(n[d], p[d], ctx->dir[d]);
        // kernels compiled for one degree: a lower-degree direction runs on
        // tables widened with exact zeros
        if (e.empty()) e = ads::pad_direction(ctx->dir[d], ctx->p, ctx->kdir[d]);
        if (!e.empty()) {
            int code = e.find("singular") != std::string::npos ? ADS_ERR_SINGULAR : ADS_ERR_INVALID;
            delete ctx;
            return fail(nullptr, code, "direction " + std::to_string(d) + ": " + e);
        }
    }
    ctx->D = static_cast<long>(ctx->dir[0].dofs) * ctx->dir[1].dofs * ctx->dir[2].dofs;
    ctx->E = static_cast<long>(ctx->dir[0].nelem) * ctx->dir[1].nelem * ctx->dir[2].nelem;
    if (ctx->D > (1L << 31) - 1) {
        delete ctx;
        return fail(nullptr, ADS_ERR_INVALID, "problem too large (more than 2^31 DOFs)");
    }
    int rc = ADS_OK;
    hipError_t e = hipStreamCreate(&ctx->own);
    if (e != hipSuccess) {
        delete ctx;
        return fail(nullptr, ADS_ERR_HIP, std::string("hipStreamCreate: ") + hipGetErrorString(e));
    }
    ctx->stream = ctx->own;
    for (int d = 0; d < 3 && !rc; ++d) {
        rc = upload_dir(ctx, ctx->kdir[d], ctx->d64[d], ctx->tb64.d[d], ctx->sw64[d]);
        if (!rc) rc = upload_dir(ctx, ctx->kdir[d], ctx->d32[d], ctx->tb32.d[d], ctx->sw32[d]);
        if (!rc) rc = upload(ctx, &ctx->mass[d], ctx->kdir[d].mass_rows);
        if (!rc) rc = upload(ctx, &ctx->stiff[d], ctx->kdir[d].stiff_rows);
    }
    for (int d = 0; d < 3 && !rc; ++d) rc = upload(ctx, &ctx->knots[d], ctx->dir[d].knots);
    if (!rc) rc = build_part(ctx);
    if (!rc) rc = scratch_address(ctx, &ctx->grid_scratch);
    if (!rc) rc = dev_alloc(ctx, &ctx->c, ctx->D);
    if (!rc) rc = dev_alloc(ctx, &ctx->F, ctx->D);
    {
        // partial sums: at most one per (64 x elements, y row, z layer) in the
        // assembly, and one per (64 x DOFs, 4 y DOFs, z DOF) in the residual
        const long asm_cap = static_cast<long>((ctx->dir[0].nelem + 63) / 64) * ctx->dir[1].nelem * ctx->dir[2].nelem;
        ctx->l2_cap = static_cast<int>(std::max<long>(asm_cap, ads::kron_residual_partials_bound(band_set(ctx))));
    }
    if (!rc) rc = dev_alloc(ctx, &ctx->l2_part, ctx->l2_cap);
    if (!rc) rc = dev_alloc(ctx, &ctx->l2, 1);
    if (!rc) rc = zero(ctx, ctx->c, ctx->D);
    if (!rc) rc = zero(ctx, ctx->l2, 1);
    if (!rc && hipStreamSynchronize(ctx->stream) != hipSuccess) rc = fail(ctx, ADS_ERR_HIP, "init sync failed");
    if (rc) {
        g_init_error = ctx->err.empty() ? g_init_error : ctx->err;
        ads_destroy(ctx);
        return rc;
    }
    *out = ctx;
    return ADS_OK;
}

int ads_destroy(ads_ctx* ctx)
{
    if (!ctx) return ADS_OK;
    if (ctx->own) (void)hipStreamSynchronize(ctx->own);
    if (ctx->stream && ctx->stream != ctx->own) (void)hipStreamSynchronize(ctx->stream);
    for (void* p : ctx->allocs) (void)hipFree(p);
    if (ctx->own) (void)hipStreamDestroy(ctx->own);
    delete ctx;
    return ADS_OK;
}

const char* ads_last_error(const ads_ctx* ctx) { return ctx ? ctx->err.c_str() : g_init_error.c_str(); }

int ads_set_stream(ads_ctx* ctx, void* hip_stream)
{
    if (check_ctx(ctx)) return ADS_ERR_INVALID;
    ctx->stream = hip_stream ? static_cast<hipStream_t>(hip_stream) : ctx->own;
    return ADS_OK;
}

int ads_synchronize(ads_ctx* ctx)
{
    if (check_ctx(ctx)) return ADS_ERR_INVALID;
    ADS_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return ADS_OK;
}

int ads_get_sizes(const ads_ctx* ctx, int dofs[3], int nelem[3], int p[3])
{
    if (check_ctx(ctx)) return ADS_ERR_INVALID;
    for (int d = 0; d < 3; ++d) {
        if (dofs) dofs[d] = ctx->dir[d].dofs;
        if (nelem) nelem[d] = ctx->dir[d].nelem;
        if (p) p[d] = ctx->dir[d].p;
    }
    return ADS_OK;
}

int ads_get_knots(const ads_ctx* ctx, int dim, double* U)
{
    if (check_ctx(ctx) || dim < 0 || dim > 2 || !U) return ADS_ERR_INVALID;
    std::memcpy(U, ctx->dir[dim].knots.data(), sizeof(double) * ctx->dir[dim].knots.size());
    return ADS_OK;
}

int ads_get_basis(const ads_ctx* ctx, int dim, int* O, double* J, double* W, double* X, double* NN)
{
    if (check_ctx(ctx) || dim < 0 || dim > 2) return ADS_ERR_INVALID;
    const ads::Direction& d = ctx->dir[dim];
    if (O) std::memcpy(O, d.first_dof.data(), sizeof(int) * d.first_dof.size());
    if (J) std::memcpy(J, d.jac.data(), sizeof(double) * d.jac.size());
    if (W) std::memcpy(W, d.wts.data(), sizeof(double) * d.wts.size());
    if (X) std::memcpy(X, d.pts.data(), sizeof(double) * d.pts.size());
    if (NN) std::memcpy(NN, d.tab.data(), sizeof(double) * d.tab.size());
    return ADS_OK;
}

int ads_get_matrix(const ads_ctx* ctx, int dim, const double mix[4], double* dense)
{
    if (check_ctx(ctx) || dim < 0 || dim > 2 || !mix || !dense) return ADS_ERR_INVALID;
    ads::dense_matrix(ctx->dir[dim], mix, dense);
    return ADS_OK;
}

int ads_set_coeffs(ads_ctx* ctx, const double* host)
{
    if (check_ctx(ctx) || !host) return ADS_ERR_INVALID;
    ++ctx->version;
    ADS_HIP(ctx, hipMemcpyAsync(ctx->c, host, sizeof(double) * ctx->D, hipMemcpyHostToDevice, ctx->stream));
    ADS_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return ADS_OK;
}

int ads_get_coeffs(ads_ctx* ctx, double* host)
{
    if (check_ctx(ctx) || !host) return ADS_ERR_INVALID;
    ADS_HIP(ctx, hipMemcpyAsync(host, ctx->c, sizeof(double) * ctx->D, hipMemcpyDeviceToHost, ctx->stream));
    ADS_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return ADS_OK;
}

int ads_coeffs_device(ads_ctx* ctx, double** dptr)
{
    if (check_ctx(ctx) || !dptr) return ADS_ERR_INVALID;
    *dptr = ctx->c;
    return ADS_OK;
}

int ads_set_kappa(ads_ctx* ctx, const double* host)
{
    if (check_ctx(ctx)) return ADS_ERR_INVALID;
    if (!host) {
        ctx->kappa = nullptr;
        return ADS_OK;
    }
    for (long i = 0; i < ctx->E; ++i)
        if (!std::isfinite(host[i])) return fail(ctx, ADS_ERR_NUMERIC, "kappa has a non-finite entry");
    static_assert(sizeof(double) == 8, "");
    if (!ctx->kappa) {
        // keep the buffer across set/unset cycles
        double* k = nullptr;
        int rc = dev_alloc(ctx, &k, ctx->E);
        if (rc) return rc;
        ctx->kappa = k;
    }
    ADS_HIP(ctx, hipMemcpyAsync(ctx->kappa, host, sizeof(double) * ctx->E, hipMemcpyHostToDevice, ctx->stream));
    ADS_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return ADS_OK;
}

int ads_project(ads_ctx* ctx, int f0_kind, double* l2norm)
{
    if (check_ctx(ctx)) return ADS_ERR_INVALID;
    if (f0_kind < ADS_F0_CONST || f0_kind > ADS_F0_BUMP3D) return fail(ctx, ADS_ERR_UNSUPPORTED, "unknown f0 kind");
    int rc = assemble<double>(ctx, ads::kFormProject, f0_kind, nullptr, ctx->F, 0.0, nullptr, nullptr);
    if (rc) return rc;
    if ((rc = mass_solve64(ctx, ctx->F, false))) return rc;
    std::swap(ctx->c, ctx->F);
    ++ctx->version;
    return read_l2(ctx, l2norm);
}

int ads_step(ads_ctx* ctx, double dt, int prec, int refine, double* l2norm)
{
    if (check_ctx(ctx)) return ADS_ERR_INVALID;
    int rc = ads_step_d(ctx, ctx->c, ctx->F, ctx->kappa, dt, prec, refine, ctx->l2);
    if (rc) return rc;
    std::swap(ctx->c, ctx->F);
    ++ctx->version;
    return read_l2(ctx, l2norm);
}

int ads_form_rhs(ads_ctx* ctx, double dt, double* F_host, double* l2norm)
{
    if (check_ctx(ctx) || !F_host) return ADS_ERR_INVALID;
    int rc = assemble<double>(ctx, ads::kFormHeat, 0, ctx->c, ctx->F, dt, ctx->kappa, nullptr);
    if (rc) return rc;
    ADS_HIP(ctx, hipMemcpyAsync(F_host, ctx->F, sizeof(double) * ctx->D, hipMemcpyDeviceToHost, ctx->stream));
    ADS_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return read_l2(ctx, l2norm);
}

int ads_last_l2norm(ads_ctx* ctx, double* l2norm)
{
    if (check_ctx(ctx) || !l2norm) return ADS_ERR_INVALID;
    return read_l2(ctx, l2norm);
}

int ads_form_rhs_d(ads_ctx* ctx, const double* c_d, const double* kappa_d, double dt, double* F_d, double* l2_d)
{
    if (check_ctx(ctx) || !c_d || !F_d) return ADS_ERR_INVALID;
    if (c_d == F_d) return fail(ctx, ADS_ERR_INVALID, "ads_form_rhs_d: c and F must not alias");
    return assemble<double>(ctx, ads::kFormHeat, 0, c_d, F_d, dt, kappa_d, l2_d);
}

int ads_solve_d(ads_ctx* ctx, double* F_d, int trans)
{
    if (check_ctx(ctx) || !F_d) return ADS_ERR_INVALID;
    return mass_solve64(ctx, F_d, trans != 0);
}

int ads_step_d(ads_ctx* ctx, const double* c_in_d, double* c_out_d, const double* kappa_d, double dt, int prec,
               int refine, double* l2_d)
{
    if (check_ctx(ctx) || !c_in_d || !c_out_d) return ADS_ERR_INVALID;
    if (!std::isfinite(dt)) return fail(ctx, ADS_ERR_INVALID, "dt is not finite");
    if (prec == ADS_PREC_FP64 || prec == ADS_PREC_FP64_KRON) {
        double* target = (c_in_d == c_out_d) ? (c_out_d == ctx->F ? ctx->c : ctx->F) : c_out_d;
        if (target == c_in_d) return fail(ctx, ADS_ERR_INVALID, "internal: buffer aliasing");
        // the Kronecker right-hand side needs a constant diffusivity
        int rc = (prec == ADS_PREC_FP64_KRON && !kappa_d)
                     ? residual<double, double>(ctx, c_in_d, nullptr, nullptr, target, dt, l2_d ? l2_d : ctx->l2)
                     : assemble<double>(ctx, ads::kFormHeat, 0, c_in_d, target, dt, kappa_d, l2_d);
        if (rc) return rc;
        if ((rc = mass_solve64(ctx, target, false))) return rc;
        if (target != c_out_d)
            ADS_HIP(ctx, hipMemcpyAsync(c_out_d, target, sizeof(double) * ctx->D, hipMemcpyDeviceToDevice, ctx->stream));
        return ADS_OK;
    }
    if (prec == ADS_PREC_FP32_IR || prec == ADS_PREC_FP32_IR_KRON || prec == ADS_PREC_FP32_IR_SEG) {
        if (refine < 1 || refine > 4) return fail(ctx, ADS_ERR_INVALID, "refine must be 1..4");
        if (c_in_d == c_out_d && c_out_d == ctx->x64)
            return fail(ctx, ADS_ERR_INVALID, "internal: buffer aliasing");
        if (prec == ADS_PREC_FP32_IR) return step_ir(ctx, c_in_d, c_out_d, kappa_d, dt, refine, l2_d);
        if (prec == ADS_PREC_FP32_IR_KRON) return step_ir_kron(ctx, c_in_d, c_out_d, kappa_d, dt, refine, l2_d);
        return step_ir_seg(ctx, c_in_d, c_out_d, kappa_d, dt, refine, l2_d);
    }
    return fail(ctx, ADS_ERR_INVALID, "unknown precision form");
}

int ads_step_adjoint_d(ads_ctx* ctx, const double* c_n_d, const double* lam_next_d, const double* kappa_d, double dt,
                       int prec, int refine, double* lam_out_d, double* dkappa_d)
{
    if (check_ctx(ctx) || !c_n_d || !lam_next_d || !lam_out_d) return ADS_ERR_INVALID;
    if (lam_out_d == c_n_d) return fail(ctx, ADS_ERR_INVALID, "ads_step_adjoint_d: lam_out must not alias c_n");
    int rc;
    if (!ctx->mu && (rc = dev_alloc(ctx, &ctx->mu, ctx->D))) return rc;
    int nparts = 0;
    if (prec == ADS_PREC_FP64 || prec == ADS_PREC_FP64_KRON) {
        // mu = M^-T lam_next
        ADS_HIP(ctx, hipMemcpyAsync(ctx->mu, lam_next_d, sizeof(double) * ctx->D, hipMemcpyDeviceToDevice,
                                    ctx->stream));
        if ((rc = mass_solve64(ctx, ctx->mu, true))) return rc;
        // lam_out = A(kappa) mu (A symmetric) and dkappa += -dt c_n,e^T K_e mu_e in one pass
        if ((rc = zero(ctx, lam_out_d, ctx->D))) return rc;
        ADS_HIP(ctx, ads::launch_heat_adjoint<double>(ctx->p, ctx->mu, c_n_d, lam_out_d, ctx->tb64, dt, kappa_d,
                                                      dkappa_d, ctx->l2_part, ctx->l2_cap, &nparts, ctx->stream));
        return ADS_OK;
    }
    if (prec != ADS_PREC_FP32_IR && prec != ADS_PREC_FP32_IR_KRON && prec != ADS_PREC_FP32_IR_SEG)
        return fail(ctx, ADS_ERR_INVALID, "unknown precision form");
    if (refine < 1 || refine > 4) return fail(ctx, ADS_ERR_INVALID, "refine must be 1..4");
    if ((rc = ensure_ir_scratch(ctx))) return rc;
    // mu = M^-T lam_next: FP32 transposed sweeps (segmented solves: M is
    // symmetric) corrected with FP64 mass residuals
    if (prec == ADS_PREC_FP32_IR_SEG) rc = mass_solve_seg(ctx, lam_next_d, ctx->mu, nullptr, refine - 1);
    else rc = mass_solve_ir(ctx, lam_next_d, ctx->mu, nullptr, refine - 1, true);
    if (rc) return rc;
    // lam_out = A(kappa) mu and dkappa in FP32 arithmetic (dkappa accumulates in FP64)
    if ((rc = zero(ctx, ctx->f32a, ctx->D))) return rc;
    ADS_HIP(ctx, ads::launch_heat_adjoint<float>(ctx->p, ctx->mu, c_n_d, ctx->f32a, ctx->tb32, dt, kappa_d, dkappa_d,
                                                 ctx->l2_part, ctx->l2_cap, &nparts, ctx->stream));
    ADS_HIP(ctx, ads::launch_add_f2d(ctx->f32a, nullptr, lam_out_d, ctx->D, ctx->stream));
    return ADS_OK;
}

int ads_eval_grid(ads_ctx* ctx, const int res[3], const double lo[3], const double hi[3], double* vals)
{
    if (check_ctx(ctx) || !res || !lo || !hi || !vals) return ADS_ERR_INVALID;
    long npts = 1;
    for (int d = 0; d < 3; ++d) {
        if (res[d] < 2 || res[d] > 65536) return fail(ctx, ADS_ERR_INVALID, "ads_eval_grid: res must be 2..65536");
        if (!std::isfinite(lo[d]) || !std::isfinite(hi[d])) return fail(ctx, ADS_ERR_INVALID, "ads_eval_grid: bad range");
        npts *= res[d];
    }
    if (npts > (1L << 30)) return fail(ctx, ADS_ERR_INVALID, "ads_eval_grid: more than 2^30 points");
    std::vector<int> first[3];
    std::vector<double> tab[3];
    for (int d = 0; d < 3; ++d) grid_tables_k(ctx, d, res[d], lo[d], hi[d], first[d], tab[d]);
    // the tables and the values live only for this call
    void* buf[7] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    auto release = [&]() {
        for (void* b : buf)
            if (b) (void)hipFree(b);
    };
    hipError_t e = hipSuccess;
    for (int d = 0; d < 3 && e == hipSuccess; ++d) {
        e = hipMalloc(&buf[2 * d], sizeof(int) * first[d].size());
        if (e == hipSuccess) e = hipMalloc(&buf[2 * d + 1], sizeof(double) * tab[d].size());
        if (e == hipSuccess)
            e = hipMemcpyAsync(buf[2 * d], first[d].data(), sizeof(int) * first[d].size(), hipMemcpyHostToDevice,
                               ctx->stream);
        if (e == hipSuccess)
            e = hipMemcpyAsync(buf[2 * d + 1], tab[d].data(), sizeof(double) * tab[d].size(), hipMemcpyHostToDevice,
                               ctx->stream);
    }
    if (e == hipSuccess) e = hipMalloc(&buf[6], sizeof(double) * npts);
    if (e == hipSuccess)
        e = ads::launch_eval_grid(ctx->p, ctx->c, ctx->dir[0].dofs, ctx->dir[1].dofs, ctx->dir[2].dofs,
                                  static_cast<int*>(buf[0]),
                                  static_cast<double*>(buf[1]), res[0], static_cast<int*>(buf[2]),
                                  static_cast<double*>(buf[3]), res[1], static_cast<int*>(buf[4]),
                                  static_cast<double*>(buf[5]), res[2], static_cast<double*>(buf[6]), ctx->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(vals, buf[6], sizeof(double) * npts, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    release();
    if (e != hipSuccess) return hip_fail(ctx, e, "ads_eval_grid");
    return ADS_OK;
}

int ads_eval_grid_noalloc_d(ads_ctx* ctx, const int res[3], const double lo[3], const double hi[3], double* vals_d)
{
    if (check_ctx(ctx) || !res || !lo || !hi || !vals_d) return ADS_ERR_INVALID;
    if (ctx->p > kGridNoAllocDeg)
        return fail(ctx, ADS_ERR_UNSUPPORTED, "ads_eval_grid_noalloc_d: degree above 4");
    int rmax = 0;
    for (int d = 0; d < 3; ++d) {
        if (res[d] < 2 || res[d] > kGridNoAllocRes)
            return fail(ctx, ADS_ERR_INVALID,
                        "ads_eval_grid_noalloc_d: res must be 2..4096 per direction (the size of its static tables)");
        if (!std::isfinite(lo[d]) || !std::isfinite(hi[d]))
            return fail(ctx, ADS_ERR_INVALID, "ads_eval_grid_noalloc_d: bad range");
        rmax = std::max(rmax, res[d]);
    }
    if (!ctx->grid_scratch) return fail(ctx, ADS_ERR_HIP, "ads_eval_grid_noalloc_d: no scratch address");
    unsigned char* base = static_cast<unsigned char*>(ctx->grid_scratch);
    GridArgs g{};
    g.p = ctx->p;
    for (int d = 0; d < 3; ++d) {
        g.d[d].U = ctx->knots[d];
        g.d[d].n = ctx->dir[d].n;
        g.d[d].p = ctx->dir[d].p;
        g.d[d].res = res[d];
        g.d[d].lo = lo[d];
        g.d[d].hi = hi[d];
        g.d[d].vals = reinterpret_cast<double*>(base + d * kGridValBytes);
        g.d[d].first = reinterpret_cast<int*>(base + 3 * kGridValBytes + d * kGridIdxBytes);
    }
    hipLaunchKernelGGL(k_grid_tables, dim3(static_cast<unsigned>((rmax + 255) / 256), 3), dim3(256), 0, ctx->stream, g);
    ADS_HIP(ctx, hipGetLastError());
    ADS_HIP(ctx, ads::launch_eval_grid(ctx->p, ctx->c, ctx->dir[0].dofs, ctx->dir[1].dofs, ctx->dir[2].dofs,
                                       g.d[0].first, g.d[0].vals,
                                       res[0], g.d[1].first, g.d[1].vals, res[1], g.d[2].first, g.d[2].vals, res[2],
                                       vals_d, ctx->stream));
    return ADS_OK;
}

int ads_host_grid_tables(int n, int p, int res, double lo, double hi, int* first, double* vals)
{
    if (res < 2 || !first || !vals) return fail(nullptr, ADS_ERR_INVALID, "ads_host_grid_tables: bad arguments");
    ads::Direction d;
    std::string e = ads::build_direction(n, p, d);
    if (!e.empty()) return fail(nullptr, e.find("singular") != std::string::npos ? ADS_ERR_SINGULAR : ADS_ERR_INVALID, e);
    std::vector<int> f;
    std::vector<double> v;
    ads::grid_tables(d, res, lo, hi, f, v);
    std::memcpy(first, f.data(), sizeof(int) * f.size());
    std::memcpy(vals, v.data(), sizeof(double) * v.size());
    return ADS_OK;
}

int ads_host_basis(int n, int p, double* U, int* O, double* J, double* W, double* X, double* NN)
{
    ads::Direction d;
    std::string e = ads::build_direction(n, p, d);
    if (!e.empty()) return fail(nullptr, e.find("singular") != std::string::npos ? ADS_ERR_SINGULAR : ADS_ERR_INVALID, e);
    if (U) std::memcpy(U, d.knots.data(), sizeof(double) * d.knots.size());
    if (O) std::memcpy(O, d.first_dof.data(), sizeof(int) * d.first_dof.size());
    if (J) std::memcpy(J, d.jac.data(), sizeof(double) * d.jac.size());
    if (W) std::memcpy(W, d.wts.data(), sizeof(double) * d.wts.size());
    if (X) std::memcpy(X, d.pts.data(), sizeof(double) * d.pts.size());
    if (NN) std::memcpy(NN, d.tab.data(), sizeof(double) * d.tab.size());
    return ADS_OK;
}

int ads_fill_uniform(uint64_t seed, int64_t n, double lo, double hi, double* out)
{
    if (n < 0 || (n > 0 && !out)) return ADS_ERR_INVALID;
    ads::fill_uniform(seed, n, lo, hi, out);
    return ADS_OK;
}

int ads_fill_lognormal(uint64_t seed, int64_t n, double mu, double sigma, double* out)
{
    if (n < 0 || (n > 0 && !out)) return ADS_ERR_INVALID;
    ads::fill_lognormal(seed, n, mu, sigma, out);
    return ADS_OK;
}

namespace {

// The factors of mix(1)M + mix(2)K + mix(3)B + mix(4)B^T per direction, built
// on first use (MKBBT + band LU, as ComputeMatrix + SolveOneDirection do in
// every substep) and kept for the context's lifetime.
int mix_set(ads_ctx* ctx, const double mix[4], const ads::SweepTables<double>** out)
{
    for (const MixSet& m : ctx->mixsets)
        if (m.mix[0] == mix[0] && m.mix[1] == mix[1] && m.mix[2] == mix[2] && m.mix[3] == mix[3]) {
            *out = m.sw;
            return ADS_OK;
        }
    if (ctx->mixsets.size() >= 64) return fail(ctx, ADS_ERR_INVALID, "MultiStep: more than 64 distinct mix columns");
    MixSet m{};
    for (int k = 0; k < 4; ++k) m.mix[k] = mix[k];
    for (int d = 0; d < 3; ++d) {
        const ads::Direction& dir = ctx->kdir[d];
        std::vector<double> band, lu, fwd, bwd;
        std::vector<int> piv, fpiv;
        ads::mix_band(dir, mix, band);
        std::string e = ads::band_lu(dir.dofs, dir.p, band, lu, piv);
        if (!e.empty()) return fail(ctx, ADS_ERR_SINGULAR, "MultiStep: direction " + std::to_string(d) + ": " + e);
        ads::sweep_tables(dir.dofs, dir.p, lu, piv, fwd, fpiv, bwd);
        double *dfwd = nullptr, *dbwd = nullptr;
        int* dpiv = nullptr;
        int rc;
        if ((rc = upload(ctx, &dfwd, fwd))) return rc;
        if ((rc = upload(ctx, &dpiv, fpiv))) return rc;
        if ((rc = upload(ctx, &dbwd, bwd))) return rc;
        m.sw[d].fwd = dfwd;
        m.sw[d].piv = dpiv;
        m.sw[d].bwd = dbwd;
        m.sw[d].n = dir.dofs;
        m.sw[d].pivots = 0;
        for (int v : fpiv) m.sw[d].pivots |= (v != 0);
    }
    ctx->mixsets.push_back(m);
    *out = ctx->mixsets.back().sw;
    return ADS_OK;
}

}  // namespace

int ads_multistep(ads_ctx* ctx, const double mix[12], double dt, double* l2norm)
{
    if (check_ctx(ctx) || !mix) return ADS_ERR_INVALID;
    if (!std::isfinite(dt)) return fail(ctx, ADS_ERR_INVALID, "dt is not finite");
    int rc;
    if (!ctx->x64 && (rc = dev_alloc(ctx, &ctx->x64, ctx->D))) return rc;
    // the sets are referenced by pointer below: no reallocation within the cap
    ctx->mixsets.reserve(64);
    const ads::SweepTables<double>* sw[3];
    for (int sub = 0; sub < 3; ++sub)
        if ((rc = mix_set(ctx, mix + 4 * sub, &sw[sub]))) return rc;
    // substep outputs: F, x64, F (the last one is the new state)
    double* out[3] = {ctx->F, ctx->x64, ctx->F};
    const double* prev = ctx->c;
    for (int sub = 0; sub < 3; ++sub) {
        if (sub == 0) {
            // values and gradients from the same (initial) coefficients
            if ((rc = assemble<double>(ctx, ads::kFormHeat, 0, ctx->c, out[0], dt, ctx->kappa, ctx->l2))) return rc;
        } else {
            if ((rc = zero(ctx, out[sub], ctx->D))) return rc;
            int nparts = 0;
            ADS_HIP(ctx, ads::launch_heat_dual(ctx->p, ctx->c, prev, out[sub], ctx->tb64, dt, ctx->kappa,
                                               ctx->l2_part, ctx->l2_cap, &nparts, ctx->stream));
        }
        ADS_HIP(ctx, ads::launch_kron_solve<double>(ctx->p, out[sub], sw[sub], false, ctx->stream));
        prev = out[sub];
    }
    std::swap(ctx->c, ctx->F);
    ++ctx->version;
    return read_l2(ctx, l2norm);
}

int ads_oil_setup(ads_ctx* ctx, const double* cx, const double* cy, const double* cz, int cN, int cL,
                  const double* pumps, int npumps, const double* drains, int ndrains, const double params[6])
{
    if (check_ctx(ctx) || !cx || !cy || !cz || !params) return ADS_ERR_INVALID;
    if (cN < 1 || cL < 2 || npumps < 0 || ndrains < 0 || (npumps > 0 && !pumps) || (ndrains > 0 && !drains))
        return fail(ctx, ADS_ERR_INVALID, "ads_oil_setup: need cN >= 1 curves of cL >= 2 points and valid sources");
    for (int i = 0; i < 6; ++i)
        if (!std::isfinite(params[i])) return fail(ctx, ADS_ERR_INVALID, "ads_oil_setup: parameters must be finite");
    if (params[3] <= 0.0) return fail(ctx, ADS_ERR_INVALID, "ads_oil_setup: radius must be positive");
    if (ctx->dir[0].p != ctx->dir[1].p || ctx->dir[1].p != ctx->dir[2].p)
        return fail(ctx, ADS_ERR_UNSUPPORTED, "ads_oil_setup: the oil form needs the same degree in x, y and z");
    if (ctx->p > 4) return fail(ctx, ADS_ERR_UNSUPPORTED, "ads_oil_setup: the oil kernels are built for degrees 1..4");
    auto& o = ctx->oil;
    o.ready = false;
    const long npt = static_cast<long>(cN) * cL;
    int rc;
    if ((rc = oil_upload(ctx, &o.curves, &o.curves_cap, cx, cy, cz, npt))) return rc;
    if ((rc = oil_upload(ctx, &o.pumps, &o.pumps_cap, pumps, nullptr, nullptr, npumps))) return rc;
    if ((rc = oil_upload(ctx, &o.drains, &o.drains_cap, drains, nullptr, nullptr, ndrains))) return rc;
    const long q3 = static_cast<long>(ctx->p + 1) * (ctx->p + 1) * (ctx->p + 1);
    if (!o.kq && (rc = dev_alloc(ctx, &o.kq, static_cast<size_t>(ctx->E * q3)))) return rc;
    if (!o.dr_part && (rc = dev_alloc(ctx, &o.dr_part, static_cast<size_t>(ctx->l2_cap)))) return rc;
    if (!o.drained && (rc = dev_alloc(ctx, &o.drained, 1))) return rc;
    o.dev = ads::OilDev{o.curves, cN, cL, o.pumps, npumps, o.drains, ndrains,
                        params[0], params[1], params[2], params[3], params[4], params[5]};
    ADS_HIP(ctx, ads::launch_oil_kq(ctx->p, ctx->tb64, o.dev, o.kq, ctx->stream));
    // the host arrays may be released once this returns
    ADS_HIP(ctx, hipStreamSynchronize(ctx->stream));
    o.ready = true;
    return ADS_OK;
}

int ads_oil_project(ads_ctx* ctx, double* l2norm)
{
    if (check_ctx(ctx)) return ADS_ERR_INVALID;
    if (!ctx->oil.ready) return fail(ctx, ADS_ERR_INVALID, "ads_oil_project: call ads_oil_setup first");
    int rc = zero(ctx, ctx->F, ctx->D);
    if (rc) return rc;
    int nparts = 0;
    ADS_HIP(ctx, ads::launch_oil(ctx->p, true, nullptr, ctx->F, ctx->tb64, ctx->oil.dev, ctx->oil.kq, 0.0,
                                 ctx->l2_part, nullptr, ctx->l2_cap, &nparts, ctx->stream));
    ADS_HIP(ctx, ads::launch_reduce_partials(ctx->l2_part, nparts, ctx->l2, ctx->stream));
    if ((rc = mass_solve64(ctx, ctx->F, false))) return rc;
    std::swap(ctx->c, ctx->F);
    ++ctx->version;
    return read_l2(ctx, l2norm);
}

int ads_oil_step(ads_ctx* ctx, double dt, double* l2norm, double* drained)
{
    if (check_ctx(ctx)) return ADS_ERR_INVALID;
    if (!ctx->oil.ready) return fail(ctx, ADS_ERR_INVALID, "ads_oil_step: call ads_oil_setup first");
    if (!std::isfinite(dt)) return fail(ctx, ADS_ERR_INVALID, "dt is not finite");
    int rc = zero(ctx, ctx->F, ctx->D);
    if (rc) return rc;
    int nparts = 0;
    ADS_HIP(ctx, ads::launch_oil(ctx->p, false, ctx->c, ctx->F, ctx->tb64, ctx->oil.dev, ctx->oil.kq, dt,
                                 ctx->l2_part, ctx->oil.dr_part, ctx->l2_cap, &nparts, ctx->stream));
    ADS_HIP(ctx, ads::launch_reduce_partials(ctx->l2_part, nparts, ctx->l2, ctx->stream));
    ADS_HIP(ctx, ads::launch_reduce_partials(ctx->oil.dr_part, nparts, ctx->oil.drained, ctx->stream));
    if ((rc = mass_solve64(ctx, ctx->F, false))) return rc;
    std::swap(ctx->c, ctx->F);
    ++ctx->version;
    if ((rc = read_l2(ctx, l2norm))) return rc;
    return read_scalar(ctx, ctx->oil.drained, drained, "drained");
}

int ads_oil_get_kq(ads_ctx* ctx, double* kq_host)
{
    if (check_ctx(ctx) || !kq_host) return ADS_ERR_INVALID;
    if (!ctx->oil.ready) return fail(ctx, ADS_ERR_INVALID, "ads_oil_get_kq: call ads_oil_setup first");
    const long q3 = static_cast<long>(ctx->p + 1) * (ctx->p + 1) * (ctx->p + 1);
    // the device cache is point-major (kq[q E + e]); the caller gets the
    // reference's Kqvals order (point fastest, then element)
    std::vector<double> dev;
    try {
        dev.resize(static_cast<size_t>(ctx->E) * q3);
    } catch (const std::bad_alloc&) {
        return fail(ctx, ADS_ERR_NOMEM, "ads_oil_get_kq: host allocation failed");
    }
    ADS_HIP(ctx, hipMemcpyAsync(dev.data(), ctx->oil.kq, sizeof(double) * dev.size(), hipMemcpyDeviceToHost,
                                ctx->stream));
    ADS_HIP(ctx, hipStreamSynchronize(ctx->stream));
    for (long e = 0; e < ctx->E; ++e)
        for (long q = 0; q < q3; ++q) kq_host[q + q3 * e] = dev[static_cast<size_t>(q * ctx->E + e)];
    return ADS_OK;
}

int ads_host_seg_tables(int n, int p, int* w, int* lo, int* hi, float* tab)
{
    if (!w || !lo || !hi || !tab || n < p || p < 1 || p > 5) return ADS_ERR_INVALID;
    ads::Direction d;
    if (!ads::build_direction(n, p, d).empty()) return ADS_ERR_INVALID;
    std::vector<float> t;
    if (!ads::seg_lu_tables(d, kSegTol, kSegTolW, ads::kSegPadLo, ads::kSegMax, t, *w, *lo, *hi).empty())
        return ADS_ERR_UNSUPPORTED;
    std::copy(t.begin(), t.end(), tab);
    return ADS_OK;
}

int ads_oil_curves(uint64_t seed, int cN, int cL, double* cx, double* cy, double* cz)
{
    if (cN < 1 || cL < 2 || !cx || !cy || !cz) return ADS_ERR_INVALID;
    ads::oil_curves(seed, cN, cL, cx, cy, cz);
    return ADS_OK;
}

int ads_time_kernel(ads_ctx* ctx, int which, double dt, int launches, double* ms_per_launch)
{
    if (check_ctx(ctx) || !ms_per_launch || launches < 1) return ADS_ERR_INVALID;
    if (which < ADS_KERNEL_HEAT_FP64 || which > ADS_KERNEL_SOLVE_SEG_FP32)
        return fail(ctx, ADS_ERR_INVALID, "ads_time_kernel: unknown kernel");
    if (which == ADS_KERNEL_OIL && !ctx->oil.ready)
        return fail(ctx, ADS_ERR_INVALID, "ads_time_kernel: call ads_oil_setup first");
    int rc = ensure_ir_scratch(ctx);
    if (rc) return rc;
    if (which == ADS_KERNEL_SOLVE_SEG_FP32 && (rc = ensure_seg(ctx))) return rc;
    if (!ctx->mu && (rc = dev_alloc(ctx, &ctx->mu, ctx->D))) return rc;
    const ads::BandSet bs = band_set(ctx);
    if ((rc = zero(ctx, ctx->mu, ctx->D))) return rc;
    if ((rc = zero(ctx, ctx->f32a, ctx->D))) return rc;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    ADS_HIP(ctx, hipEventCreate(&e0));
    hipError_t e = hipEventCreate(&e1);
    if (e != hipSuccess) {
        (void)hipEventDestroy(e0);
        return hip_fail(ctx, e, "hipEventCreate");
    }
    e = hipEventRecord(e0, ctx->stream);
    int nparts = 0;
    for (int i = 0; i < launches && e == hipSuccess; ++i) {
        if (which == ADS_KERNEL_HEAT_FP64)
            e = ads::launch_assemble<double>(ctx->p, ads::kFormHeat, 0, ctx->c, ctx->mu, ctx->tb64, dt, ctx->kappa,
                                             ctx->l2_part, ctx->l2_cap, &nparts, ctx->stream);
        else if (which == ADS_KERNEL_OIL)
            e = ads::launch_oil(ctx->p, false, ctx->c, ctx->mu, ctx->tb64, ctx->oil.dev, ctx->oil.kq, dt,
                                ctx->l2_part, ctx->oil.dr_part, ctx->l2_cap, &nparts, ctx->stream);
        else if (which == ADS_KERNEL_SOLVE_FP64)
            e = ctx->part_ok ? ads::launch_part_solve<double>(ctx->p, ctx->mu, ctx->part, ctx->stream)
                             : ads::launch_kron_solve<double>(ctx->p, ctx->mu, ctx->sw64, false, ctx->stream);
        else if (which == ADS_KERNEL_SOLVE_FP64_LU)
            e = ads::launch_kron_solve<double>(ctx->p, ctx->mu, ctx->sw64, false, ctx->stream);
        else if (which == ADS_KERNEL_SOLVE_SEG_FP32)
            e = ads::launch_seg_solve(ctx->p, ctx->f32a, ctx->f32b, ctx->f32c, ctx->seg, ads::SegOut{}, ctx->stream);
        else
            e = ads::launch_kron_residual<float, float>(ctx->p, ctx->c, nullptr, nullptr, ctx->f32b, bs, dt,
                                                        ctx->l2_part, ctx->l2_cap, &nparts, ctx->stream);
    }
    if (e == hipSuccess) e = hipEventRecord(e1, ctx->stream);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    float ms = 0.0f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (e != hipSuccess) return hip_fail(ctx, e, "ads_time_kernel");
    *ms_per_launch = static_cast<double>(ms) / launches;
    return ADS_OK;
}

int ads_form_un(ads_ctx* ctx, int subun)
{
    if (check_ctx(ctx)) return ADS_ERR_INVALID;
    if (subun < 1 || subun > 3) return fail(ctx, ADS_ERR_INVALID, "FormUn: wrong substep (subun must be 1..3)");
    if (subun > 1) return ADS_OK;  // the gradients always come from the current coefficients
    int rc;
    if (!ctx->un && (rc = dev_alloc(ctx, &ctx->un, ctx->D))) return rc;
    ADS_HIP(ctx, hipMemcpyAsync(ctx->un, ctx->c, sizeof(double) * ctx->D, hipMemcpyDeviceToDevice, ctx->stream));
    ctx->un_kept = true;
    ctx->un_version = ctx->version;
    return ADS_OK;
}

int ads_form_rhs_sub(ads_ctx* ctx, int form, int f0_kind, int direction, int substep, double dt, double* F_host,
                     double* l2norm)
{
    if (check_ctx(ctx) || !F_host) return ADS_ERR_INVALID;
    if (direction < 1 || direction > 3 || substep < 1 || substep > 3)
        return fail(ctx, ADS_ERR_INVALID, "Form3DRHS: direction and substep must be 1..3");
    int rc = form_rhs(ctx, form, f0_kind, dt, ctx->F, ctx->l2);
    if (rc) return rc;
    ADS_HIP(ctx, hipMemcpyAsync(F_host, ctx->F, sizeof(double) * ctx->D, hipMemcpyDeviceToHost, ctx->stream));
    ADS_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return read_l2(ctx, l2norm);
}

int ads_sub_step(ads_ctx* ctx, int form, int f0_kind, const double mix[4], int direction, int substep, double dt,
                 double* l2norm)
{
    if (check_ctx(ctx) || !mix) return ADS_ERR_INVALID;
    if (direction < 1 || direction > 3 || substep < 1 || substep > 3)
        return fail(ctx, ADS_ERR_INVALID, "Sub_Step: direction and substep must be 1..3");
    for (int k = 0; k < 4; ++k)
        if (!std::isfinite(mix[k])) return fail(ctx, ADS_ERR_INVALID, "Sub_Step: mix must be finite");
    const bool mass = mix[0] == 1.0 && mix[1] == 0.0 && mix[2] == 0.0 && mix[3] == 0.0;
    const ads::SweepTables<double>* sw = ctx->sw64;
    int rc;
    if (!mass) {
        // the sets are referenced by pointer: no reallocation within the cap
        ctx->mixsets.reserve(64);
        if ((rc = mix_set(ctx, mix, &sw))) return rc;
    }
    if ((rc = form_rhs(ctx, form, f0_kind, dt, ctx->F, ctx->l2))) return rc;
    if (mass) rc = mass_solve64(ctx, ctx->F, false);
    else ADS_HIP(ctx, ads::launch_kron_solve<double>(ctx->p, ctx->F, sw, false, ctx->stream));
    if (rc) return rc;
    std::swap(ctx->c, ctx->F);
    ++ctx->version;
    return read_l2(ctx, l2norm);
}

int ads_last_drained(ads_ctx* ctx, double* drained)
{
    if (check_ctx(ctx) || !drained) return ADS_ERR_INVALID;
    if (!ctx->oil.ready) return fail(ctx, ADS_ERR_INVALID, "ads_last_drained: call ads_oil_setup first");
    return read_scalar(ctx, ctx->oil.drained, drained, "drained");
}

int ads_set_solver(ads_ctx* ctx, int solver)
{
    if (check_ctx(ctx)) return ADS_ERR_INVALID;
    if (solver != ADS_SOLVER_AUTO && solver != ADS_SOLVER_LU && solver != ADS_SOLVER_PARTITIONED)
        return fail(ctx, ADS_ERR_INVALID, "ads_set_solver: unknown solver");
    if (solver == ADS_SOLVER_PARTITIONED && !ctx->part_ok)
        return fail(ctx, ADS_ERR_UNSUPPORTED,
                    "ads_set_solver: the partitioned solve takes lines of at most " +
                        std::to_string(ads::kPartMaxLine) + " entries");
    ctx->solver = solver;
    return ADS_OK;
}

int ads_host_part_tables(int n, int p, int* K, int* S, double* lrow, double* urow, double* gsp, double* hsp)
{
    if (!K || !S) return fail(nullptr, ADS_ERR_INVALID, "ads_host_part_tables: K and S are required");
    ads::Direction d;
    std::string e = ads::build_direction(n, p, d);
    if (!e.empty()) return fail(nullptr, ADS_ERR_INVALID, e);
    const int seg = ads::part_seg(d.dofs);
    if (seg == 0)
        return fail(nullptr, ADS_ERR_UNSUPPORTED,
                    "ads_host_part_tables: lines of at most " + std::to_string(ads::kPartMaxLine) + " entries");
    ads::PartHost ph;
    e = ads::part_tables(d, seg, ph);
    if (!e.empty()) return fail(nullptr, ADS_ERR_SINGULAR, e);
    *K = ph.K;
    *S = ph.S;
    if (lrow) std::copy(ph.lrow.begin(), ph.lrow.end(), lrow);
    if (urow) std::copy(ph.urow.begin(), ph.urow.end(), urow);
    if (gsp) std::copy(ph.gsp.begin(), ph.gsp.end(), gsp);
    if (hsp) std::copy(ph.hsp.begin(), ph.hsp.end(), hsp);
    return ADS_OK;
}

int ads_matrix_create(const double* U, int n, int p, int nelem, const double mix[4], ads_matrix** out)
{
    if (!U || !mix || !out) return fail(nullptr, ADS_ERR_INVALID, "ads_matrix_create: null argument");
    *out = nullptr;
    for (int k = 0; k < 4; ++k)
        if (!std::isfinite(mix[k])) return fail(nullptr, ADS_ERR_INVALID, "ads_matrix_create: mix must be finite");
    ads_matrix* m = new (std::nothrow) ads_matrix();
    if (!m) return fail(nullptr, ADS_ERR_NOMEM, "host allocation failed");
    std::string e = ads::build_direction_knots(U, n, p, nelem, m->dir);
    if (!e.empty()) {
        delete m;
        return fail(nullptr, ADS_ERR_INVALID, "ads_matrix_create: " + e);
    }
    for (int k = 0; k < 4; ++k) m->mix[k] = mix[k];
    ads::mix_band(m->dir, mix, m->band);
    e = ads::band_lu(m->dir.dofs, p, m->band, m->lu, m->piv);
    if (!e.empty()) {
        delete m;
        return fail(nullptr, ADS_ERR_SINGULAR, "ads_matrix_create: " + e);
    }
    *out = m;
    return ADS_OK;
}

int ads_matrix_destroy(ads_matrix* m)
{
    if (!m) return ADS_OK;
    if (m->stream) (void)hipStreamSynchronize(m->stream);
    for (void* ptr : {static_cast<void*>(m->fwd), static_cast<void*>(m->bwd), static_cast<void*>(m->dpiv),
                      static_cast<void*>(m->rhs)})
        if (ptr) (void)hipFree(ptr);
    if (m->stream) (void)hipStreamDestroy(m->stream);
    delete m;
    return ADS_OK;
}

const char* ads_matrix_last_error(const ads_matrix* m) { return m ? m->err.c_str() : g_init_error.c_str(); }

int ads_matrix_dense(const ads_matrix* m, double* dense)
{
    if (!m || !dense) return ADS_ERR_INVALID;
    ads::dense_matrix(m->dir, m->mix, dense);
    return ADS_OK;
}

int ads_matrix_solve(ads_matrix* m, double* rhs, int eqnum)
{
    if (!m || (!rhs && eqnum > 0) || eqnum < 0) return ADS_ERR_INVALID;
    auto mfail = [&](int code, const std::string& msg) {
        m->err = msg;
        return code;
    };
    if (eqnum == 0) return ADS_OK;
    const int N = m->dir.dofs, p = m->dir.p;
    if (p > 9) return mfail(ADS_ERR_UNSUPPORTED, "ads_matrix_solve: the GPU sweeps are built for degrees 1..9");
    hipError_t e = hipSuccess;
    if (!m->stream) {
        std::vector<double> fwd, bwd;
        std::vector<int> fpiv;
        ads::sweep_tables(N, p, m->lu, m->piv, fwd, fpiv, bwd);
        e = hipStreamCreate(&m->stream);
        if (e == hipSuccess) e = hipMalloc(&m->fwd, sizeof(double) * fwd.size());
        if (e == hipSuccess) e = hipMalloc(&m->bwd, sizeof(double) * bwd.size());
        if (e == hipSuccess) e = hipMalloc(&m->dpiv, sizeof(int) * fpiv.size());
        if (e == hipSuccess) e = hipMemcpy(m->fwd, fwd.data(), sizeof(double) * fwd.size(), hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(m->bwd, bwd.data(), sizeof(double) * bwd.size(), hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(m->dpiv, fpiv.data(), sizeof(int) * fpiv.size(), hipMemcpyHostToDevice);
        m->pivots = 0;
        for (int v : fpiv) m->pivots |= (v != 0);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            return mfail(ADS_ERR_HIP, std::string("ads_matrix_solve: set-up: ") + hipGetErrorString(e));
        }
    }
    const size_t count = static_cast<size_t>(N) * eqnum;
    if (count > m->rhs_cap) {
        if (m->rhs) (void)hipFree(m->rhs);
        m->rhs = nullptr;
        m->rhs_cap = 0;
        e = hipMalloc(&m->rhs, sizeof(double) * count);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            return mfail(ADS_ERR_NOMEM, std::string("ads_matrix_solve: hipMalloc: ") + hipGetErrorString(e));
        }
        m->rhs_cap = count;
    }
    const ads::SweepTables<double> st{m->fwd, m->dpiv, m->bwd, N, m->pivots};
    e = hipMemcpyAsync(m->rhs, rhs, sizeof(double) * count, hipMemcpyHostToDevice, m->stream);
    if (e == hipSuccess) e = ads::launch_rows_solve(p, m->rhs, eqnum, st, false, m->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(rhs, m->rhs, sizeof(double) * count, hipMemcpyDeviceToHost, m->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(m->stream);
    if (e != hipSuccess) return mfail(ADS_ERR_HIP, std::string("ads_matrix_solve: ") + hipGetErrorString(e));
    return ADS_OK;
}

}  // extern "C"
