// sweep.hip -- gfx950 kernels of the Kronecker mass-matrix solve.
//
//   k_sweep_lines      banded LU substitution along y or z (SolveOneDirection,
//                      ADS.F90:294-379): one lane per line, lanes along x so
//                      every load/store is a coalesced 512-byte wave access,
//                      software-prefetched in chunks.
//   k_sweep_rows       the same along x, where a line is a contiguous row: a
//                      wave stages 64 rows through LDS tiles (the on-device
//                      replacement of Gather/ReorderRHSFor*/Scatter).
//   k_band_*           FP64 Kronecker residual r = A c - M x for the
//                      FP32 + refinement form.
#include "kernels.hpp"

#include <hip/hip_runtime.h>

namespace ads {

namespace {

constexpr int kWave = 64;

// ---------------------------------------------------------------------------
// banded substitution passes (dgbtrs 'N' and 'T' with dtbsv, restated as
// streaming recurrences over a window of registers).  KL = p, K = 2p.
//   NF  L solve with row interchanges, ascending:   in(j+KL+1) -> out(j)
//   NB  U solve (column form), descending:           in(j-K-1)  -> out(j)
//   TF  U^T solve (row form), ascending:              in(j)      -> out(j)
//   TB  L^T solve + interchanges, descending:         in(j)      -> out(j+KL)
// ---------------------------------------------------------------------------
template <typename T, int KL>
struct PassNF {
    static constexpr bool kAscending = true;
    static constexpr int kIn = KL + 1, kOut = 0;
    T w[KL + 1];
    template <class LD>
    __device__ void prologue(LD ld, int n)
    {
#pragma unroll
        for (int t = 0; t <= KL; ++t) w[t] = t < n ? ld(t) : T(0);
    }
    __device__ T step(int j, T in, const SweepTables<T>& st)
    {
        const int l = st.piv[j];
        if (l != 0) {
#pragma unroll
            for (int t = 1; t <= KL; ++t)
                if (t == l) {
                    const T tmp = w[0];
                    w[0] = w[t];
                    w[t] = tmp;
                }
        }
        const T xj = w[0];
#pragma unroll
        for (int t = 1; t <= KL; ++t) w[t] -= st.fwd[j * KL + t - 1] * xj;
#pragma unroll
        for (int t = 0; t < KL; ++t) w[t] = w[t + 1];
        w[KL] = in;
        return xj;
    }
    template <class ST>
    __device__ void epilogue(ST, int)
    {
    }
};

template <typename T, int KL>
struct PassNB {
    static constexpr bool kAscending = false;
    static constexpr int K = 2 * KL;
    static constexpr int kIn = -(K + 1), kOut = 0;
    T v[K + 1];
    template <class LD>
    __device__ void prologue(LD ld, int n)
    {
#pragma unroll
        for (int t = 0; t <= K; ++t) v[K - t] = (n - 1 - t >= 0) ? ld(n - 1 - t) : T(0);
    }
    __device__ T step(int j, T in, const SweepTables<T>& st)
    {
        const T* b = st.bwd + static_cast<size_t>(j) * (K + 1);
        const T xj = v[K] * b[0];
#pragma unroll
        for (int t = 1; t <= K; ++t) v[K - t] -= b[t] * xj;
#pragma unroll
        for (int t = K; t > 0; --t) v[t] = v[t - 1];
        v[0] = in;
        return xj;
    }
    template <class ST>
    __device__ void epilogue(ST, int)
    {
    }
};

template <typename T, int KL>
struct PassTF {
    static constexpr bool kAscending = true;
    static constexpr int K = 2 * KL;
    static constexpr int kIn = 0, kOut = 0;
    T h[K];
    template <class LD>
    __device__ void prologue(LD, int)
    {
#pragma unroll
        for (int t = 0; t < K; ++t) h[t] = T(0);
    }
    __device__ T step(int j, T in, const SweepTables<T>& st)
    {
        const T* b = st.bwd + static_cast<size_t>(j) * (K + 1);
        T s = in;
#pragma unroll
        for (int t = K; t >= 1; --t) s -= b[t] * h[t - 1];
        const T xj = s * b[0];
#pragma unroll
        for (int t = K - 1; t > 0; --t) h[t] = h[t - 1];
        h[0] = xj;
        return xj;
    }
    template <class ST>
    __device__ void epilogue(ST, int)
    {
    }
};

template <typename T, int KL>
struct PassTB {
    static constexpr bool kAscending = false;
    static constexpr int kIn = 0, kOut = KL;
    T w[KL + 1];
    template <class LD>
    __device__ void prologue(LD, int)
    {
#pragma unroll
        for (int t = 0; t <= KL; ++t) w[t] = T(0);
    }
    __device__ T step(int j, T in, const SweepTables<T>& st)
    {
#pragma unroll
        for (int t = KL; t > 0; --t) w[t] = w[t - 1];
        T s = T(0);
#pragma unroll
        for (int t = 1; t <= KL; ++t) s += w[t] * st.fwd[j * KL + t - 1];
        w[0] = in - s;
        const int l = st.piv[j];
        if (l != 0) {
#pragma unroll
            for (int t = 1; t <= KL; ++t)
                if (t == l) {
                    const T tmp = w[0];
                    w[0] = w[t];
                    w[t] = tmp;
                }
        }
        return w[KL];  // b(j + KL) is final once step j is done
    }
    template <class ST>
    __device__ void epilogue(ST st, int n)
    {
#pragma unroll
        for (int t = 0; t < KL; ++t)
            if (t < n) st(t, w[t]);
    }
};

// One lane per line; element k of the line at base[k*stride].  Inputs are
// fetched CH at a time one chunk ahead of the recurrence.
template <typename T, int KL, template <typename, int> class PASS, int CH>
__device__ void run_line(T* __restrict__ base, long stride, int n, const SweepTables<T>& st)
{
    PASS<T, KL> ps;
    auto ld = [&](int k) { return base[static_cast<long>(k) * stride]; };
    ps.prologue(ld, n);
    const int nch = (n + CH - 1) / CH;
    auto fetch = [&](int ch, T* buf) {
#pragma unroll
        for (int i = 0; i < CH; ++i) {
            const int j = PASS<T, KL>::kAscending ? ch * CH + i : n - 1 - (ch * CH + i);
            const int k = j + PASS<T, KL>::kIn;
            buf[i] = (ch < nch && j >= 0 && j < n && k >= 0 && k < n) ? base[static_cast<long>(k) * stride] : T(0);
        }
    };
    T cur[CH], nxt[CH];
    fetch(0, cur);
    for (int ch = 0; ch < nch; ++ch) {
        fetch(ch + 1, nxt);
#pragma unroll
        for (int i = 0; i < CH; ++i) {
            const int j = PASS<T, KL>::kAscending ? ch * CH + i : n - 1 - (ch * CH + i);
            if (j >= 0 && j < n) {
                const T o = ps.step(j, cur[i], st);
                const int k = j + PASS<T, KL>::kOut;
                if (k < n) base[static_cast<long>(k) * stride] = o;
            }
        }
#pragma unroll
        for (int i = 0; i < CH; ++i) cur[i] = nxt[i];
    }
    ps.epilogue([&](int k, T v) { base[static_cast<long>(k) * stride] = v; }, n);
}

// Lines along a strided axis.  Line L = (i, j) with i in [0, ni) fastest:
// base = data + i + j*stride_j; line element k at base + k*stride_k.
template <typename T, int KL, template <typename, int> class PASS>
__global__ __launch_bounds__(256) void k_sweep_lines(T* __restrict__ data, int ni, long stride_j, long stride_k,
                                                     int nlines, SweepTables<T> st)
{
    const int L = blockIdx.x * blockDim.x + threadIdx.x;
    if (L >= nlines) return;
    const int i = L % ni, j = L / ni;
    run_line<T, KL, PASS, 16>(data + i + static_cast<long>(j) * stride_j, stride_k, st.n, st);
}

// Contiguous rows (the x direction): a wave owns 64 rows and moves them
// through LDS in CH-wide tiles so every global access is row-contiguous.
template <typename T, int KL, template <typename, int> class PASS>
__global__ __launch_bounds__(kWave) void k_sweep_rows(T* __restrict__ data, int nrows, SweepTables<T> st)
{
    constexpr int CH = 16;
    constexpr int LD = CH + 1;
    using PS = PASS<T, KL>;
    __shared__ T tin[kWave * LD];
    __shared__ T tout[kWave * LD];
    const int lane = threadIdx.x;
    const int row0 = blockIdx.x * kWave;
    const int n = st.n;
    const long rowlen = n;
    const int myrow = row0 + lane;
    const bool active = myrow < nrows;
    T* __restrict__ mine = data + static_cast<long>(active ? myrow : row0) * rowlen;

    PS ps;
    ps.prologue([&](int k) { return active ? mine[k] : T(0); }, n);
    const int nch = (n + CH - 1) / CH;
    constexpr int RPI = kWave / CH;  // rows per cooperative instruction
    for (int ch = 0; ch < nch; ++ch) {
        // j of tile column c: ascending j0 + c, descending j0 - (CH-1) + c
        const int jlo = PS::kAscending ? ch * CH : n - 1 - (ch * CH + CH - 1);
#pragma unroll
        for (int it = 0; it < kWave / RPI; ++it) {
            const int r = it * RPI + lane / CH;
            const int cidx = lane % CH;
            const int j = jlo + cidx;
            const int k = j + PS::kIn;
            const int grow = row0 + r;
            T v = T(0);
            if (grow < nrows && j >= 0 && j < n && k >= 0 && k < n) v = data[static_cast<long>(grow) * rowlen + k];
            tin[r * LD + cidx] = v;
        }
        __syncthreads();
#pragma unroll
        for (int i = 0; i < CH; ++i) {
            const int cidx = PS::kAscending ? i : CH - 1 - i;
            const int j = jlo + cidx;
            if (j >= 0 && j < n) tout[lane * LD + cidx] = ps.step(j, tin[lane * LD + cidx], st);
        }
        __syncthreads();
#pragma unroll
        for (int it = 0; it < kWave / RPI; ++it) {
            const int r = it * RPI + lane / CH;
            const int cidx = lane % CH;
            const int j = jlo + cidx;
            const int k = j + PS::kOut;
            const int grow = row0 + r;
            if (grow < nrows && j >= 0 && j < n && k < n) data[static_cast<long>(grow) * rowlen + k] = tout[r * LD + cidx];
        }
        __syncthreads();
    }
    if (active) ps.epilogue([&](int k, T v) { mine[k] = v; }, n);
}

// ---------------------------------------------------------------------------
// Kronecker residual r = M(c - x) - dtk K c, staged per axis:
//   X: e = Mx(c - x) - dtk Kx c,   m = Mx c
//   Y: f = My e - dtk Ky m,        g = My m
//   Z: r = Mz f - dtk Kz g
// one thread per output entry; neighbours along the axis come from cache.
// ---------------------------------------------------------------------------
template <int P>
__device__ __forceinline__ void band_row(const double* __restrict__ band, int i, int n, double* a)
{
    constexpr int KL = P - 1, BW = 2 * KL + 1;
#pragma unroll
    for (int t = -KL; t <= KL; ++t) {
        const int j = i + t;
        a[t + KL] = (j >= 0 && j < n) ? band[static_cast<long>(j) * BW + (KL + i - j)] : 0.0;
    }
}

template <int P, typename TX>
__global__ __launch_bounds__(256) void k_band_x(const double* __restrict__ c, const TX* __restrict__ x,
                                                 double* __restrict__ e, double* __restrict__ m, BandSet bs,
                                                 double dtk, long D)
{
    constexpr int KL = P - 1;
    const long idx = blockIdx.x * static_cast<long>(blockDim.x) + threadIdx.x;
    if (idx >= D) return;
    const int nx = bs.n[0];
    const int i = static_cast<int>(idx % nx);
    double am[2 * KL + 1], ak[2 * KL + 1];
    band_row<P>(bs.mass[0], i, nx, am);
    band_row<P>(bs.stiff[0], i, nx, ak);
    double se = 0.0, sm = 0.0, sk = 0.0;
#pragma unroll
    for (int t = -KL; t <= KL; ++t) {
        const int j = i + t;
        if (j < 0 || j >= nx) continue;
        const double cv = c[idx + t];
        const double dv = cv - static_cast<double>(x[idx + t]);
        se += am[t + KL] * dv;
        sm += am[t + KL] * cv;
        sk += ak[t + KL] * cv;
    }
    e[idx] = se - dtk * sk;
    m[idx] = sm;
}

template <int P, typename TR>
__global__ __launch_bounds__(256) void k_band_yz(const double* __restrict__ a_in, const double* __restrict__ b_in,
                                                  double* __restrict__ out0, double* __restrict__ out1,
                                                  TR* __restrict__ r_out, const double* __restrict__ mass,
                                                  const double* __restrict__ stiff, int naxis, long stride,
                                                  long inner, double dtk, long D)
{
    // out0 = Maxis a - dtk Kaxis b ; out1 = Maxis b (when out1 != nullptr)
    // r_out (when != nullptr) receives out0 instead of out0.
    constexpr int KL = P - 1;
    const long idx = blockIdx.x * static_cast<long>(blockDim.x) + threadIdx.x;
    if (idx >= D) return;
    const int i = static_cast<int>((idx / inner) % naxis);
    double am[2 * KL + 1], ak[2 * KL + 1];
    band_row<P>(mass, i, naxis, am);
    band_row<P>(stiff, i, naxis, ak);
    double s0 = 0.0, s1 = 0.0, sk = 0.0;
#pragma unroll
    for (int t = -KL; t <= KL; ++t) {
        const int j = i + t;
        if (j < 0 || j >= naxis) continue;
        const double av = a_in[idx + t * stride];
        const double bv = b_in[idx + t * stride];
        s0 += am[t + KL] * av;
        sk += ak[t + KL] * bv;
        s1 += am[t + KL] * bv;
    }
    const double o0 = s0 - dtk * sk;
    if (r_out) r_out[idx] = static_cast<TR>(o0);
    else out0[idx] = o0;
    if (out1) out1[idx] = s1;
}

__global__ __launch_bounds__(256) void k_convert_d2f(const double* __restrict__ in, float* __restrict__ out, long n)
{
    const long i = blockIdx.x * static_cast<long>(blockDim.x) + threadIdx.x;
    if (i < n) out[i] = static_cast<float>(in[i]);
}

__global__ __launch_bounds__(256) void k_add_f2d(const float* __restrict__ a, const float* __restrict__ b,
                                                  double* __restrict__ out, long n)
{
    const long i = blockIdx.x * static_cast<long>(blockDim.x) + threadIdx.x;
    if (i >= n) return;
    const double base = a ? static_cast<double>(a[i]) : out[i];
    out[i] = base + static_cast<double>(b[i]);
}

template <typename T, int KL>
hipError_t kron_solve_p(T* F, const SweepTables<T> sw[3], bool trans, hipStream_t s)
{
    const int nx = sw[0].n, ny = sw[1].n, nz = sw[2].n;
    const long sxy = static_cast<long>(nx) * ny;
    // x: rows of length nx, nrows = ny*nz
    {
        const int nrows = ny * nz;
        dim3 grid((nrows + kWave - 1) / kWave);
        if (!trans) {
            hipLaunchKernelGGL((k_sweep_rows<T, KL, PassNF>), grid, dim3(kWave), 0, s, F, nrows, sw[0]);
            hipLaunchKernelGGL((k_sweep_rows<T, KL, PassNB>), grid, dim3(kWave), 0, s, F, nrows, sw[0]);
        } else {
            hipLaunchKernelGGL((k_sweep_rows<T, KL, PassTF>), grid, dim3(kWave), 0, s, F, nrows, sw[0]);
            hipLaunchKernelGGL((k_sweep_rows<T, KL, PassTB>), grid, dim3(kWave), 0, s, F, nrows, sw[0]);
        }
    }
    // y: lines (x, z), element stride nx
    {
        const int nlines = nx * nz;
        dim3 grid((nlines + 255) / 256);
        if (!trans) {
            hipLaunchKernelGGL((k_sweep_lines<T, KL, PassNF>), grid, dim3(256), 0, s, F, nx, sxy, static_cast<long>(nx),
                               nlines, sw[1]);
            hipLaunchKernelGGL((k_sweep_lines<T, KL, PassNB>), grid, dim3(256), 0, s, F, nx, sxy, static_cast<long>(nx),
                               nlines, sw[1]);
        } else {
            hipLaunchKernelGGL((k_sweep_lines<T, KL, PassTF>), grid, dim3(256), 0, s, F, nx, sxy, static_cast<long>(nx),
                               nlines, sw[1]);
            hipLaunchKernelGGL((k_sweep_lines<T, KL, PassTB>), grid, dim3(256), 0, s, F, nx, sxy, static_cast<long>(nx),
                               nlines, sw[1]);
        }
    }
    // z: lines (x, y) = contiguous index in [0, nx*ny), element stride nx*ny
    {
        const int nlines = nx * ny;
        dim3 grid((nlines + 255) / 256);
        if (!trans) {
            hipLaunchKernelGGL((k_sweep_lines<T, KL, PassNF>), grid, dim3(256), 0, s, F, nlines, 0L, sxy, nlines, sw[2]);
            hipLaunchKernelGGL((k_sweep_lines<T, KL, PassNB>), grid, dim3(256), 0, s, F, nlines, 0L, sxy, nlines, sw[2]);
        } else {
            hipLaunchKernelGGL((k_sweep_lines<T, KL, PassTF>), grid, dim3(256), 0, s, F, nlines, 0L, sxy, nlines, sw[2]);
            hipLaunchKernelGGL((k_sweep_lines<T, KL, PassTB>), grid, dim3(256), 0, s, F, nlines, 0L, sxy, nlines, sw[2]);
        }
    }
    return hipGetLastError();
}

template <int P, typename TX, typename TR>
hipError_t residual_p(const double* c, const TX* x, TR* r, const BandSet& bs, double dtk, double* w0, double* w1,
                      double* w2, double* w3, hipStream_t s)
{
    const long nx = bs.n[0], ny = bs.n[1], nz = bs.n[2];
    const long D = nx * ny * nz;
    dim3 grid(static_cast<unsigned>((D + 255) / 256));
    hipLaunchKernelGGL((k_band_x<P, TX>), grid, dim3(256), 0, s, c, x, w0, w1, bs, dtk, D);
    // Y: f = My e - dtk Ky m -> w2 ; g = My m -> w3
    hipLaunchKernelGGL((k_band_yz<P, TR>), grid, dim3(256), 0, s, w0, w1, w2, w3, static_cast<TR*>(nullptr),
                       bs.mass[1], bs.stiff[1], static_cast<int>(ny), nx, nx, dtk, D);
    // Z: r = Mz f - dtk Kz g
    hipLaunchKernelGGL((k_band_yz<P, TR>), grid, dim3(256), 0, s, w2, w3, static_cast<double*>(nullptr),
                       static_cast<double*>(nullptr), r, bs.mass[2], bs.stiff[2], static_cast<int>(nz), nx * ny,
                       nx * ny, dtk, D);
    return hipGetLastError();
}

}  // namespace

template <typename T>
hipError_t launch_kron_solve(int p, T* F, const SweepTables<T> sw[3], bool trans, hipStream_t s)
{
    switch (p) {
    case 1: return kron_solve_p<T, 1>(F, sw, trans, s);
    case 2: return kron_solve_p<T, 2>(F, sw, trans, s);
    case 3: return kron_solve_p<T, 3>(F, sw, trans, s);
    case 4: return kron_solve_p<T, 4>(F, sw, trans, s);
    default: return hipErrorInvalidValue;
    }
}

template hipError_t launch_kron_solve<double>(int, double*, const SweepTables<double>[3], bool, hipStream_t);
template hipError_t launch_kron_solve<float>(int, float*, const SweepTables<float>[3], bool, hipStream_t);

template <typename TX, typename TR>
hipError_t launch_kron_residual(int p, const double* c, const TX* x, TR* r, const BandSet& bs, double dtk,
                                double* w0, double* w1, double* w2, hipStream_t s)
{
    // the Y stage writes (f, g); r's own storage is not used as a temporary,
    // so a fourth buffer is carved from w2's tail by the caller: w2 holds 2*D.
    const long D = static_cast<long>(bs.n[0]) * bs.n[1] * bs.n[2];
    double* w3 = w2 + D;
    switch (p) {
    case 1: return residual_p<2, TX, TR>(c, x, r, bs, dtk, w0, w1, w2, w3, s);
    case 2: return residual_p<3, TX, TR>(c, x, r, bs, dtk, w0, w1, w2, w3, s);
    case 3: return residual_p<4, TX, TR>(c, x, r, bs, dtk, w0, w1, w2, w3, s);
    case 4: return residual_p<5, TX, TR>(c, x, r, bs, dtk, w0, w1, w2, w3, s);
    default: return hipErrorInvalidValue;
    }
}

template hipError_t launch_kron_residual<float, float>(int, const double*, const float*, float*, const BandSet&,
                                                       double, double*, double*, double*, hipStream_t);
template hipError_t launch_kron_residual<double, float>(int, const double*, const double*, float*, const BandSet&,
                                                        double, double*, double*, double*, hipStream_t);
template hipError_t launch_kron_residual<double, double>(int, const double*, const double*, double*,
                                                         const BandSet&, double, double*, double*, double*,
                                                         hipStream_t);

hipError_t launch_convert_d2f(const double* in, float* out, int64_t n, hipStream_t s)
{
    hipLaunchKernelGGL(k_convert_d2f, dim3(static_cast<unsigned>((n + 255) / 256)), dim3(256), 0, s, in, out,
                       static_cast<long>(n));
    return hipGetLastError();
}

hipError_t launch_add_f2d(const float* a, const float* b, double* out, int64_t n, hipStream_t s)
{
    hipLaunchKernelGGL(k_add_f2d, dim3(static_cast<unsigned>((n + 255) / 256)), dim3(256), 0, s, a, b, out,
                       static_cast<long>(n));
    return hipGetLastError();
}

}  // namespace ads
