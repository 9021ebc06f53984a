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

// The per-step factor entries are the same for every line of a direction:
// the block copies them into LDS once, so the recurrences read them with
// broadcast LDS loads instead of serialised scalar loads.
template <typename T, int KL>
constexpr size_t sweep_table_bytes(int n)
{
    return static_cast<size_t>(n) * (3 * KL + 1) * sizeof(T) + static_cast<size_t>(n) * sizeof(int);
}

template <typename T, int KL>
__device__ __forceinline__ SweepTables<T> stage_tables(const SweepTables<T>& g, unsigned char* smem)
{
    const int n = g.n;
    T* fwd = reinterpret_cast<T*>(smem);
    T* bwd = fwd + n * KL;
    int* piv = reinterpret_cast<int*>(bwd + n * (2 * KL + 1));
    for (int i = threadIdx.x; i < n * KL; i += blockDim.x) fwd[i] = g.fwd[i];
    for (int i = threadIdx.x; i < n * (2 * KL + 1); i += blockDim.x) bwd[i] = g.bwd[i];
    for (int i = threadIdx.x; i < n; i += blockDim.x) piv[i] = g.piv[i];
    __syncthreads();
    SweepTables<T> t;
    t.fwd = fwd;
    t.bwd = bwd;
    t.piv = piv;
    t.n = n;
    return t;
}

// One lane per line; element k of the line at base[k*stride].  Inputs are
// fetched CH at a time one chunk ahead of the recurrence (double buffer in
// VGPRs), so every lane keeps CH loads in flight.
template <typename T, int KL, template <typename, int> class PASS, int CH>
__device__ __forceinline__ void run_line(T* __restrict__ base, long stride, int n, const SweepTables<T>& st)
{
    using PS = PASS<T, KL>;
    PS ps;
    ps.prologue([&](int k) { return base[static_cast<long>(k) * stride]; }, n);
    const int nch = (n + CH - 1) / CH;
    auto fetch = [&](int ch, T* buf) {
#pragma unroll
        for (int i = 0; i < CH; ++i) {
            const int j = PS::kAscending ? ch * CH + i : n - 1 - (ch * CH + i);
            const int k = j + PS::kIn;
            buf[i] = (ch < nch && j >= 0 && j < n && k >= 0 && k < n) ? base[static_cast<long>(k) * stride] : T(0);
        }
    };
    T cur[CH], nxt[CH];
    fetch(0, cur);
    for (int ch = 0; ch < nch; ++ch) {
        fetch(ch + 1, nxt);
#pragma unroll
        for (int i = 0; i < CH; ++i) {
            const int j = PS::kAscending ? ch * CH + i : n - 1 - (ch * CH + i);
            if (j >= 0 && j < n) {
                const T o = ps.step(j, cur[i], st);
                const int k = j + PS::kOut;
                if (k < n) base[static_cast<long>(k) * stride] = o;
            }
        }
#pragma unroll
        for (int i = 0; i < CH; ++i) cur[i] = nxt[i];
    }
    ps.epilogue([&](int k, T v) { base[static_cast<long>(k) * stride] = v; }, n);
}

template <typename T>
struct LineChunk {
    static constexpr int value = sizeof(T) == 8 ? 32 : 48;
};

// Lines along a strided axis, both passes of the solve in one launch (the
// second pass re-reads what the lane itself just wrote, partly from cache).
// Line L = (i, j) with i in [0, ni) fastest: base = data + i + j*stride_j;
// line element k at base + k*stride_k.
template <typename T, int KL, bool TRANS>
__global__ __launch_bounds__(256) void k_sweep_lines(T* __restrict__ data, int ni, long stride_j, long stride_k,
                                                     int nlines, SweepTables<T> gst)
{
    extern __shared__ unsigned char smem[];
    const SweepTables<T> st = stage_tables<T, KL>(gst, smem);
    const int L = blockIdx.x * blockDim.x + threadIdx.x;
    if (L >= nlines) return;
    const int i = L % ni, j = L / ni;
    T* base = data + i + static_cast<long>(j) * stride_j;
    constexpr int CH = LineChunk<T>::value;
    if constexpr (!TRANS) {
        run_line<T, KL, PassNF, CH>(base, stride_k, st.n, st);
        run_line<T, KL, PassNB, CH>(base, stride_k, st.n, st);
    } else {
        run_line<T, KL, PassTF, CH>(base, stride_k, st.n, st);
        run_line<T, KL, PassTB, CH>(base, stride_k, st.n, st);
    }
}

// Contiguous rows (the x direction): a wave owns 64 rows and moves them
// through LDS in CH-wide tiles so every global access is row-contiguous; the
// next tile is fetched into VGPRs while the current one is solved.
template <typename T, int KL, template <typename, int> class PASS>
__device__ __forceinline__ void run_rows(T* __restrict__ data, int nrows, int row0, int lane,
                                         const SweepTables<T>& st, T* tin, T* tout)
{
    constexpr int CH = 16;
    constexpr int LDT = CH + 1;
    constexpr int RPI = kWave / CH;      // rows per cooperative instruction
    constexpr int NIT = kWave / RPI;     // cooperative instructions per tile
    using PS = PASS<T, KL>;
    const int n = st.n;
    const long rowlen = n;
    const int myrow = row0 + lane;
    const bool active = myrow < nrows;
    T* __restrict__ mine = data + static_cast<long>(active ? myrow : row0) * rowlen;
    const int cidx = lane % CH;
    const int rsub = lane / CH;

    PS ps;
    ps.prologue([&](int k) { return active ? mine[k] : T(0); }, n);
    const int nch = (n + CH - 1) / CH;
    auto jlo_of = [&](int ch) { return PS::kAscending ? ch * CH : n - 1 - (ch * CH + CH - 1); };
    auto fetch = [&](int ch, T* buf) {
        const int j = jlo_of(ch) + cidx;
        const int k = j + PS::kIn;
#pragma unroll
        for (int it = 0; it < NIT; ++it) {
            const int grow = row0 + it * RPI + rsub;
            buf[it] = (ch < nch && grow < nrows && j >= 0 && j < n && k >= 0 && k < n)
                          ? data[static_cast<long>(grow) * rowlen + k]
                          : T(0);
        }
    };
    T cur[NIT], nxt[NIT];
    fetch(0, cur);
    for (int ch = 0; ch < nch; ++ch) {
        fetch(ch + 1, nxt);
        const int jlo = jlo_of(ch);
#pragma unroll
        for (int it = 0; it < NIT; ++it) tin[(it * RPI + rsub) * LDT + cidx] = cur[it];
        __syncthreads();
#pragma unroll
        for (int i = 0; i < CH; ++i) {
            const int c = PS::kAscending ? i : CH - 1 - i;
            const int j = jlo + c;
            if (j >= 0 && j < n) tout[lane * LDT + c] = ps.step(j, tin[lane * LDT + c], st);
        }
        __syncthreads();
        {
            const int j = jlo + cidx;
            const int k = j + PS::kOut;
#pragma unroll
            for (int it = 0; it < NIT; ++it) {
                const int grow = row0 + it * RPI + rsub;
                if (grow < nrows && j >= 0 && j < n && k < n)
                    data[static_cast<long>(grow) * rowlen + k] = tout[(it * RPI + rsub) * LDT + cidx];
            }
        }
#pragma unroll
        for (int it = 0; it < NIT; ++it) cur[it] = nxt[it];
    }
    if (active) ps.epilogue([&](int k, T v) { mine[k] = v; }, n);
}

template <typename T, int KL, bool TRANS>
__global__ __launch_bounds__(kWave) void k_sweep_rows(T* __restrict__ data, int nrows, SweepTables<T> gst)
{
    constexpr int LDT = 16 + 1;
    __shared__ T tin[kWave * LDT];
    __shared__ T tout[kWave * LDT];
    extern __shared__ unsigned char smem[];
    const SweepTables<T> st = stage_tables<T, KL>(gst, smem);
    const int lane = threadIdx.x;
    const int row0 = blockIdx.x * kWave;
    if constexpr (!TRANS) {
        run_rows<T, KL, PassNF>(data, nrows, row0, lane, st, tin, tout);
        __syncthreads();
        run_rows<T, KL, PassNB>(data, nrows, row0, lane, st, tin, tout);
    } else {
        run_rows<T, KL, PassTF>(data, nrows, row0, lane, st, tin, tout);
        __syncthreads();
        run_rows<T, KL, PassTB>(data, nrows, row0, lane, st, tin, tout);
    }
}

// ---------------------------------------------------------------------------
// Kronecker residual r = M(c - x) - dtk K c, staged per axis:
//   X: e = Mx(c - x) - dtk Kx c,   m = Mx c       (thread per entry; the
//                                                   x neighbours share lines)
//   Y: f = My e - dtk Ky m,        g = My m       (lane per y line, streamed)
//   Z: r = Mz f - dtk Kz g                        (lane per z line, streamed)
// Row tables hold A(i, i-KL+t), t = 0..2KL, zero outside the matrix.
// ---------------------------------------------------------------------------
template <int P, typename TX>
__global__ __launch_bounds__(256) void k_band_x(const double* __restrict__ c, const TX* __restrict__ x,
                                                 double* __restrict__ e, double* __restrict__ m,
                                                 const double* __restrict__ mrow, const double* __restrict__ krow,
                                                 int nx, double dtk, long D)
{
    constexpr int KL = P - 1, BW = 2 * KL + 1;
    const long idx = blockIdx.x * static_cast<long>(blockDim.x) + threadIdx.x;
    if (idx >= D) return;
    const int i = static_cast<int>(idx % nx);
    double se = 0.0, sm = 0.0, sk = 0.0;
#pragma unroll
    for (int t = 0; t < BW; ++t) {
        const int j = i - KL + t;
        if (j < 0 || j >= nx) continue;
        const double cv = c[idx - KL + t];
        const double dv = x ? cv - static_cast<double>(x[idx - KL + t]) : cv;
        const double am = mrow[i * BW + t], ak = krow[i * BW + t];
        se = fma(am, dv, se);
        sm = fma(am, cv, sm);
        sk = fma(ak, cv, sk);
    }
    e[idx] = fma(-dtk, sk, se);
    m[idx] = sm;
}

// out0 = M a - dtk K b  (to out0, or rounded to r_out when r_out != nullptr)
// out1 = M b            (when out1 != nullptr)
template <int P, typename TR>
__global__ __launch_bounds__(256) void k_band_lines(const double* __restrict__ a_in, const double* __restrict__ b_in,
                                                     double* __restrict__ out0, double* __restrict__ out1,
                                                     TR* __restrict__ r_out, const double* __restrict__ mrow_g,
                                                     const double* __restrict__ krow_g, int n, int ni, long stride_j,
                                                     long stride_k, int nlines, double dtk)
{
    constexpr int KL = P - 1, BW = 2 * KL + 1, CH = 16;
    extern __shared__ unsigned char smem[];
    double* mrow = reinterpret_cast<double*>(smem);
    double* krow = mrow + n * BW;
    for (int q = threadIdx.x; q < n * BW; q += blockDim.x) {
        mrow[q] = mrow_g[q];
        krow[q] = krow_g[q];
    }
    __syncthreads();
    const int L = blockIdx.x * blockDim.x + threadIdx.x;
    if (L >= nlines) return;
    const long base = (L % ni) + static_cast<long>(L / ni) * stride_j;
    double wa[BW], wb[BW];
#pragma unroll
    for (int t = 0; t < BW; ++t) {
        const int k = t - KL;
        wa[t] = (k >= 0 && k < n) ? a_in[base + k * stride_k] : 0.0;
        wb[t] = (k >= 0 && k < n) ? b_in[base + k * stride_k] : 0.0;
    }
    const int nch = (n + CH - 1) / CH;
    auto fetch = [&](int ch, double* ba, double* bb) {
#pragma unroll
        for (int q = 0; q < CH; ++q) {
            const int k = ch * CH + q + KL + 1;  // input entering the window after output ch*CH+q
            const bool ok = ch < nch && k < n;
            ba[q] = ok ? a_in[base + k * stride_k] : 0.0;
            bb[q] = ok ? b_in[base + k * stride_k] : 0.0;
        }
    };
    double ca[CH], cb[CH], na[CH], nb[CH];
    fetch(0, ca, cb);
    for (int ch = 0; ch < nch; ++ch) {
        fetch(ch + 1, na, nb);
#pragma unroll
        for (int q = 0; q < CH; ++q) {
            const int j = ch * CH + q;
            if (j < n) {
                double s0 = 0.0, s1 = 0.0, sk = 0.0;
#pragma unroll
                for (int t = 0; t < BW; ++t) {
                    const double am = mrow[j * BW + t], ak = krow[j * BW + t];
                    s0 = fma(am, wa[t], s0);
                    s1 = fma(am, wb[t], s1);
                    sk = fma(ak, wb[t], sk);
                }
                const double o0 = fma(-dtk, sk, s0);
                if (r_out) r_out[base + j * stride_k] = static_cast<TR>(o0);
                else out0[base + j * stride_k] = o0;
                if (out1) out1[base + j * stride_k] = s1;
            }
#pragma unroll
            for (int t = 0; t < BW - 1; ++t) {
                wa[t] = wa[t + 1];
                wb[t] = wb[t + 1];
            }
            wa[BW - 1] = ca[q];
            wb[BW - 1] = cb[q];
        }
#pragma unroll
        for (int q = 0; q < CH; ++q) {
            ca[q] = na[q];
            cb[q] = nb[q];
        }
    }
}

__global__ __launch_bounds__(256) void k_dot_partials(const double* __restrict__ a, const double* __restrict__ b,
                                                      long n, double* __restrict__ part)
{
    __shared__ double red[256];
    double acc = 0.0;
    for (long i = blockIdx.x * 256L + threadIdx.x; i < n; i += 256L * gridDim.x) acc = fma(a[i], b[i], acc);
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

__global__ __launch_bounds__(256) void k_sum_partials(const double* __restrict__ part, int n, double* __restrict__ out)
{
    __shared__ double red[256];
    double s = 0.0;
    for (int i = threadIdx.x; i < n; i += 256) s += part[i];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = red[0];
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
    const int xrows = ny * nz, ylines = nx * nz, zlines = nx * ny;
    const size_t bx = sweep_table_bytes<T, KL>(nx), by = sweep_table_bytes<T, KL>(ny), bz = sweep_table_bytes<T, KL>(nz);
    // LDS budget: rows kernel = 17.4 KB of tiles + tables, lines kernel = tables (64 KB per block)
    if (bx > 47104 || by > 61440 || bz > 61440) return hipErrorInvalidValue;  // line too long for the LDS tables
    dim3 gx((xrows + kWave - 1) / kWave), gy((ylines + 255) / 256), gz((zlines + 255) / 256);
    if (!trans) {
        hipLaunchKernelGGL((k_sweep_rows<T, KL, false>), gx, dim3(kWave), bx, s, F, xrows, sw[0]);
        hipLaunchKernelGGL((k_sweep_lines<T, KL, false>), gy, dim3(256), by, s, F, nx, sxy, static_cast<long>(nx),
                           ylines, sw[1]);
        hipLaunchKernelGGL((k_sweep_lines<T, KL, false>), gz, dim3(256), bz, s, F, zlines, 0L, sxy, zlines, sw[2]);
    } else {
        hipLaunchKernelGGL((k_sweep_rows<T, KL, true>), gx, dim3(kWave), bx, s, F, xrows, sw[0]);
        hipLaunchKernelGGL((k_sweep_lines<T, KL, true>), gy, dim3(256), by, s, F, nx, sxy, static_cast<long>(nx),
                           ylines, sw[1]);
        hipLaunchKernelGGL((k_sweep_lines<T, KL, true>), gz, dim3(256), bz, s, F, zlines, 0L, sxy, zlines, sw[2]);
    }
    return hipGetLastError();
}

template <int P, typename TX, typename TR>
hipError_t residual_p(const double* c, const TX* x, TR* r, const BandSet& bs, double dtk, double* w0, double* w1,
                      double* w2, double* w3, hipStream_t s)
{
    const int nx = bs.n[0], ny = bs.n[1], nz = bs.n[2];
    const long D = static_cast<long>(nx) * ny * nz;
    const long sxy = static_cast<long>(nx) * ny;
    hipLaunchKernelGGL((k_band_x<P, TX>), dim3(static_cast<unsigned>((D + 255) / 256)), dim3(256), 0, s, c, x, w0, w1,
                       bs.mass[0], bs.stiff[0], nx, dtk, D);
    // Y: f = My e - dtk Ky m -> w2 ; g = My m -> w3   (lines (x, z))
    const int ylines = nx * nz;
    const size_t by = 2 * sizeof(double) * (2 * P - 1) * static_cast<size_t>(ny);
    const size_t bz = 2 * sizeof(double) * (2 * P - 1) * static_cast<size_t>(nz);
    if (by > 61440 || bz > 61440) return hipErrorInvalidValue;  // line too long for the LDS tables
    hipLaunchKernelGGL((k_band_lines<P, TR>), dim3((ylines + 255) / 256), dim3(256), by, s, w0, w1, w2, w3,
                       static_cast<TR*>(nullptr), bs.mass[1], bs.stiff[1], ny, nx, sxy, static_cast<long>(nx), ylines,
                       dtk);
    // Z: r = Mz f - dtk Kz g   (lines (x, y))
    const int zlines = nx * ny;
    hipLaunchKernelGGL((k_band_lines<P, TR>), dim3((zlines + 255) / 256), dim3(256), bz, s, w2, w3,
                       static_cast<double*>(nullptr), static_cast<double*>(nullptr), r, bs.mass[2], bs.stiff[2], nz,
                       zlines, 0L, sxy, zlines, dtk);
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
template hipError_t launch_kron_residual<float, double>(int, const double*, const float*, double*, const BandSet&,
                                                        double, double*, double*, double*, hipStream_t);

hipError_t launch_dot(const double* a, const double* b, int64_t n, double* part, int part_cap, double* out,
                      hipStream_t s)
{
    int blocks = static_cast<int>((n + 255) / 256);
    if (blocks > 2048) blocks = 2048;
    if (blocks > part_cap) blocks = part_cap;
    if (blocks < 1) blocks = 1;
    hipLaunchKernelGGL(k_dot_partials, dim3(blocks), dim3(256), 0, s, a, b, static_cast<long>(n), part);
    hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(256), 0, s, part, blocks, out);
    return hipGetLastError();
}

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
