// part.hip -- the partitioned banded solve (part.hpp) on gfx950.
//
// A workgroup owns 64 lines (one lane each) and all K partitions of them
// (one wave each), so every global access of a wave covers 64 consecutive
// lines: coalesced along x for the y and z lines, and for the x rows through
// an LDS tile per wave.  A lane keeps its partition in registers from the
// load to the store, so each value is read and written once (16 B per DOF in
// FP64).  The two edge-state scans run through LDS between two barriers.
#include "part.hpp"

namespace ads {

namespace {

constexpr int kW = 64;
constexpr int kTile = 16;  // columns per LDS tile of the row kernel

// the partition (wave) index, wave-uniform, so table addresses are scalar
__device__ __forceinline__ int wave_index() { return __builtin_amdgcn_readfirstlane(threadIdx.x / kW); }

// LDS ordering between the lanes of one wave
__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <typename T>
__device__ __forceinline__ T fmadd(T a, T b, T c)
{
    return __builtin_fma(a, b, c);
}
template <>
__device__ __forceinline__ float fmadd<float>(float a, float b, float c)
{
    return __builtin_fmaf(a, b, c);
}

// The solve of one lane's partition k, rows [a, a + len), values in v.  All
// waves of the workgroup call it together (two barriers).  E and X: LDS of
// K * KL * 64 values each.
template <typename T, int KL, int SM>
__device__ __forceinline__ void part_core(T (&v)[SM], int k, int a, int len, int lane, const PartDev<T>& pd,
                                          T* __restrict__ E, T* __restrict__ X)
{
    // forward recurrence from zero state
    T h[KL];  // h[t] = y(j - 1 - t)
#pragma unroll
    for (int t = 0; t < KL; ++t) h[t] = T(0);
#pragma unroll
    for (int i = 0; i < SM; ++i) {
        if (i < len) {
            const T* __restrict__ l = pd.lrow + static_cast<long>(a + i) * KL;
            T y = v[i];
#pragma unroll
            for (int t = KL - 1; t >= 0; --t) y = fmadd(-l[t], h[t], y);
#pragma unroll
            for (int t = KL - 1; t > 0; --t) h[t] = h[t - 1];
            h[0] = y;
            v[i] = y;
        }
    }
#pragma unroll
    for (int r = 0; r < KL; ++r) E[(k * KL + r) * kW + lane] = h[r];
    __syncthreads();
    // the true state below the partition: s[i] = y(a - 1 - i)
    T s[KL];
#pragma unroll
    for (int t = 0; t < KL; ++t) s[t] = T(0);
    for (int kk = 0; kk < k; ++kk) {
        const int b = (kk + 1) * pd.S;  // partitions below the last are full
        T sn[KL];
#pragma unroll
        for (int r = 0; r < KL; ++r) {
            const T* __restrict__ g = pd.gsp + static_cast<long>(b - 1 - r) * KL;
            T acc = E[(kk * KL + r) * kW + lane];
#pragma unroll
            for (int i = 0; i < KL; ++i) acc = fmadd(g[i], s[i], acc);
            sn[r] = acc;
        }
#pragma unroll
        for (int r = 0; r < KL; ++r) s[r] = sn[r];
    }
    // y = y_loc + sum_i g_i s_i, then the backward recurrence from zero state
    T g2[KL];  // g2[t] = x(j + 1 + t)
#pragma unroll
    for (int t = 0; t < KL; ++t) g2[t] = T(0);
#pragma unroll
    for (int i = 0; i < SM; ++i) {
        if (i < len) {
            const T* __restrict__ g = pd.gsp + static_cast<long>(a + i) * KL;
            T y = v[i];
#pragma unroll
            for (int t = 0; t < KL; ++t) y = fmadd(g[t], s[t], y);
            v[i] = y;
        }
    }
#pragma unroll
    for (int i = SM - 1; i >= 0; --i) {
        if (i < len) {
            const T* __restrict__ u = pd.urow + static_cast<long>(a + i) * (KL + 1);
            T x = v[i];
#pragma unroll
            for (int t = KL; t >= 1; --t) x = fmadd(-u[t], g2[t - 1], x);
            x = x * u[0];
#pragma unroll
            for (int t = KL - 1; t > 0; --t) g2[t] = g2[t - 1];
            g2[0] = x;
            v[i] = x;
        }
    }
#pragma unroll
    for (int r = 0; r < KL; ++r) X[(k * KL + r) * kW + lane] = g2[r];
    __syncthreads();
    // the true state above the partition: tt[i] = x(b + i)
    T tt[KL];
#pragma unroll
    for (int t = 0; t < KL; ++t) tt[t] = T(0);
    for (int kk = pd.K - 1; kk > k; --kk) {
        const int a2 = kk * pd.S;
        T tn[KL];
#pragma unroll
        for (int r = 0; r < KL; ++r) {
            const T* __restrict__ hh = pd.hsp + static_cast<long>(a2 + r) * KL;
            T acc = X[(kk * KL + r) * kW + lane];
#pragma unroll
            for (int i = 0; i < KL; ++i) acc = fmadd(hh[i], tt[i], acc);
            tn[r] = acc;
        }
#pragma unroll
        for (int r = 0; r < KL; ++r) tt[r] = tn[r];
    }
#pragma unroll
    for (int i = 0; i < SM; ++i) {
        if (i < len) {
            const T* __restrict__ hh = pd.hsp + static_cast<long>(a + i) * KL;
            T x = v[i];
#pragma unroll
            for (int t = 0; t < KL; ++t) x = fmadd(hh[t], tt[t], x);
            v[i] = x;
        }
    }
}

// Lines along a strided axis (y, z): line L = (i, j), i in [0, ni) fastest,
// element m at F[i + j*stride_j + m*stride_k].
template <typename T, int KL, int SM>
__global__ __launch_bounds__(kW* kPartMaxK) void k_part_lines(T* __restrict__ F, int ni, long stride_j,
                                                              long stride_k, int nlines, PartDev<T> pd)
{
    extern __shared__ unsigned char smem[];
    T* E = reinterpret_cast<T*>(smem);
    T* X = E + pd.K * KL * kW;
    const int lane = threadIdx.x % kW, k = wave_index();
    const int L = blockIdx.x * kW + lane;
    const bool active = L < nlines;
    const long base = active ? static_cast<long>(L % ni) + static_cast<long>(L / ni) * stride_j : 0;
    const int a = k * pd.S, len = min(pd.S, pd.n - a);
    T v[SM];
#pragma unroll
    for (int i = 0; i < SM; ++i) v[i] = (active && i < len) ? F[base + static_cast<long>(a + i) * stride_k] : T(0);
    part_core<T, KL, SM>(v, k, a, len, lane, pd, E, X);
    if (active) {
#pragma unroll
        for (int i = 0; i < SM; ++i)
            if (i < len) F[base + static_cast<long>(a + i) * stride_k] = v[i];
    }
}

// Contiguous rows (x): row r = blockIdx.x*64 + lane, element m at F[r*n + m].
// A wave moves its partition of the 64 rows through an LDS tile of kTile
// columns (coalesced row pieces in, one row per lane out), and back.  The
// tiles share the LDS with the scans: barriers keep them apart.
template <typename T, int KL, int SM>
__global__ __launch_bounds__(kW* kPartMaxK) void k_part_rows(T* __restrict__ F, int nrows, PartDev<T> pd)
{
    static_assert(SM % kTile == 0, "partitions of whole tiles");
    constexpr int LDT = kTile + 1, RPI = kW / kTile, NIT = kW / RPI;
    extern __shared__ unsigned char smem[];
    T* E = reinterpret_cast<T*>(smem);
    T* X = E + pd.K * KL * kW;
    const int lane = threadIdx.x % kW, k = wave_index();
    T* tile = reinterpret_cast<T*>(smem) + static_cast<long>(k) * kW * LDT;
    const int row0 = blockIdx.x * kW;
    const int n = pd.n, a = k * pd.S, len = min(pd.S, n - a);
    const int cidx = lane % kTile, rsub = lane / kTile;
    T v[SM];
#pragma unroll
    for (int c0 = 0; c0 < SM; c0 += kTile) {
#pragma unroll
        for (int it = 0; it < NIT; ++it) {
            const int r = row0 + it * RPI + rsub, m = c0 + cidx;
            tile[(it * RPI + rsub) * LDT + cidx] =
                (r < nrows && m < len) ? F[static_cast<long>(r) * n + a + m] : T(0);
        }
        wave_sync();
#pragma unroll
        for (int i = 0; i < kTile; ++i) v[c0 + i] = tile[lane * LDT + i];
        wave_sync();
    }
    __syncthreads();  // the tiles overlap the scan buffers
    part_core<T, KL, SM>(v, k, a, len, lane, pd, E, X);
    __syncthreads();
#pragma unroll
    for (int c0 = 0; c0 < SM; c0 += kTile) {
#pragma unroll
        for (int i = 0; i < kTile; ++i) tile[lane * LDT + i] = v[c0 + i];
        wave_sync();
#pragma unroll
        for (int it = 0; it < NIT; ++it) {
            const int r = row0 + it * RPI + rsub, m = c0 + cidx;
            if (r < nrows && m < len) F[static_cast<long>(r) * n + a + m] = tile[(it * RPI + rsub) * LDT + cidx];
        }
        wave_sync();
    }
}

template <typename T, int KL>
size_t lines_lds(const PartDev<T>& pd)
{
    return static_cast<size_t>(2) * pd.K * KL * kW * sizeof(T);
}
template <typename T, int KL>
size_t rows_lds(const PartDev<T>& pd)
{
    const size_t tiles = static_cast<size_t>(pd.K) * kW * (kTile + 1) * sizeof(T);
    return tiles > lines_lds<T, KL>(pd) ? tiles : lines_lds<T, KL>(pd);
}

template <typename T, int KL>
hipError_t part_p(T* F, const PartDev<T> pd[3], hipStream_t s)
{
    for (int d = 0; d < 3; ++d)
        if (pd[d].K < 1 || pd[d].K > kPartMaxK || pd[d].S > kPartSeg || (pd[d].K - 1) * pd[d].S >= pd[d].n)
            return hipErrorInvalidValue;
    const int nx = pd[0].n, ny = pd[1].n, nz = pd[2].n;
    const long sxy = static_cast<long>(nx) * ny;
    const int rows = ny * nz, ylines = nx * nz, zlines = nx * ny;
    hipLaunchKernelGGL((k_part_rows<T, KL, kPartSeg>), dim3((rows + kW - 1) / kW), dim3(kW * pd[0].K),
                       (rows_lds<T, KL>(pd[0])), s, F, rows, pd[0]);
    hipLaunchKernelGGL((k_part_lines<T, KL, kPartSeg>), dim3((ylines + kW - 1) / kW), dim3(kW * pd[1].K),
                       (lines_lds<T, KL>(pd[1])), s, F, nx, sxy, static_cast<long>(nx), ylines, pd[1]);
    hipLaunchKernelGGL((k_part_lines<T, KL, kPartSeg>), dim3((zlines + kW - 1) / kW), dim3(kW * pd[2].K),
                       (lines_lds<T, KL>(pd[2])), s, F, zlines, 0L, sxy, zlines, pd[2]);
    return hipGetLastError();
}

}  // namespace

template <typename T>
hipError_t launch_part_solve(int p, T* F, const PartDev<T> pd[3], hipStream_t s)
{
    switch (p) {
    case 1: return part_p<T, 1>(F, pd, s);
    case 2: return part_p<T, 2>(F, pd, s);
    case 3: return part_p<T, 3>(F, pd, s);
    case 4: return part_p<T, 4>(F, pd, s);
    default: return hipErrorInvalidValue;
    }
}

template hipError_t launch_part_solve<double>(int, double*, const PartDev<double>[3], hipStream_t);

}  // namespace ads
