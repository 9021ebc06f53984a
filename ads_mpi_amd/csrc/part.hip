// part.hip -- the partitioned banded solve (part.hpp) on gfx950.
//
// A workgroup solves groups of 64 lines (one lane each) with one wave per
// partition, so every global access of a wave covers 64 consecutive lines:
// coalesced along x for the y and z lines, and for the x rows through an LDS
// tile per wave.  A lane keeps its partition in registers from the load to
// the store, so each value is read and written once (16 B per DOF in FP64).
// The two edge-state scans run through LDS between two barriers.  The
// workgroups are persistent and load the next group while they solve the
// current one; a wave keeps its partition's factor rows spread over its
// lanes for the whole launch.
#include <type_traits>

#include "part.hpp"

namespace ads {

namespace {

constexpr int kW = 64;

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

// The solve of one lane's partition k, rows [a, a + SM), values in v (past
// the line the tables are zero, so the backward recurrence leaves exact zeros
// there for any finite values).  All waves of the workgroup call it together (two barriers).  E and
// X: LDS of K * KL * 64 values each.
// A uniform value (index, stride or global pointer) made opaque to the
// optimiser, in scalar registers: the loads and stores of the persistent
// loops then take a scalar base per element and one shared VGPR offset, and
// their addresses are not hoisted out of the loop (32 of them would spill).
template <typename I>
__device__ __forceinline__ I opaque_s(I x)
{
    asm volatile("" : "+s"(x));
    return x;
}
template <typename T>
using gptr = __attribute__((address_space(1))) T*;
// The tables are read through the constant address space: they are never
// written while a kernel runs, so their wave-uniform loads stay scalar loads
// also after the loop's stores (which a global pointer could alias).
template <typename T>
using cptr = const __attribute__((address_space(4))) T*;
template <typename T>
__device__ __forceinline__ cptr<T> as_const(const T* p)
{
    return (cptr<T>)p;
}
template <typename T>
__device__ __forceinline__ gptr<T> opaque_g(T* p)
{
    gptr<T> q = (gptr<T>)p;
    asm volatile("" : "+s"(q));
    return q;
}

// The scans' tables, copied to LDS once per workgroup (2 K KL^2 values):
//   Gs[(kk*KL + r)*KL + i] = g_i(b_kk - 1 - r)   (the last rows of partition kk)
//   Hs[(kk*KL + r)*KL + i] = h_i(a_kk + r)       (its first rows)
// Read from global memory they would be scalar loads, one L2 round trip per
// step of the scans.  The first barrier of part_core orders the copy before
// the first read.
template <typename T, int KL, int SM>
__device__ __forceinline__ void stage_edges(const PartDev<T>& pd, T* __restrict__ Gs, T* __restrict__ Hs)
{
    const int n = pd.K * KL * KL;
    for (int q = threadIdx.x; q < n; q += blockDim.x) {
        const int kk = q / (KL * KL), r = (q / KL) % KL, i = q % KL;
        Gs[q] = pd.gsp[static_cast<long>((kk + 1) * SM - 1 - r) * KL + i];
        Hs[q] = pd.hsp[static_cast<long>(kk * SM + r) * KL + i];
    }
}

template <typename T, int KL>
constexpr size_t edge_lds(int K)
{
    return static_cast<size_t>(2) * K * KL * KL * sizeof(T);
}

// A block of a partition's factor rows held across the lanes of its wave:
// value q sits in lane q % 64 of slot q / 64, and a step takes it with two
// v_readlane into a scalar operand.  Scalar loads of the tables put an L2
// round trip on every step of the recurrences (the tables of a 259-entry
// line are 30 KB, beyond the scalar cache), and the persistent kernels load
// these once.
template <typename T, int N>
struct LaneBlock {
    static_assert(sizeof(T) == 8, "two 32-bit halves");
    static constexpr int kSlots = (N + kW - 1) / kW;
    T v[kSlots];
    __device__ __forceinline__ void load(cptr<T> src, int lane)
    {
#pragma unroll
        for (int s = 0; s < kSlots; ++s) {
            const int q = s * kW + lane;
            v[s] = q < N ? src[q] : T(0);
        }
    }
    // The slots made to depend on dep (a value of the recurrence): the
    // readlanes after this point wait for dep, so the scheduler cannot hoist
    // them (out of the persistent loop, or far ahead of their steps, where
    // hundreds of scalar values would spill).
    __device__ __forceinline__ void tie(const T& dep)
    {
#pragma unroll
        for (int s = 0; s < kSlots; ++s) asm volatile("" : "+v"(v[s]) : "v"(dep));
    }
    __device__ __forceinline__ T operator[](int q) const
    {
        const long bits = __builtin_bit_cast(long, v[q / kW]);
        const int lo = __builtin_amdgcn_readlane(static_cast<int>(bits), q % kW);
        const int hi = __builtin_amdgcn_readlane(static_cast<int>(bits >> 32), q % kW);
        return __builtin_bit_cast(T, (static_cast<long>(hi) << 32) | static_cast<unsigned>(lo));
    }
};

// Steps between two ties of the table registers to the recurrence (the
// readlanes of this many steps may be scheduled ahead).
constexpr int kTieSteps = 4;

// Partition k's rows [a, a + SM) of the factor tables (part.hpp).
template <typename T, int KL, int SM>
struct PartTabs {
    LaneBlock<T, SM * KL> l;
    LaneBlock<T, SM*(KL + 1)> u;
    __device__ __forceinline__ void load(const PartDev<T>& pd, int a, int lane)
    {
        l.load(as_const(pd.lrow) + static_cast<long>(a) * KL, lane);
        u.load(as_const(pd.urow) + static_cast<long>(a) * (KL + 1), lane);
    }

};

template <typename T, int KL, int SM>
__device__ __forceinline__ void part_core(T (&v)[SM], int k, int a, int lane, const PartDev<T>& pd,
                                          PartTabs<T, KL, SM> tb, T* __restrict__ E, T* __restrict__ X,
                                          const T* __restrict__ Gs, const T* __restrict__ Hs)
{
    // forward recurrence from zero state
    T h[KL];  // h[t] = y(j - 1 - t)
#pragma unroll
    for (int t = 0; t < KL; ++t) h[t] = T(0);
#pragma unroll
    for (int i = 0; i < SM; ++i) {
        if (i % kTieSteps == 0) tb.l.tie(h[0]);
        T y = v[i];
#pragma unroll
        for (int t = KL - 1; t >= 0; --t) y = fmadd(-tb.l[i * KL + t], h[t], y);
#pragma unroll
        for (int t = KL - 1; t > 0; --t) h[t] = h[t - 1];
        h[0] = y;
        v[i] = y;
    }
#pragma unroll
    for (int r = 0; r < KL; ++r) E[(k * KL + r) * kW + lane] = h[r];
    __syncthreads();
    // the true state below the partition: s[i] = y(a - 1 - i)
    T s[KL];
#pragma unroll
    for (int t = 0; t < KL; ++t) s[t] = T(0);
    for (int kk = 0; kk < k; ++kk) {
        T sn[KL];
#pragma unroll
        for (int r = 0; r < KL; ++r) {
            const T* g = Gs + (kk * KL + r) * KL;
            T acc = E[(kk * KL + r) * kW + lane];
#pragma unroll
            for (int i = 0; i < KL; ++i) acc = fmadd(g[i], s[i], acc);
            sn[r] = acc;
        }
#pragma unroll
        for (int r = 0; r < KL; ++r) s[r] = sn[r];
    }
    // y = y_loc + d: d is the homogeneous forward recurrence started from
    // the true state s (by linearity the response to it, sum_i g_i s_i,
    // computed with the factor rows already held instead of the g tables)
    {
        T d[KL];  // d[t] = d(j - 1 - t)
#pragma unroll
        for (int t = 0; t < KL; ++t) d[t] = s[t];
#pragma unroll
        for (int i = 0; i < SM; ++i) {
            if (i % kTieSteps == 0) tb.l.tie(d[0]);
            T e = -tb.l[i * KL + KL - 1] * d[KL - 1];
#pragma unroll
            for (int t = KL - 2; t >= 0; --t) e = fmadd(-tb.l[i * KL + t], d[t], e);
#pragma unroll
            for (int t = KL - 1; t > 0; --t) d[t] = d[t - 1];
            d[0] = e;
            v[i] += e;
        }
    }
    // the backward recurrence from zero state
    T g2[KL];  // g2[t] = x(j + 1 + t)
#pragma unroll
    for (int t = 0; t < KL; ++t) g2[t] = T(0);
#pragma unroll
    for (int i = SM - 1; i >= 0; --i) {
        if ((SM - 1 - i) % kTieSteps == 0) tb.u.tie(g2[0]);
        T x = v[i];
#pragma unroll
        for (int t = KL; t >= 1; --t) x = fmadd(-tb.u[i * (KL + 1) + t], g2[t - 1], x);
        x = x * tb.u[i * (KL + 1)];
#pragma unroll
        for (int t = KL - 1; t > 0; --t) g2[t] = g2[t - 1];
        g2[0] = x;
        v[i] = x;
    }
#pragma unroll
    for (int r = 0; r < KL; ++r) X[(k * KL + r) * kW + lane] = g2[r];
    __syncthreads();
    // the true state above the partition: tt[i] = x(b + i)
    T tt[KL];
#pragma unroll
    for (int t = 0; t < KL; ++t) tt[t] = T(0);
    for (int kk = pd.K - 1; kk > k; --kk) {
        T tn[KL];
#pragma unroll
        for (int r = 0; r < KL; ++r) {
            const T* hh = Hs + (kk * KL + r) * KL;
            T acc = X[(kk * KL + r) * kW + lane];
#pragma unroll
            for (int i = 0; i < KL; ++i) acc = fmadd(hh[i], tt[i], acc);
            tn[r] = acc;
        }
#pragma unroll
        for (int r = 0; r < KL; ++r) tt[r] = tn[r];
    }
    // x = x_loc + e: e is the homogeneous backward recurrence started from
    // the true state tt above the partition
    T e[KL];  // e[t] = e(j + 1 + t)
#pragma unroll
    for (int t = 0; t < KL; ++t) e[t] = tt[t];
#pragma unroll
    for (int i = SM - 1; i >= 0; --i) {
        if ((SM - 1 - i) % kTieSteps == 0) tb.u.tie(e[0]);
        T w = -tb.u[i * (KL + 1) + KL] * e[KL - 1];
#pragma unroll
        for (int t = KL - 1; t >= 1; --t) w = fmadd(-tb.u[i * (KL + 1) + t], e[t - 1], w);
        w = w * tb.u[i * (KL + 1)];
#pragma unroll
        for (int t = KL - 1; t > 0; --t) e[t] = e[t - 1];
        e[0] = w;
        v[i] += w;
    }
}

// Partitions per workgroup up to which the kernels double-buffer: the next
// group of 64 lines is loaded into registers while the current one is
// solved and stored (about 170 VGPRs: three waves per SIMD).  Longer lines
// run one group at a time.
constexpr int kPipeMaxK = 12;

// Both kernels are persistent: workgroup b solves groups b, b + grid, ...
// of 64 lines (one wave per partition), so the loads of one group overlap
// the recurrences and stores of the one before, and the grid is one round of
// resident workgroups (no tail of half-filled rounds).

// Lines along a strided axis (y, z): line L = (i, j), i in [0, ni) fastest,
// element m at F[i + j*stride_j + m*stride_k].
template <typename T, int KL, int SM, int KMAX>
__global__ __launch_bounds__(kW* KMAX) void k_part_lines(T* __restrict__ F, int ni, long stride_j, long stride_k,
                                                         int nlines, PartDev<T> pd)
{
    constexpr bool PIPE = KMAX <= kPipeMaxK;
    extern __shared__ unsigned char smem[];
    T* E = reinterpret_cast<T*>(smem);
    T* X = E + pd.K * KL * kW;
    T* Gs = X + pd.K * KL * kW;
    T* Hs = Gs + pd.K * KL * KL;
    const int lane = threadIdx.x % kW, k = wave_index();
    const int a = k * SM, n = pd.n;
    const int ngroups = (nlines + kW - 1) / kW;
    int g = blockIdx.x;
    if (g >= ngroups) return;  // the whole workgroup
    PartTabs<T, KL, SM> tb;
    tb.load(pd, a, lane);
    stage_edges<T, KL, SM>(pd, Gs, Hs);
    // lanes past the last line read the last line and store nothing
    // A group's lines start at a uniform address (its first line) plus a
    // lane offset below 4 GiB (the host checks it), so every load and store
    // takes a scalar base and one shared 32-bit VGPR offset.  Lanes past the
    // last line read the last line and store nothing.
    auto first_of = [&](int grp) {
        const int L = grp * kW;
        return static_cast<long>(L % ni) + static_cast<long>(L / ni) * stride_j;
    };
    auto lane_off = [&](int grp) {
        const int L = min(grp * kW + lane, nlines - 1);
        const long b = static_cast<long>(L % ni) + static_cast<long>(L / ni) * stride_j;
        return static_cast<unsigned>((b - first_of(grp)) * static_cast<long>(sizeof(T)));
    };
    // Every load issued, its row clamped inside the line: rows past the line
    // see its last value, which the zero table rows there turn into exact
    // zeros (x = (y - U x) / U(j, j) with 1/U(j, j) = 0).  The element
    // addresses are recomputed per group from opaque copies of the stride
    // and the clamp: hoisted out of the loop they would take 64 scalar
    // registers and spill.
    auto load = [&](int grp, T(&v)[SM]) {
        const char* gb = reinterpret_cast<const char*>(F + first_of(grp));
        const unsigned off = lane_off(grp);
        const long sk = opaque_s(stride_k * static_cast<long>(sizeof(T)));
        const int last = opaque_s(n - 1 - a);
#pragma unroll
        for (int i = 0; i < SM; ++i)
            v[i] = *reinterpret_cast<gptr<const T>>(
                reinterpret_cast<gptr<const char>>(opaque_g(gb + static_cast<long>(a + min(i, last)) * sk)) + off);
    };
    auto store = [&](int grp, const T(&v)[SM]) {
        if (grp * kW + lane >= nlines) return;
        char* gb = reinterpret_cast<char*>(F + first_of(grp));
        const unsigned off = lane_off(grp);
        const long sk = opaque_s(stride_k * static_cast<long>(sizeof(T)));
        const int last = opaque_s(n - 1 - a);
#pragma unroll
        for (int i = 0; i < SM; ++i)
            if (i <= last)
                *reinterpret_cast<gptr<T>>(reinterpret_cast<gptr<char>>(opaque_g(gb + static_cast<long>(a + i) * sk)) +
                                           off) = v[i];
    };
    T v[SM];
    load(g, v);
    for (;;) {
        const int gn = g + gridDim.x;  // the same in every wave of the group
        T nv[SM];
        if (PIPE && gn < ngroups) load(gn, nv);
        // the loads stay ahead of the recurrence (the scheduler would
        // otherwise interleave them with it, a few at a time)
        __builtin_amdgcn_sched_barrier(0);
        part_core<T, KL, SM>(v, k, a, lane, pd, tb, E, X, Gs, Hs);
        store(g, v);
        if (gn >= ngroups) break;
        g = gn;
        if constexpr (PIPE) {
#pragma unroll
            for (int i = 0; i < SM; ++i) v[i] = nv[i];
        } else {
            load(g, v);
        }
    }
}

// Contiguous rows (x): row r = group*64 + lane, element m at F[r*n + m].  A
// wave moves its partition of the 64 rows through an LDS tile of TILE
// columns (coalesced row pieces in, one row per lane out), and back; the
// next group's row pieces wait in registers (untransposed) meanwhile.
template <typename T, int KL, int SM, int KMAX, int TILE>
__global__ __launch_bounds__(kW* KMAX) void k_part_rows(T* __restrict__ F, int nrows, PartDev<T> pd)
{
    static_assert(SM % TILE == 0 && kW % TILE == 0, "partitions of whole tiles");
    constexpr bool PIPE = KMAX <= kPipeMaxK;
    constexpr int LDT = TILE + 1, RPI = kW / TILE, NIT = kW / RPI, NT = SM / TILE;
    extern __shared__ unsigned char smem[];
    T* E = reinterpret_cast<T*>(smem);
    T* X = E + pd.K * KL * kW;
    T* Gs = X + pd.K * KL * kW;
    T* Hs = Gs + pd.K * KL * KL;
    const int lane = threadIdx.x % kW, k = wave_index();
    T* tile = Hs + pd.K * KL * KL + static_cast<long>(k) * kW * LDT;  // private to the wave
    const int n = pd.n, a = k * SM;
    const int cidx = lane % TILE, rsub = lane / TILE;
    const int ngroups = (nrows + kW - 1) / kW;
    int g = blockIdx.x;
    if (g >= ngroups) return;  // the whole workgroup
    PartTabs<T, KL, SM> tb;
    tb.load(pd, a, lane);
    stage_edges<T, KL, SM>(pd, Gs, Hs);
    // raw[c*NIT + it]: row row0 + it*RPI + rsub, column a + c*TILE + cidx,
    // clamped inside the array (as in k_part_lines); a uniform piece base
    // and a small lane offset, so the loads take scalar bases
    auto load = [&](int grp, T(&raw)[SM]) {
        const int row0 = grp * kW;
        const int lastr = opaque_s(nrows - 1 - row0), lastc = opaque_s(n - 1 - a);
#pragma unroll
        for (int c = 0; c < NT; ++c)
#pragma unroll
            for (int it = 0; it < NIT; ++it) {
                const int rb = min(it * RPI, lastr), cb = min(c * TILE, lastc);
                gptr<const T> e = opaque_g(static_cast<const T*>(F) + static_cast<long>(row0 + rb) * n + a + cb);
                const int lo = min(rsub, lastr - rb) * n + min(cidx, lastc - cb);
                raw[c * NIT + it] = *reinterpret_cast<gptr<const T>>(
                    reinterpret_cast<gptr<const char>>(e) + static_cast<unsigned>(lo * static_cast<int>(sizeof(T))));
            }
    };
    auto transpose_in = [&](const T(&raw)[SM], T(&v)[SM]) {
#pragma unroll
        for (int c = 0; c < NT; ++c) {
#pragma unroll
            for (int it = 0; it < NIT; ++it) tile[(it * RPI + rsub) * LDT + cidx] = raw[c * NIT + it];
            wave_sync();
#pragma unroll
            for (int i = 0; i < TILE; ++i) v[c * TILE + i] = tile[lane * LDT + i];
            wave_sync();
        }
    };
    auto store = [&](int grp, const T(&v)[SM]) {
        const int row0 = grp * kW;
#pragma unroll
        for (int c = 0; c < NT; ++c) {
#pragma unroll
            for (int i = 0; i < TILE; ++i) tile[lane * LDT + i] = v[c * TILE + i];
            wave_sync();
#pragma unroll
            for (int it = 0; it < NIT; ++it) {
                const int r = row0 + it * RPI + rsub, m = a + c * TILE + cidx;
                gptr<char> e = reinterpret_cast<gptr<char>>(opaque_g(F + static_cast<long>(row0 + it * RPI) * n + a + c * TILE));
                if (r < nrows && m < n)
                    *reinterpret_cast<gptr<T>>(e + static_cast<unsigned>((rsub * n + cidx) * static_cast<int>(sizeof(T)))) =
                        tile[(it * RPI + rsub) * LDT + cidx];
            }
            wave_sync();
        }
    };
    T raw[SM], v[SM];
    load(g, raw);
    for (;;) {
        const int gn = g + gridDim.x;  // the same in every wave of the group
        transpose_in(raw, v);
        if (PIPE && gn < ngroups) load(gn, raw);
        __builtin_amdgcn_sched_barrier(0);
        part_core<T, KL, SM>(v, k, a, lane, pd, tb, E, X, Gs, Hs);
        store(g, v);
        if (gn >= ngroups) break;
        g = gn;
        if constexpr (!PIPE) load(g, raw);
    }
}

template <typename T, int KL>
size_t scan_lds(const PartDev<T>& pd)
{
    return static_cast<size_t>(2) * pd.K * KL * kW * sizeof(T) + edge_lds<T, KL>(pd.K);
}
template <int TILE, typename T, int KL>
size_t rows_lds(const PartDev<T>& pd)
{
    return scan_lds<T, KL>(pd) + static_cast<size_t>(pd.K) * kW * (TILE + 1) * sizeof(T);
}

// One round of resident workgroups (at most one per group).
template <class KERNEL>
hipError_t persistent_grid(KERNEL kernel, int block, size_t lds, int ngroups, int* grid)
{
    int dev = 0, cus = 0, occ = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (e == hipSuccess)
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, reinterpret_cast<const void*>(kernel), block, lds);
    if (e != hipSuccess) return e;
    if (occ < 1) return hipErrorInvalidConfiguration;  // does not fit a CU
    const long g = static_cast<long>(cus) * occ;
    *grid = static_cast<int>(g < ngroups ? g : ngroups);
    return hipSuccess;
}

// One direction's launch, the instance chosen by its partition length and
// count: double-buffered up to kPipeMaxK partitions of kPartSeg, one group at
// a time beyond (kPartMaxK partitions of kPartSeg or kPartSegLong).
template <typename T, int KL, int SM, int KMAX, int TILE>
hipError_t part_rows_launch(T* F, int rows, const PartDev<T>& pd, hipStream_t s)
{
    auto kr = k_part_rows<T, KL, SM, KMAX, TILE>;
    const size_t lds = rows_lds<TILE, T, KL>(pd);
    const int block = kW * pd.K, ng = (rows + kW - 1) / kW;
    int grid = 0;
    if (hipError_t e = persistent_grid(kr, block, lds, ng, &grid)) return e;
    hipLaunchKernelGGL(kr, dim3(grid), dim3(block), lds, s, F, rows, pd);
    return hipGetLastError();
}

template <typename T, int KL, int SM, int KMAX>
hipError_t part_lines_launch(T* F, int ni, long stride_j, long stride_k, int nlines, const PartDev<T>& pd,
                             hipStream_t s)
{
    auto kl = k_part_lines<T, KL, SM, KMAX>;
    const size_t lds = scan_lds<T, KL>(pd);
    int grid = 0;
    if (hipError_t e = persistent_grid(kl, kW * pd.K, lds, (nlines + kW - 1) / kW, &grid)) return e;
    hipLaunchKernelGGL(kl, dim3(grid), dim3(kW * pd.K), lds, s, F, ni, stride_j, stride_k, nlines, pd);
    return hipGetLastError();
}

// f(integral_constant SM, integral_constant KMAX) for a direction's tables
template <typename T, class Fn>
hipError_t part_instance(const PartDev<T>& pd, Fn&& f)
{
    if (pd.S == kPartSeg && pd.K <= kPipeMaxK)
        return f(std::integral_constant<int, kPartSeg>{}, std::integral_constant<int, kPipeMaxK>{});
    if (pd.S == kPartSeg) return f(std::integral_constant<int, kPartSeg>{}, std::integral_constant<int, kPartMaxK>{});
    return f(std::integral_constant<int, kPartSegLong>{}, std::integral_constant<int, kPartMaxK>{});
}

template <typename T, int KL>
hipError_t part_p(T* F, const PartDev<T> pd[3], hipStream_t s)
{
    if (!part_planes_fit(pd[0].n, pd[1].n, sizeof(T))) return hipErrorInvalidValue;
    for (int d = 0; d < 3; ++d) {
        if (pd[d].K < 1 || pd[d].K > kPartMaxK || pd[d].S != part_seg(pd[d].n) ||
            (pd[d].K - 1) * pd[d].S >= pd[d].n)
            return hipErrorInvalidValue;
    }
    const int nx = pd[0].n, ny = pd[1].n, nz = pd[2].n;
    const long sxy = static_cast<long>(nx) * ny;
    hipError_t e = part_instance(pd[0], [&](auto sm, auto kmax) {
        return part_rows_launch<T, KL, decltype(sm)::value, decltype(kmax)::value, 8>(F, ny * nz, pd[0], s);
    });
    if (e != hipSuccess) return e;
    e = part_instance(pd[1], [&](auto sm, auto kmax) {
        return part_lines_launch<T, KL, decltype(sm)::value, decltype(kmax)::value>(F, nx, sxy, static_cast<long>(nx),
                                                                                    nx * nz, pd[1], s);
    });
    if (e != hipSuccess) return e;
    return part_instance(pd[2], [&](auto sm, auto kmax) {
        return part_lines_launch<T, KL, decltype(sm)::value, decltype(kmax)::value>(F, nx * ny, 0L, sxy, nx * ny,
                                                                                    pd[2], s);
    });
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
    case 5: return part_p<T, 5>(F, pd, s);
    default: return hipErrorInvalidValue;
    }
}

template hipError_t launch_part_solve<double>(int, double*, const PartDev<double>[3], hipStream_t);

}  // namespace ads
