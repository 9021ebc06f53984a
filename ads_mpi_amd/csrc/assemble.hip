// assemble.hip -- gfx950 kernels of the RHS assembly (FormUn + Form3DRHS).
//
// Heat form (problems/heat/RHS_eq.F90:74-81, with a per-element diffusivity
// kappa).  The integrand J W (v u - dt kappa grad u . grad v) is bilinear in
// (u, v), and J = Jx Jy Jz, W = Wx Wy Wz and kappa are constant on an
// element, so the Gauss sum over the Q^3 points of an element (Form3DRHS,
// projection_engine.F90:154-284, over the point values of FormUn, :306-441)
// factorises per direction into the 1-D element matrices
//   Me(a, b) = sum_q N_a N_b J W_q,   Ke(a, b) = sum_q N'_a N'_b J W_q
// (built once on the host from the same basis values):
//   F_e  = (Mz(x)My(x)Mx - dt kappa (Mz(x)My(x)Kx + Mz(x)Ky(x)Mx + Kz(x)My(x)Mx)) c_e
//   l2_e = c_e^T (Mz(x)My(x)Mx) c_e          (= sum over the points of J W u^2)
// k_heat applies this operator one direction at a time: 8 P^4 multiply-adds
// per element, against 18 P^4 for values and gradients at the points followed
// by the transposed contractions.
//   * a workgroup of four waves owns 64 consecutive elements along x, four
//     consecutive y rows (one wave each) and a segment of z; one lane per
//     element;
//   * the coefficient planes z = ez..ez+p the current element layer touches
//     sit in an LDS ring shared by the four waves ((64+p) x (4+p) values per
//     plane), so a coefficient is read from HBM about (4+p)/4 times;
//   * contributions accumulate in an LDS ring of right-hand-side planes by LDS
//     atomic adds; a plane is complete once the element layer below it is
//     done and is then added to F with coalesced global atomics (neighbouring
//     workgroups in x, y and z share its edges).
//
// Projection form (t = 0, RHS_eq.F90:82-86): f0 is sampled at the Gauss
// points, so k_project keeps the point-wise quadrature, sum-factorised one
// direction at a time.
//
// The adjoint instance of k_heat (kHeatGrad) also accumulates the per-element
// diffusivity gradient; the MultiStep instance (kHeatDual) takes the mass
// part from one coefficient field and the stiffness part from another.
#include "kernels.hpp"

#include <hip/hip_runtime.h>

namespace ads {

namespace {

constexpr int kWave = 64;
constexpr int kRowsWg = 4;  // y rows (one wave each) per heat workgroup

template <typename T>
__device__ __forceinline__ void atomic_add(T* p, T v)
{
    unsafeAtomicAdd(p, v);
}

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
    return v;
}

// position of (a, b) in a packed upper triangle of a symmetric P x P matrix
template <int P>
__host__ __device__ constexpr int tri(int a, int b)
{
    return a <= b ? a * P - a * (a - 1) / 2 + (b - a) : b * P - b * (b - 1) / 2 + (a - b);
}

// initial states of the t = 0 projection
template <typename T>
__device__ __forceinline__ T f0_eval(int kind, T x, T y, T z)
{
    if (kind == 0) return T(1);
    if (kind == 1) {
        const double pi = 3.141592653589793;
        return T(sin(pi * double(x)) * sin(pi * double(y)) * sin(pi * double(z)));
    }
    // 2*bump3d(0.05, 0.4, x, y, z) = 2*falloff(0.025, 0.2, |X - (0.5,0.5,0.5)|)
    const double dx = double(x) - 0.5, dy = double(y) - 0.5, dz = double(z) - 0.5;
    const double t = sqrt(dx * dx + dy * dy + dz * dz);
    const double r = 0.05 / 2.0, R = 0.4 / 2.0;
    double fv;
    if (t < r) fv = 1.0;
    else if (t > R) fv = 0.0;
    else {
        const double h = (t - r) / (R - r);
        const double g = (h - 1.0) * (h + 1.0);
        fv = g * g;
    }
    return T(2.0 * fv);
}

struct AsmDims {
    int nex, ney, nez;  // elements per direction
    int sx, sy;         // DOFs along x and y
    int zseg;           // element layers per workgroup
    int f0;             // initial-state kind (projection form)
    int xchunks;        // workgroups along x
    int yblocks;        // workgroups along y
};

// MI355X hands consecutive workgroups to its 8 XCDs round robin.  Remapping
// the 1-D block index gives XCD k a contiguous run of (x chunk, y block, z
// segment) tiles, so workgroups that add into the same right-hand-side rows
// (neighbouring y blocks) share one L2.  The map is a bijection for any count.
constexpr int kXcds = 8;
__device__ __forceinline__ int xcd_tile(int b, int nb)
{
    const int q = nb / kXcds, r = nb % kXcds, k = b % kXcds, s = b / kXcds;
    return k * q + min(k, r) + s;
}

// Every table pointer is a separate __restrict__ argument so the compiler
// may read the wave-uniform y and z element matrices with scalar loads.
// MODE kHeatGrad (adjoint): c is mu = M^-T lam and, in the same pass,
// dkappa[e] += -dt cn_e^T K_e mu_e (K_e symmetric) from a third LDS ring
// holding the forward state cn.
// MODE kHeatDual (MultiStep, src/ADS.F90:391-429): F = M c - dt kappa K cn,
// the point values from c (Un) and the gradients from cn (dUn of the latest
// substep); cn sits in the third ring.
constexpr int kHeatPlain = 0, kHeatGrad = 1, kHeatDual = 2;

// A wave's x positions move up by one lane (lane i takes lane i-1's value,
// lane 0 takes 0): DPP wave_shr:1 on both halves of a double.
__device__ __forceinline__ double lane_up(double v)
{
    const long long b = __builtin_bit_cast(long long, v);
    const int lo = __builtin_amdgcn_update_dpp(0, static_cast<int>(b), 0x138, 0xf, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(0, static_cast<int>(b >> 32), 0x138, 0xf, 0xf, false);
    return __builtin_bit_cast(double, (static_cast<long long>(hi) << 32) | static_cast<unsigned>(lo));
}

// DPPR (plain mode only): the P x-offsets of an output are summed across
// lanes (DPP shifts into P^2 accumulators per lane) before the LDS add, so
// a wave issues P^2 full-width ds_add per element layer instead of P^3 (the
// lanes whose x position lies past the wave add theirs directly).
template <typename T, int P, int MODE, bool DPPR = false>
__global__ __launch_bounds__(kWave* kRowsWg) void k_heat(const double* __restrict__ c, T* __restrict__ F,
                                                         const T* __restrict__ emx, const T* __restrict__ emy,
                                                         const T* __restrict__ emz, const double* __restrict__ kappa,
                                                         T ndt, AsmDims dm, double* __restrict__ l2_part,
                                                         const double* __restrict__ cn, double* __restrict__ dkappa)
{
    constexpr int KL = P - 1, W = kWave + KL, R = kRowsWg + KL, NT = P * (P + 1) / 2;
    constexpr int NTH = kWave * kRowsWg;
    constexpr bool GRAD = MODE == kHeatGrad, DUAL = MODE == kHeatDual, TWO = GRAD || DUAL;
    // The right-hand-side ring accumulates in FP64 also for T = float: on
    // gfx950 ds_add_f32 issues about three times slower than ds_add_f64
    // (LDS-active cycles 302M against 108M for the same instruction count).
    using A = double;
    __shared__ T cs[P][R][W];  // coefficient planes, ring by z
    __shared__ A fs[P][R][W];  // right-hand-side planes, ring by z
    __shared__ T ns[TWO ? P : 1][TWO ? R : 1][TWO ? W : 1];  // second field (cn)
    __shared__ double l2w[kRowsWg];

    const int tid = threadIdx.x;
    const int lane = tid & (kWave - 1);
    const int wv = __builtin_amdgcn_readfirstlane(tid / kWave);
    const int tile = xcd_tile(blockIdx.x, gridDim.x);
    const int xc = tile % dm.xchunks;
    const int yb = (tile / dm.xchunks) % dm.yblocks;
    const int seg = tile / (dm.xchunks * dm.yblocks);
    const int ex0 = xc * kWave, ey0 = yb * kRowsWg;
    const int ex = ex0 + lane, ey = ey0 + wv;
    const int ez0 = seg * dm.zseg;
    const int ez1 = min(ez0 + dm.zseg, dm.nez);
    const bool row_ok = ey < dm.ney;
    const bool active = row_ok && ex < dm.nex;
    const int exc = min(ex, dm.nex - 1), eyc = min(ey, dm.ney - 1);
    const int sx = dm.sx;
    const long plane = static_cast<long>(sx) * dm.sy;
    const int ncol = min(W, sx - ex0);
    const int nrow = min(R, dm.sy - ey0);

    // first DOF of element e is e (open knots, checked on the host)
    const T* __restrict__ Ex = emx + static_cast<size_t>(exc) * P * 2 * P;
    const T* __restrict__ Ey = emy + static_cast<size_t>(eyc) * 2 * NT;

    for (int i = tid; i < P * R * W; i += NTH) (&fs[0][0][0])[i] = A(0);

    // plane z of c (and cn) is fetched into registers (NPT entries per
    // thread) and written to its ring slot later, so the loads of the next
    // plane overlap the current layer's work
    constexpr int NE = R * W, NPT = (NE + NTH - 1) / NTH;
    T pc[NPT], pn[TWO ? NPT : 1];
    auto fetch = [&](int z) {
        const long off = ex0 + static_cast<long>(sx) * ey0 + plane * z;
#pragma unroll
        for (int k = 0; k < NPT; ++k) {
            const int i = tid + k * NTH;
            const int r = i / W, col = i - r * W;
            const bool in = i < NE && col < ncol && r < nrow;
            pc[k] = in ? T(c[off + col + static_cast<long>(sx) * r]) : T(0);
            if constexpr (TWO) pn[k] = in ? T(cn[off + col + static_cast<long>(sx) * r]) : T(0);
        }
    };
    auto stash = [&](int z) {
        const int slot = z % P;
#pragma unroll
        for (int k = 0; k < NPT; ++k) {
            const int i = tid + k * NTH;
            if (i < NE) {
                const int r = i / W, col = i - r * W;
                cs[slot][r][col] = pc[k];
                if constexpr (TWO) ns[slot][r][col] = pn[k];
            }
        }
    };
    // each entry is always handled by the same thread, so consecutive
    // flushes need no barrier between them
    auto flush_plane = [&](int z) {
        const int slot = z % P;
        T* dst = F + ex0 + static_cast<long>(sx) * ey0 + plane * z;
        for (int i = tid; i < R * W; i += NTH) {
            const int r = i / W, col = i - r * W;
            if (col < ncol && r < nrow) atomic_add(&dst[col + static_cast<long>(sx) * r], T(fs[slot][r][col]));
            fs[slot][r][col] = A(0);
        }
    };

    for (int z = ez0; z < ez0 + KL; ++z) {
        fetch(z);
        stash(z);
    }
    fetch(ez0 + KL);

    T l2 = T(0);
    for (int ez = ez0; ez < ez1; ++ez) {
        // plane ez+KL goes to the slot of plane ez-1, whose readers finished
        // before the previous iteration's second barrier; then the loads of
        // plane ez+KL+1 are issued
        stash(ez + KL);
        if (ez + 1 < ez1) fetch(ez + KL + 1);
        __syncthreads();
        if (row_ok) {
            const T* __restrict__ Ez = emz + static_cast<size_t>(ez) * 2 * NT;
            const T kap =
                kappa ? T(kappa[exc + static_cast<size_t>(dm.nex) * (eyc + static_cast<size_t>(dm.ney) * ez)]) : T(1);
            const T s = ndt * kap;
            int slot[P];
#pragma unroll
            for (int az = 0; az < P; ++az) slot[az] = (ez + az) % P;
            T gacc = T(0);  // cn_e^T K_e c_e (GRAD)
            A racc[DPPR ? P : 1][DPPR ? P : 1];  // [azo][ayo] at x position lane (DPPR)
            if constexpr (DPPR) {
#pragma unroll
                for (int azo = 0; azo < P; ++azo)
#pragma unroll
                    for (int ayo = 0; ayo < P; ++ayo) racc[azo][ayo] = A(0);
            }

#pragma unroll 1
            for (int axo = 0; axo < P; ++axo) {
                // row axo of the lane's x element matrices; zero rows make an
                // inactive lane add exact zeros
                T mr[P], kr[P];
#pragma unroll
                for (int b = 0; b < P; ++b) {
                    mr[b] = active ? Ex[(axo * 2) * P + b] : T(0);
                    kr[b] = active ? Ex[(axo * 2 + 1) * P + b] : T(0);
                }
                // x: mx = (Mx c)[axo], kx = (Kx c)[axo]   [ay][az]; the
                // stiffness source is cn in DUAL, with mk = (Mx cn)[axo]
                T mx[P][P], kx[P][P], mk[DUAL ? P : 1][DUAL ? P : 1];
#pragma unroll
                for (int ay = 0; ay < P; ++ay)
#pragma unroll
                    for (int az = 0; az < P; ++az) {
                        const T* col = &cs[slot[az]][wv + ay][lane];
                        const T* colk = col;
                        if constexpr (DUAL) colk = &ns[slot[az]][wv + ay][lane];
                        T a = mr[0] * col[0];
                        T k = kr[0] * colk[0];
#pragma unroll
                        for (int b = 1; b < P; ++b) {
                            a = fma(mr[b], col[b], a);
                            k = fma(kr[b], colk[b], k);
                        }
                        mx[ay][az] = a;
                        kx[ay][az] = k;
                        if constexpr (DUAL) {
                            T q = mr[0] * colk[0];
#pragma unroll
                            for (int b = 1; b < P; ++b) q = fma(mr[b], colk[b], q);
                            mk[ay][az] = q;
                        }
                    }
#pragma unroll
                for (int azo = 0; azo < P; ++azo) {
                    // z: zm = Mz mx, zk = Mz kx + Kz mk, zq = Mz mk (DUAL; mk = mx otherwise)   [ay]
                    T zm[P], zk[P], zq[DUAL ? P : 1];
#pragma unroll
                    for (int ay = 0; ay < P; ++ay) {
                        const T* mkr = &mx[ay][0];
                        if constexpr (DUAL) mkr = &mk[ay][0];
                        T m = Ez[tri<P>(azo, 0)] * mx[ay][0];
                        T k = fma(Ez[NT + tri<P>(azo, 0)], mkr[0], Ez[tri<P>(azo, 0)] * kx[ay][0]);
#pragma unroll
                        for (int az = 1; az < P; ++az) {
                            m = fma(Ez[tri<P>(azo, az)], mx[ay][az], m);
                            k = fma(Ez[tri<P>(azo, az)], kx[ay][az], k);
                            k = fma(Ez[NT + tri<P>(azo, az)], mkr[az], k);
                        }
                        zm[ay] = m;
                        zk[ay] = k;
                        if constexpr (DUAL) {
                            T q = Ez[tri<P>(azo, 0)] * mkr[0];
#pragma unroll
                            for (int az = 1; az < P; ++az) q = fma(Ez[tri<P>(azo, az)], mkr[az], q);
                            zq[ay] = q;
                        }
                    }
                    // y: M c = My zm, K c = My zk + Ky zq   (output row ayo)
#pragma unroll
                    for (int ayo = 0; ayo < P; ++ayo) {
                        const T* zqr = &zm[0];
                        if constexpr (DUAL) zqr = &zq[0];
                        T om = Ey[tri<P>(ayo, 0)] * zm[0];
                        T ok = fma(Ey[NT + tri<P>(ayo, 0)], zqr[0], Ey[tri<P>(ayo, 0)] * zk[0]);
#pragma unroll
                        for (int ay = 1; ay < P; ++ay) {
                            om = fma(Ey[tri<P>(ayo, ay)], zm[ay], om);
                            ok = fma(Ey[tri<P>(ayo, ay)], zk[ay], ok);
                            ok = fma(Ey[NT + tri<P>(ayo, ay)], zqr[ay], ok);
                        }
                        l2 = fma(cs[slot[azo]][wv + ayo][lane + axo], om, l2);
                        if constexpr (GRAD) gacc = fma(ns[slot[azo]][wv + ayo][lane + axo], ok, gacc);
                        const A val = A(fma(s, ok, om));
                        if constexpr (DPPR) {
                            if (lane + axo >= kWave) atomic_add(&fs[slot[azo]][wv + ayo][lane + axo], val);
                            A sh = val;
#pragma unroll
                            for (int t = 1; t < P; ++t)
                                if (axo >= t) sh = lane_up(sh);
                            racc[azo][ayo] += sh;
                        } else {
                            // lanes of one wave address distinct columns here; the
                            // LDS applies one wave's adds in program order
                            atomic_add(&fs[slot[azo]][wv + ayo][lane + axo], val);
                        }
                    }
                    // keep the scheduler from computing every plane's values up front
                    __builtin_amdgcn_sched_barrier(0);
                }
            }
            if constexpr (DPPR) {
#pragma unroll
                for (int azo = 0; azo < P; ++azo)
#pragma unroll
                    for (int ayo = 0; ayo < P; ++ayo) atomic_add(&fs[slot[azo]][wv + ayo][lane], racc[azo][ayo]);
            }
            if constexpr (GRAD)
                if (active)
                    dkappa[exc + static_cast<size_t>(dm.nex) * (eyc + static_cast<size_t>(dm.ney) * ez)] +=
                        double(ndt * gacc);
        }
        __syncthreads();
        flush_plane(ez);
    }
    // the last KL planes (their contributions finished before the loop's last barrier)
    for (int z = ez1; z < ez1 + KL; ++z) flush_plane(z);

    const double l2d = wave_sum(double(l2));
    if (lane == 0) l2w[wv] = l2d;
    __syncthreads();
    if (tid == 0) {
        double sum = 0.0;
#pragma unroll
        for (int w = 0; w < kRowsWg; ++w) sum += l2w[w];
        l2_part[blockIdx.x] = sum;
    }
}

// LDS ordering between the lanes of one wave
__device__ __forceinline__ void wave_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Position-first heat assembly (the operator of k_heat, its three modes).
// The y and z contractions do not involve the element's x index, so they are
// applied once per x position j of a wave's row instead of once per element
// and x test index:
//   GM[j](ayo, azo) = sum_{ay,az} My(ayo,ay) Mz(azo,az) c(j, ay, az)
//   GS[j](ayo, azo) = sum_{ay,az} (Ky(ayo,ay) Mz(azo,az) + My(ayo,ay) Kz(azo,az)) c(j, ay, az)
// and element e (lane e, DOFs e..e+p along x) takes the x contractions over
// positions e..e+p:
//   F_e(axo, q) = sum_b Mx(axo,b) GM[e+b](q)
//                 - dt kappa_e (sum_b Kx(axo,b) GM[e+b](q) + sum_b Mx(axo,b) GS[e+b](q))
//   l2_e = sum_{axo,q} c(e+axo, q) sum_b Mx(axo,b) GM[e+b](q)
// 5 P^3 (y, z) + 3 P^4 (x) + 2 P^3 (combine, l2) multiply-adds per element
// against 8 P^4 + 2 P^3.  A lane computes GM, GS of its own position; the p
// positions past the wave's 64 (the halo) of every row go into LDS before a
// barrier, one (row, position, azo) triple per lane on one or two waves.
// Neighbouring positions' values go through a per-wave LDS slot, one (ayo,
// azo) at a time.  The plain instances up to p = 3 are held to 4 waves per
// SIMD (the halo pass would otherwise take them to 3).
// kHeatGrad: c is mu, and dkappa[e] += -dt cn_e^T (K_e mu_e) from a second
// coefficient ring.  kHeatDual: GM from c (values), GM and GS from cn
// (gradients) for the stiffness part.
template <typename T, int P, int MODE>
__global__ __launch_bounds__(kWave* kRowsWg, P <= 4 && MODE == 0 ? 4 : 1) void k_heat_pf(const double* __restrict__ c, T* __restrict__ F,
                                                            const T* __restrict__ emx, const T* __restrict__ emy,
                                                            const T* __restrict__ emz,
                                                            const double* __restrict__ kappa, T ndt, AsmDims dm,
                                                            double* __restrict__ l2_part,
                                                            const double* __restrict__ cn,
                                                            double* __restrict__ dkappa)
{
    constexpr int KL = P - 1, W = kWave + KL, R = kRowsWg + KL, NT = P * (P + 1) / 2;
    constexpr int NTH = kWave * kRowsWg;
    constexpr bool GRAD = MODE == kHeatGrad, DUAL = MODE == kHeatDual, TWO = GRAD || DUAL;
    constexpr int NV = DUAL ? 3 : 2;  // exchanged values per position: GM(c) [, GM(cn)], GS
    constexpr int NHC = kRowsWg * KL * P;  // halo (row, position, azo) triples per element layer
    constexpr int NHW = (NHC + kWave - 1) / kWave;  // waves that compute them
    static_assert(NHW <= kRowsWg, "the halo triples fit the workgroup's waves");
    using A = double;
    __shared__ T cs[P][R][W];                            // coefficient planes, ring by z
    __shared__ A fs[P][R][W];                            // right-hand-side planes, ring by z
    __shared__ T ns[TWO ? P : 1][TWO ? R : 1][TWO ? W : 1];  // second field (cn)
    __shared__ T xb[kRowsWg][kWave][NV];                 // one (ayo, azo) per position
    __shared__ T hal[kRowsWg][P][KL][P][NV];             // halo: [row][azo][position - 64][ayo][.]
    __shared__ T eyl[kRowsWg][2 * NT];                   // y element matrices of the rows
    __shared__ double l2w[kRowsWg];

    const int tid = threadIdx.x;
    const int lane = tid & (kWave - 1);
    const int wv = __builtin_amdgcn_readfirstlane(tid / kWave);
    const int tile = xcd_tile(blockIdx.x, gridDim.x);
    const int xc = tile % dm.xchunks;
    const int yb = (tile / dm.xchunks) % dm.yblocks;
    const int seg = tile / (dm.xchunks * dm.yblocks);
    const int ex0 = xc * kWave, ey0 = yb * kRowsWg;
    const int ex = ex0 + lane, ey = ey0 + wv;
    const int ez0 = seg * dm.zseg;
    const int ez1 = min(ez0 + dm.zseg, dm.nez);
    const bool row_ok = ey < dm.ney;
    const bool active = row_ok && ex < dm.nex;
    const int exc = min(ex, dm.nex - 1), eyc = min(ey, dm.ney - 1);
    const int sx = dm.sx;
    const long plane = static_cast<long>(sx) * dm.sy;
    const int ncol = min(W, sx - ex0);
    const int nrow = min(R, dm.sy - ey0);

    // the lane's x element matrices (packed upper triangles; zero rows for an
    // inactive lane, which then adds exact zeros)
    const T* __restrict__ Ex = emx + static_cast<size_t>(exc) * 2 * NT;
    const T* __restrict__ Ey = emy + static_cast<size_t>(eyc) * 2 * NT;
    for (int i = tid; i < kRowsWg * 2 * NT; i += NTH) {
        const int r = i / (2 * NT);
        (&eyl[0][0])[i] = emy[static_cast<size_t>(min(ey0 + r, dm.ney - 1)) * 2 * NT + i % (2 * NT)];
    }

    for (int i = tid; i < P * R * W; i += NTH) (&fs[0][0][0])[i] = A(0);

    constexpr int NE = R * W, NPT = (NE + NTH - 1) / NTH;
    T pc[NPT], pn[TWO ? NPT : 1];
    auto fetch = [&](int z) {
        const long off = ex0 + static_cast<long>(sx) * ey0 + plane * z;
#pragma unroll
        for (int k = 0; k < NPT; ++k) {
            const int i = tid + k * NTH;
            const int r = i / W, col = i - r * W;
            const bool in = i < NE && col < ncol && r < nrow;
            pc[k] = in ? T(c[off + col + static_cast<long>(sx) * r]) : T(0);
            if constexpr (TWO) pn[k] = in ? T(cn[off + col + static_cast<long>(sx) * r]) : T(0);
        }
    };
    auto stash = [&](int z) {
        const int slot = z % P;
#pragma unroll
        for (int k = 0; k < NPT; ++k) {
            const int i = tid + k * NTH;
            if (i < NE) {
                const int r = i / W, col = i - r * W;
                cs[slot][r][col] = pc[k];
                if constexpr (TWO) ns[slot][r][col] = pn[k];
            }
        }
    };
    auto flush_plane = [&](int z) {
        const int slot = z % P;
        T* dst = F + ex0 + static_cast<long>(sx) * ey0 + plane * z;
        for (int i = tid; i < R * W; i += NTH) {
            const int r = i / W, col = i - r * W;
            if (col < ncol && r < nrow) atomic_add(&dst[col + static_cast<long>(sx) * r], T(fs[slot][r][col]));
            fs[slot][r][col] = A(0);
        }
    };
    // z then y contraction of a column cc[ay][az] for output plane azo:
    // gm[ayo] = (My (x) Mz cc)(ayo, azo), gs[ayo] = ((Ky (x) Mz + My (x) Kz) cc)(ayo, azo)
    auto yz = [&](const T(&cc)[P][P], const T* __restrict__ Ez, const T* my, const T* ky, int azo, T(&gm)[P],
                  T(&gs)[P]) {
        T tm[P], tk[P];
#pragma unroll
        for (int ay = 0; ay < P; ++ay) {
            T m = Ez[tri<P>(azo, 0)] * cc[ay][0];
            T k = Ez[NT + tri<P>(azo, 0)] * cc[ay][0];
#pragma unroll
            for (int az = 1; az < P; ++az) {
                m = fma(Ez[tri<P>(azo, az)], cc[ay][az], m);
                k = fma(Ez[NT + tri<P>(azo, az)], cc[ay][az], k);
            }
            tm[ay] = m;
            tk[ay] = k;
        }
#pragma unroll
        for (int ayo = 0; ayo < P; ++ayo) {
            T m = my[tri<P>(ayo, 0)] * tm[0];
            T k = fma(ky[tri<P>(ayo, 0)], tm[0], my[tri<P>(ayo, 0)] * tk[0]);
#pragma unroll
            for (int ay = 1; ay < P; ++ay) {
                m = fma(my[tri<P>(ayo, ay)], tm[ay], m);
                k = fma(ky[tri<P>(ayo, ay)], tm[ay], k);
                k = fma(my[tri<P>(ayo, ay)], tk[ay], k);
            }
            gm[ayo] = m;
            gs[ayo] = k;
        }
    };
    // the exchanged values of a column at (row r0, position col) for plane azo:
    // v[ayo][0] = GM(c), [1] = GS (of c, or of cn in DUAL), [2] = GM(cn) (DUAL)
    auto column_values = [&](int r0, int col, const T* __restrict__ Ez, const T* my, const T* ky, int azo,
                             const int (&slot)[P], T(&v)[P][NV]) {
        T cc[P][P];
#pragma unroll
        for (int ay = 0; ay < P; ++ay)
#pragma unroll
            for (int az = 0; az < P; ++az) cc[ay][az] = cs[slot[az]][r0 + ay][col];
        T gm[P], gs[P];
        yz(cc, Ez, my, ky, azo, gm, gs);
        if constexpr (DUAL) {
#pragma unroll
            for (int ay = 0; ay < P; ++ay)
#pragma unroll
                for (int az = 0; az < P; ++az) cc[ay][az] = ns[slot[az]][r0 + ay][col];
            T gmn[P], gsn[P];
            yz(cc, Ez, my, ky, azo, gmn, gsn);
#pragma unroll
            for (int ayo = 0; ayo < P; ++ayo) {
                v[ayo][0] = gm[ayo];
                v[ayo][1] = gsn[ayo];
                v[ayo][2] = gmn[ayo];
            }
        } else {
#pragma unroll
            for (int ayo = 0; ayo < P; ++ayo) {
                v[ayo][0] = gm[ayo];
                v[ayo][1] = gs[ayo];
            }
        }
    };

    for (int z = ez0; z < ez0 + KL; ++z) {
        fetch(z);
        stash(z);
    }
    fetch(ez0 + KL);

    T l2 = T(0);
    for (int ez = ez0; ez < ez1; ++ez) {
        stash(ez + KL);
        if (ez + 1 < ez1) fetch(ez + KL + 1);
        __syncthreads();
        const T* __restrict__ Ez = emz + static_cast<size_t>(ez) * 2 * NT;
        int slot[P];
#pragma unroll
        for (int az = 0; az < P; ++az) slot[az] = (ez + az) % P;
        // the halo positions' values: the (row, position, azo) triples of the
        // layer, one per lane, on ceil(triples / 64) waves (one at p <= 3, two
        // at p = 4); which waves take them rotates with the layer and the
        // workgroup, so that no SIMD takes them all
        {
            const int hw = (wv + kRowsWg - (ez + static_cast<int>(blockIdx.x)) % kRowsWg) % kRowsWg;
            if (hw < NHW) {
                // the other waves of the workgroup wait for this at the next
                // barrier: issue it ahead of the co-resident workgroups' waves
                __builtin_amdgcn_s_setprio(2);
                const int q0 = hw * kWave + lane;
                const bool ha = q0 < NHC;
                const int q = ha ? q0 : 0;
                const int hz = q / (kRowsWg * KL), hrw = (q / KL) % kRowsWg, hp = q % KL;
                T v[P][NV];
                column_values(hrw, kWave + hp, Ez, &eyl[hrw][0], &eyl[hrw][NT], hz, slot, v);
                if (ha) {
#pragma unroll
                    for (int ayo = 0; ayo < P; ++ayo)
#pragma unroll
                        for (int k = 0; k < NV; ++k) hal[hrw][hz][hp][ayo][k] = v[ayo][k];
                }
                __builtin_amdgcn_s_setprio(0);
            }
        }
        __syncthreads();
        if (row_ok) {
            const size_t eidx = exc + static_cast<size_t>(dm.nex) * (eyc + static_cast<size_t>(dm.ney) * ez);
            const T kap = kappa ? T(kappa[eidx]) : T(1);
            const T s = ndt * kap;
            T gacc = T(0);  // cn_e^T K_e mu_e (GRAD)
            // read per layer (cache hits): kept across the layers they would
            // take 2 P(P+1)/2 VGPRs through the halo pass too
            const T* ex_ = Ex;
            asm volatile("" : "+v"(ex_));
            T mxs[NT], kxs[NT];
#pragma unroll
            for (int t = 0; t < NT; ++t) {
                mxs[t] = active ? ex_[t] : T(0);
                kxs[t] = active ? ex_[NT + t] : T(0);
            }
#pragma unroll 1
            for (int azo = 0; azo < P; ++azo) {
                T vo[P][NV];
                column_values(wv, lane, Ez, Ey, Ey + NT, azo, slot, vo);
#pragma unroll
                for (int ayo = 0; ayo < P; ++ayo) {
#pragma unroll
                    for (int k = 0; k < NV; ++k) xb[wv][lane][k] = vo[ayo][k];
                    wave_lds_sync();
                    T g[P][NV];
#pragma unroll
                    for (int k = 0; k < NV; ++k) g[0][k] = vo[ayo][k];
#pragma unroll
                    for (int b = 1; b < P; ++b) {
                        const int src = lane + b;
                        const T* q = src < kWave ? &xb[wv][src][0] : &hal[wv][azo][src - kWave][ayo][0];
#pragma unroll
                        for (int k = 0; k < NV; ++k) g[b][k] = q[k];
                    }
                    wave_lds_sync();
#pragma unroll
                    for (int axo = 0; axo < P; ++axo) {
                        // om = (Mx GM(c)), ok = (Kx GM(cn|c) + Mx GS)
                        constexpr int KG = DUAL ? 2 : 0;
                        T om = mxs[tri<P>(axo, 0)] * g[0][0];
                        T ok = fma(kxs[tri<P>(axo, 0)], g[0][KG], mxs[tri<P>(axo, 0)] * g[0][1]);
#pragma unroll
                        for (int b = 1; b < P; ++b) {
                            om = fma(mxs[tri<P>(axo, b)], g[b][0], om);
                            ok = fma(kxs[tri<P>(axo, b)], g[b][KG], ok);
                            ok = fma(mxs[tri<P>(axo, b)], g[b][1], ok);
                        }
                        l2 = fma(cs[slot[azo]][wv + ayo][lane + axo], om, l2);
                        if constexpr (GRAD) gacc = fma(ns[slot[azo]][wv + ayo][lane + axo], ok, gacc);
                        // lanes of one wave address distinct columns here; the
                        // LDS applies one wave's adds in program order
                        atomic_add(&fs[slot[azo]][wv + ayo][lane + axo], A(fma(s, ok, om)));
                    }
                }
            }
            if constexpr (GRAD)
                if (active) dkappa[eidx] += double(ndt * gacc);
        }
        __syncthreads();
        flush_plane(ez);
    }
    for (int z = ez1; z < ez1 + KL; ++z) flush_plane(z);

    const double l2d = wave_sum(double(l2));
    if (lane == 0) l2w[wv] = l2d;
    __syncthreads();
    if (tid == 0) {
        double sum = 0.0;
#pragma unroll
        for (int w = 0; w < kRowsWg; ++w) sum += l2w[w];
        l2_part[blockIdx.x] = sum;
    }
}

// t = 0 projection: F_a = sum_e sum_q J W f0(X_q) v_a(X_q), l2 = sum J W f0^2.
// A wave owns 64 consecutive elements along x at one ey and walks a segment
// of ez; f0 at the points is applied and the transposed contractions run
// z, y, x into an LDS ring of right-hand-side planes (as in k_heat).
template <typename T, int P>
__global__ __launch_bounds__(kWave) void k_project(T* __restrict__ F, const T* __restrict__ tabx,
                                                   const T* __restrict__ taby, const T* __restrict__ tabz,
                                                   const T* __restrict__ jacx, const T* __restrict__ jacy,
                                                   const T* __restrict__ jacz, const T* __restrict__ Wx,
                                                   const T* __restrict__ Wy, const T* __restrict__ Wz,
                                                   const T* __restrict__ ptsx, const T* __restrict__ ptsy,
                                                   const T* __restrict__ ptsz, AsmDims dm,
                                                   double* __restrict__ l2_part)
{
    constexpr int Q = P, KL = P - 1, W = kWave + KL;
    __shared__ T fs[P][P][W];  // right-hand-side planes, ring by z

    const int lane = threadIdx.x;
    const int tile = xcd_tile(blockIdx.x, gridDim.x);
    const int xc = tile % dm.xchunks;
    const int ey = (tile / dm.xchunks) % dm.ney;
    const int seg = tile / (dm.xchunks * dm.ney);
    const int ex0 = xc * kWave;
    const int ex = ex0 + lane;
    const int ez0 = seg * dm.zseg;
    const int ez1 = min(ez0 + dm.zseg, dm.nez);
    const bool active = ex < dm.nex;
    const int exc = active ? ex : dm.nex - 1;
    const int sx = dm.sx;
    const long plane = static_cast<long>(sx) * dm.sy;
    const int ncol = min(W, sx - ex0);

    const T* __restrict__ Nx = tabx + static_cast<size_t>(exc) * Q * P * 2;
    const T* __restrict__ Ny = taby + static_cast<size_t>(ey) * Q * P * 2;
    const T Jxy = jacx[exc] * jacy[ey];

    for (int i = lane; i < P * P * W; i += kWave) (&fs[0][0][0])[i] = T(0);
    auto flush_plane = [&](int z) {
        const int slot = z % P;
        T* dst = F + ex0 + static_cast<long>(sx) * ey + plane * z;
        for (int i = lane; i < P * W; i += kWave) {
            const int ay = i / W, col = i - ay * W;
            if (col < ncol) atomic_add(&dst[col + static_cast<long>(sx) * ay], fs[slot][ay][col]);
            fs[slot][ay][col] = T(0);
        }
    };
    __syncthreads();

    T l2 = T(0);
    for (int ez = ez0; ez < ez1; ++ez) {
        const T* __restrict__ Nz = tabz + static_cast<size_t>(ez) * Q * P * 2;
        const T J = Jxy * jacz[ez];
        const int s0 = ez % P;
        int slot[P];
#pragma unroll
        for (int az = 0; az < P; ++az) slot[az] = (s0 + az) % P;

#pragma unroll 1
        for (int qx = 0; qx < Q; ++qx) {
            T n0x[P];
#pragma unroll
            for (int a = 0; a < P; ++a) n0x[a] = Nx[(qx * P + a) * 2];
            T t0[P][P];
#pragma unroll
            for (int ay = 0; ay < P; ++ay)
#pragma unroll
                for (int az = 0; az < P; ++az) t0[ay][az] = T(0);
            const T xq = ptsx[static_cast<size_t>(exc) * Q + qx];
            const T wxq = Wx[qx];
#pragma unroll 1
            for (int qy = 0; qy < Q; ++qy) {
                T ny0[P];
#pragma unroll
                for (int a = 0; a < P; ++a) ny0[a] = Ny[(qy * P + a) * 2];
                T h00[P];
#pragma unroll
                for (int az = 0; az < P; ++az) h00[az] = T(0);
                const T yq = ptsy[static_cast<size_t>(ey) * Q + qy];
                const T wxy = wxq * Wy[qy];
#pragma unroll
                for (int qz = 0; qz < Q; ++qz) {
                    const T w = J * (wxy * Wz[qz]);
                    const T zq = ptsz[static_cast<size_t>(ez) * Q + qz];
                    const T fv = f0_eval<T>(dm.f0, xq, yq, zq);
                    const T g0 = w * fv;
                    l2 = fma(g0, fv, l2);
#pragma unroll
                    for (int az = 0; az < P; ++az) h00[az] = fma(Nz[(qz * P + az) * 2], g0, h00[az]);
                }
#pragma unroll
                for (int ay = 0; ay < P; ++ay)
#pragma unroll
                    for (int az = 0; az < P; ++az) t0[ay][az] = fma(ny0[ay], h00[az], t0[ay][az]);
            }
            if (active) {
#pragma unroll
                for (int ax = 0; ax < P; ++ax)
#pragma unroll
                    for (int ay = 0; ay < P; ++ay)
#pragma unroll
                        for (int az = 0; az < P; ++az) atomic_add(&fs[slot[az]][ay][lane + ax], n0x[ax] * t0[ay][az]);
            }
        }
        __syncthreads();
        flush_plane(ez);
        __syncthreads();
    }
    for (int z = ez1; z < ez1 + KL; ++z) flush_plane(z);

    double l2d = active ? double(l2) : 0.0;
    l2d = wave_sum(l2d);
    if (lane == 0) l2_part[blockIdx.x] = l2d;
}

// ---------------------------------------------------------------------------
// The oil problem (problems/oil): integrand at t > 0 (RHS_eq.F90:76-87)
//   J W (v u + Dt (-kq(X) exp(mi u) grad u . grad v + v (pump(X) - max(0, drain(X) u))))
// It is nonlinear in u, so there are no element matrices: u and grad u are
// evaluated at the Gauss points (FormUn, sum-factorised x, y, z), the point
// weights of v and of grad v formed, and contracted back z, y, x into the
// right-hand-side ring (as k_project).  kq is cached per point (CacheKqValues,
// input_data.F90:333-380).  The distance and kq arithmetic runs without
// contraction, in the reference's order, so the cache matches the oracle
// bit for bit.
// ---------------------------------------------------------------------------
__device__ __forceinline__ double oil_falloff(double r, double R, double t)
{
#pragma clang fp contract(off)
    if (t < r) return 1.0;
    if (t > R) return 0.0;
    const double h = (t - r) / (R - r);
    const double g = (h - 1.0) * (h + 1.0);
    return g * g;
}

// Segment s = 0 .. cN (cL - 1) - 1 of the curves: its end point index k
// (points k - 1 and k).
__device__ __forceinline__ int oil_seg_end(const OilDev& o, int s)
{
    return (s / (o.cL - 1)) * o.cL + s % (o.cL - 1) + 1;
}

// Whether segment (k - 1, k) can lie within sqrt(cut2) of a point of the box
// [lo, hi]: the distance between the boxes is a lower bound of the
// segment's distance to any point of the box.
__device__ __forceinline__ bool oil_seg_near(const OilDev& o, int k, const double (&lo)[3], const double (&hi)[3],
                                             double cut2)
{
    const int npt = o.cN * o.cL;
    double d2 = 0.0;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const double u = o.curves[a * npt + k - 1], v = o.curves[a * npt + k];
        const double gap = fmax(0.0, fmax(lo[a] - fmax(u, v), fmin(u, v) - hi[a]));
        d2 += gap * gap;
    }
    return d2 <= cut2;
}

// The candidate segments of a falloff of radius R: within R (with a margin
// for rounding) of the box.  Every segment whose computed distance to a
// point of the box is below R^2 (1 + 1e-9) is listed, so the minimum over
// the list equals the minimum over all segments wherever falloff(0, R, .)
// of it is not 0, and both give 0 elsewhere: the results are identical.
__device__ __forceinline__ double oil_cut2(double R)
{
    return R * R * (1.0 + 1e-9);
}

// squared distance from the curves (dist_from_curves / dist_from_segment,
// input_data.F90:171-238), at most 1e3, over the segments ends[0 .. nlist)
// (their order does not matter: the minimum is exact)
__device__ double oil_dist_curves(const OilDev& o, const int* __restrict__ ends, int nlist, double px, double py,
                                  double pz)
{
#pragma clang fp contract(off)
    const int npt = o.cN * o.cL;
    const double* __restrict__ cx = o.curves;
    const double* __restrict__ cy = o.curves + npt;
    const double* __restrict__ cz = o.curves + 2 * npt;
    double fval = 1e3;
    for (int s = 0; s < nlist; ++s) {
        const int k = ends[s];
        const double ax = cx[k - 1], ay = cy[k - 1], az = cz[k - 1];
        const double dx = cx[k] - ax, dy = cy[k] - ay, dz = cz[k] - az;
        const double qx = px - ax, qy = py - ay, qz = pz - az;
        const double dot = dx * qx + dy * qy + dz * qz;
        const double len2 = dx * dx + dy * dy + dz * dz;
        const double proj = dot / len2;
        double xx, yy, zz;
        if (proj < 0) {
            xx = ax; yy = ay; zz = az;
        } else if (proj > 1) {
            xx = cx[k]; yy = cy[k]; zz = cz[k];
        } else {
            xx = ax + proj * dx; yy = ay + proj * dy; zz = az + proj * dz;
        }
        const double ex = px - xx, ey = py - yy, ez = pz - zz;
        const double d = ex * ex + ey * ey + ez * ez;
        if (d < fval) fval = d;
    }
    return fval;
}

// sum over the sources at (x, y, z) of strength * falloff(0, radius, |q - X|).
// A source with |q - X|^2 > radius^2 (1 + 1e-12) has sqrt > radius after
// rounding, so falloff is exactly 0 and adding its term leaves the sum as
// it is: such sources are skipped before the square root and the division.
__device__ __forceinline__ double oil_sources(const double* __restrict__ q, int nq, double strength, double radius,
                                              double x, double y, double z)
{
#pragma clang fp contract(off)
    const double reject = radius * radius * (1.0 + 1e-12);
    double f = 0.0;
    for (int i = 0; i < nq; ++i) {
        const double dx = q[3 * i] - x, dy = q[3 * i + 1] - y, dz = q[3 * i + 2] - z;
        const double d2 = dx * dx + dy * dy + dz * dz;
        if (d2 > reject) continue;
        f = f + strength * oil_falloff(0.0, radius, sqrt(d2));
    }
    return f;
}

// Whether a source of q can lie within its radius of a point of the box
// [lo, hi]: the box distance is a lower bound of every point's distance, and
// the margin keeps the test conservative under rounding, so when it says no,
// oil_sources returns exactly 0 for every point of the box.
__device__ __forceinline__ bool oil_sources_near(const double* __restrict__ q, int nq, double radius,
                                                 const double (&lo)[3], const double (&hi)[3])
{
    const double cut = radius * radius * (1.0 + 1e-9);
    bool near = false;
    for (int i = 0; i < nq; ++i) {
        double d2 = 0.0;
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            const double v = q[3 * i + a];
            const double gap = fmax(0.0, fmax(lo[a] - v, v - hi[a]));
            d2 += gap * gap;
        }
        near = near || d2 <= cut;
    }
    return near;
}

// initial_state (input_data.F90:317-329): 0.1 lerp(falloff(0, .1, dist), 0, 1) bump3d(.2, .6, X),
// the distance over the candidate segments of radius 0.1
__device__ double oil_initial(const OilDev& o, const int* __restrict__ ends, int nlist, double x, double y, double z)
{
#pragma clang fp contract(off)
    const double dist = sqrt(oil_dist_curves(o, ends, nlist, x, y, z));
    const double bx = x - 0.5, by = y - 0.5, bz = z - 0.5;
    const double b = oil_falloff(0.2 / 2.0, 0.6 / 2.0, sqrt(bx * bx + by * by + bz * bz));
    const double t = oil_falloff(0.0, 0.1, dist);
    return 0.1 * ((1.0 - t) * 0.0 + t * 1.0) * b;
}

// One thread per Gauss point.  The block lists the candidate segments of
// radius 0.06 (the falloff of kq) for the box of its points first (in LDS,
// cN (cL - 1) entries at most), so a point scans a few segments instead of
// all of them.
template <int Q>
__global__ __launch_bounds__(256) void k_oil_kq(const double* __restrict__ ptsx, const double* __restrict__ ptsy,
                                                const double* __restrict__ ptsz, int nex, int ney, int nez, OilDev o,
                                                double* __restrict__ kq)
{
#pragma clang fp contract(off)
    extern __shared__ int ends[];
    __shared__ double wlo[4][3], whi[4][3];
    __shared__ int nlist;
    const long n = static_cast<long>(nex) * ney * nez * Q * Q * Q;
    const long i0 = blockIdx.x * static_cast<long>(blockDim.x) + threadIdx.x;
    const long i = i0 < n ? i0 : n - 1;  // past the end: the last point (not stored)
    const int kx = static_cast<int>(i % Q), ky = static_cast<int>((i / Q) % Q), kz = static_cast<int>((i / (Q * Q)) % Q);
    const long e = i / (Q * Q * Q);
    const int ex = static_cast<int>(e % nex), ey = static_cast<int>((e / nex) % ney), ez = static_cast<int>(e / (static_cast<long>(nex) * ney));
    const double pt[3] = {ptsx[ex * Q + kx], ptsy[ey * Q + ky], ptsz[ez * Q + kz]};
    // the block's bounding box: per wave by shuffles, then across the waves
    const int lane = threadIdx.x % kWave, w = threadIdx.x / kWave;
    if (threadIdx.x == 0) nlist = 0;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        double lo = pt[a], hi = pt[a];
#pragma unroll
        for (int off = kWave / 2; off > 0; off >>= 1) {
            lo = fmin(lo, __shfl_xor(lo, off, kWave));
            hi = fmax(hi, __shfl_xor(hi, off, kWave));
        }
        if (lane == 0) {
            wlo[w][a] = lo;
            whi[w][a] = hi;
        }
    }
    __syncthreads();
    double lo[3], hi[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        lo[a] = wlo[0][a];
        hi[a] = whi[0][a];
        for (int v = 1; v < static_cast<int>(blockDim.x / kWave); ++v) {
            lo[a] = fmin(lo[a], wlo[v][a]);
            hi[a] = fmax(hi[a], whi[v][a]);
        }
    }
    const int nseg = o.cN * (o.cL - 1);
    const double cut2 = oil_cut2(0.06);
    for (int s = threadIdx.x; s < nseg; s += blockDim.x) {
        const int k = oil_seg_end(o, s);
        if (oil_seg_near(o, k, lo, hi, cut2)) ends[atomicAdd(&nlist, 1)] = k;
    }
    __syncthreads();
    const double dist = sqrt(oil_dist_curves(o, ends, nlist, pt[0], pt[1], pt[2]));
    const double t = oil_falloff(0.0, 0.06, dist);
    // stored point-major (kq[(kx + Q (ky + Q kz)) E + e]), so the step kernel's
    // lanes (consecutive elements) read each point's values coalesced
    if (i0 < n) kq[(kx + Q * (ky + Q * kz)) * (n / (Q * Q * Q)) + e] = (1.0 - t) * o.kq_min + t * o.kq_max;
}

template <int P, bool TZERO>
__global__ __launch_bounds__(kWave) void k_oil(const double* __restrict__ c, double* __restrict__ F,
                                               const double* __restrict__ tabx, const double* __restrict__ taby,
                                               const double* __restrict__ tabz, const double* __restrict__ jacx,
                                               const double* __restrict__ jacy, const double* __restrict__ jacz,
                                               const double* __restrict__ Wx, const double* __restrict__ Wy,
                                               const double* __restrict__ Wz, const double* __restrict__ ptsx,
                                               const double* __restrict__ ptsy, const double* __restrict__ ptsz,
                                               AsmDims dm, OilDev o, const double* __restrict__ kq, double dt,
                                               double* __restrict__ l2_part, double* __restrict__ dr_part)
{
    constexpr int Q = P, KL = P - 1, W = kWave + KL;
    __shared__ double fs[P][P][W];  // right-hand-side planes, ring by z

    const int lane = threadIdx.x;
    const int tile = xcd_tile(blockIdx.x, gridDim.x);
    const int xc = tile % dm.xchunks;
    const int ey = (tile / dm.xchunks) % dm.ney;
    const int seg = tile / (dm.xchunks * dm.ney);
    const int ex0 = xc * kWave;
    const int ex = ex0 + lane;
    const int ez0 = seg * dm.zseg;
    const int ez1 = min(ez0 + dm.zseg, dm.nez);
    const bool active = ex < dm.nex;
    const int exc = active ? ex : dm.nex - 1;
    const int sx = dm.sx;
    const long plane = static_cast<long>(sx) * dm.sy;
    const int ncol = min(W, sx - ex0);
    // TZERO: the candidate segments of radius 0.1 (initial_state's falloff)
    // for the current element layer's points (cN (cL - 1) entries at most)
    extern __shared__ int ends[];
    int nlist = 0;

    const double* __restrict__ Nx = tabx + static_cast<size_t>(exc) * Q * P * 2;
    const double* __restrict__ Ny = taby + static_cast<size_t>(ey) * Q * P * 2;
    const double Jxy = jacx[exc] * jacy[ey];
    const long ne = static_cast<long>(dm.nex) * dm.ney * dm.nez;  // elements (the kq cache's point stride)

    for (int i = lane; i < P * P * W; i += kWave) (&fs[0][0][0])[i] = 0.0;
    auto flush_plane = [&](int z) {
        const int slot = z % P;
        double* dst = F + ex0 + static_cast<long>(sx) * ey + plane * z;
        for (int i = lane; i < P * W; i += kWave) {
            const int ay = i / W, col = i - ay * W;
            if (col < ncol) atomic_add(&dst[col + static_cast<long>(sx) * ay], fs[slot][ay][col]);
            fs[slot][ay][col] = 0.0;
        }
    };
    __syncthreads();

    double l2 = 0.0, dr = 0.0;
    for (int ez = ez0; ez < ez1; ++ez) {
        const double* __restrict__ Nz = tabz + static_cast<size_t>(ez) * Q * P * 2;
        const double J = Jxy * jacz[ez];
        const int s0 = ez % P;
        int slot[P];
#pragma unroll
        for (int az = 0; az < P; ++az) slot[az] = (s0 + az) % P;
        const long e = exc + static_cast<long>(dm.nex) * (ey + static_cast<long>(dm.ney) * ez);
        if constexpr (TZERO) {
            // the layer's points span x of elements ex0 .. ex0+63 and y, z of (ey, ez)
            const int exl = min(ex0 + kWave - 1, dm.nex - 1);
            const double lo[3] = {ptsx[static_cast<size_t>(ex0) * Q], ptsy[static_cast<size_t>(ey) * Q],
                                  ptsz[static_cast<size_t>(ez) * Q]};
            const double hi[3] = {ptsx[static_cast<size_t>(exl) * Q + Q - 1], ptsy[static_cast<size_t>(ey) * Q + Q - 1],
                                  ptsz[static_cast<size_t>(ez) * Q + Q - 1]};
            const int nseg = o.cN * (o.cL - 1);
            const double cut2 = oil_cut2(0.1);
            __syncthreads();  // the previous layer's readers are done
            nlist = 0;
            for (int s0 = 0; s0 < nseg; s0 += kWave) {
                const int s = s0 + lane;
                const int k = s < nseg ? oil_seg_end(o, s) : 0;
                const bool near = s < nseg && oil_seg_near(o, k, lo, hi, cut2);
                const unsigned long long m = __ballot(near);
                if (near) ends[nlist + __popcll(m & ((1ull << lane) - 1))] = k;
                nlist += __popcll(m);
            }
            __syncthreads();
        }
        // the element's coefficients: first DOF of element e is e (open knots)
        const double* __restrict__ ce = c ? c + exc + static_cast<long>(sx) * ey + plane * ez : nullptr;
        // pumps and drains near the element's points (most elements have
        // none, and their point loops over the sources are skipped)
        bool pump_near = false, drain_near = false;
        if constexpr (!TZERO) {
            const double* pt[3] = {ptsx + static_cast<size_t>(exc) * Q, ptsy + static_cast<size_t>(ey) * Q,
                                   ptsz + static_cast<size_t>(ez) * Q};
            double lo[3], hi[3];
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                lo[a] = hi[a] = pt[a][0];
#pragma unroll
                for (int q = 1; q < Q; ++q) {
                    lo[a] = fmin(lo[a], pt[a][q]);
                    hi[a] = fmax(hi[a], pt[a][q]);
                }
            }
            pump_near = oil_sources_near(o.pumps, o.npumps, o.radius, lo, hi);
            drain_near = oil_sources_near(o.drains, o.ndrains, o.radius, lo, hi);
        }

#pragma unroll 1
        for (int qx = 0; qx < Q; ++qx) {
            double nx0[P], nx1[P];
#pragma unroll
            for (int a = 0; a < P; ++a) {
                nx0[a] = Nx[(qx * P + a) * 2];
                nx1[a] = Nx[(qx * P + a) * 2 + 1];
            }
            // x stage of FormUn: u and du/dx along x at qx, for every (ay, az)
            double ux0[P][P], ux1[P][P];
            if constexpr (!TZERO) {
#pragma unroll
                for (int ay = 0; ay < P; ++ay)
#pragma unroll
                    for (int az = 0; az < P; ++az) {
                        const double* col = ce + static_cast<long>(sx) * ay + plane * az;
                        double a0 = nx0[0] * col[0], a1 = nx1[0] * col[0];
#pragma unroll
                        for (int b = 1; b < P; ++b) {
                            a0 = fma(nx0[b], col[b], a0);
                            a1 = fma(nx1[b], col[b], a1);
                        }
                        ux0[ay][az] = a0;
                        ux1[ay][az] = a1;
                    }
            }
            double t0[P][P], tx[P][P];
#pragma unroll
            for (int ay = 0; ay < P; ++ay)
#pragma unroll
                for (int az = 0; az < P; ++az) t0[ay][az] = tx[ay][az] = 0.0;
            const double xq = ptsx[static_cast<size_t>(exc) * Q + qx];
            const double wxq = Wx[qx];
#pragma unroll 1
            for (int qy = 0; qy < Q; ++qy) {
                double ny0[P], ny1[P];
#pragma unroll
                for (int a = 0; a < P; ++a) {
                    ny0[a] = Ny[(qy * P + a) * 2];
                    ny1[a] = Ny[(qy * P + a) * 2 + 1];
                }
                double u00[P], u10[P], u01[P];  // y stage: values, d/dx, d/dy along z
                if constexpr (!TZERO) {
#pragma unroll
                    for (int az = 0; az < P; ++az) {
                        double a = ny0[0] * ux0[0][az], b = ny0[0] * ux1[0][az], g = ny1[0] * ux0[0][az];
#pragma unroll
                        for (int ay = 1; ay < P; ++ay) {
                            a = fma(ny0[ay], ux0[ay][az], a);
                            b = fma(ny0[ay], ux1[ay][az], b);
                            g = fma(ny1[ay], ux0[ay][az], g);
                        }
                        u00[az] = a;
                        u10[az] = b;
                        u01[az] = g;
                    }
                }
                double h0[P], hx[P], hy[P];
#pragma unroll
                for (int az = 0; az < P; ++az) h0[az] = hx[az] = hy[az] = 0.0;
                const double yq = ptsy[static_cast<size_t>(ey) * Q + qy];
                const double wxy = wxq * Wy[qy];
#pragma unroll
                for (int qz = 0; qz < Q; ++qz) {
                    double nz0[P], nz1[P];
#pragma unroll
                    for (int a = 0; a < P; ++a) {
                        nz0[a] = Nz[(qz * P + a) * 2];
                        nz1[a] = Nz[(qz * P + a) * 2 + 1];
                    }
                    const double w = J * (wxy * Wz[qz]);
                    const double zq = ptsz[static_cast<size_t>(ez) * Q + qz];
                    double w0, gx = 0.0, gy = 0.0, gz = 0.0;
                    if constexpr (TZERO) {
                        const double f0 = oil_initial(o, ends, nlist, xq, yq, zq);
                        w0 = w * f0;
                        l2 = fma(w0, f0, l2);
                    } else {
                        double u = nz0[0] * u00[0], dux = nz0[0] * u10[0], duy = nz0[0] * u01[0],
                               duz = nz1[0] * u00[0];
#pragma unroll
                        for (int az = 1; az < P; ++az) {
                            u = fma(nz0[az], u00[az], u);
                            dux = fma(nz0[az], u10[az], dux);
                            duy = fma(nz0[az], u01[az], duy);
                            duz = fma(nz1[az], u00[az], duz);
                        }
                        const double kqv = kq[(qx + Q * (qy + Q * qz)) * ne + e];
                        const double pump =
                            pump_near ? oil_sources(o.pumps, o.npumps, o.pump_strength, o.radius, xq, yq, zq) : 0.0;
                        const double dw =
                            drain_near ? oil_sources(o.drains, o.ndrains, o.drain_strength, o.radius, xq, yq, zq) : 0.0;
                        const double vd = fmax(0.0, dw * u);
                        w0 = w * (u + dt * (pump - vd));
                        const double g = -w * dt * kqv * exp(o.mi * u);
                        gx = g * dux;
                        gy = g * duy;
                        gz = g * duz;
                        l2 = fma(w * u, u, l2);
                        dr = fma(w * dt, vd, dr);
                    }
                    // z stage of the transposed contraction
#pragma unroll
                    for (int az = 0; az < P; ++az) {
                        h0[az] = fma(nz1[az], gz, fma(nz0[az], w0, h0[az]));
                        hx[az] = fma(nz0[az], gx, hx[az]);
                        hy[az] = fma(nz0[az], gy, hy[az]);
                    }
                }
#pragma unroll
                for (int ay = 0; ay < P; ++ay)
#pragma unroll
                    for (int az = 0; az < P; ++az) {
                        t0[ay][az] = fma(ny1[ay], hy[az], fma(ny0[ay], h0[az], t0[ay][az]));
                        tx[ay][az] = fma(ny0[ay], hx[az], tx[ay][az]);
                    }
            }
            if (active) {
#pragma unroll
                for (int ax = 0; ax < P; ++ax)
#pragma unroll
                    for (int ay = 0; ay < P; ++ay)
#pragma unroll
                        for (int az = 0; az < P; ++az)
                            atomic_add(&fs[slot[az]][ay][lane + ax], fma(nx1[ax], tx[ay][az], nx0[ax] * t0[ay][az]));
            }
        }
        __syncthreads();
        flush_plane(ez);
        __syncthreads();
    }
    for (int z = ez1; z < ez1 + KL; ++z) flush_plane(z);

    double l2d = wave_sum(active ? l2 : 0.0);
    double drd = wave_sum(active ? dr : 0.0);
    if (lane == 0) {
        l2_part[blockIdx.x] = l2d;
        if (dr_part) dr_part[blockIdx.x] = drd;
    }
}

__global__ __launch_bounds__(256) void k_reduce_partials(const double* __restrict__ part, int n,
                                                         double* __restrict__ out)
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

// z-segment length giving at most `target` workgroups (at least one segment).
int segment_length(long tiles_xy, int nez, long target)
{
    long nseg = target / tiles_xy;
    if (nseg < 1) nseg = 1;
    if (nseg > nez) nseg = nez;
    return static_cast<int>((nez + nseg - 1) / nseg);
}

template <typename T>
AsmDims base_dims(const Tables<T>& tb)
{
    AsmDims dm{};
    dm.nex = tb.d[0].nelem;
    dm.ney = tb.d[1].nelem;
    dm.nez = tb.d[2].nelem;
    dm.sx = tb.d[0].dofs;
    dm.sy = tb.d[1].dofs;
    dm.xchunks = (dm.nex + kWave - 1) / kWave;
    return dm;
}

#ifndef ADS_HEAT_DPPR
#define ADS_HEAT_DPPR 0
#endif

template <typename T, int P, int MODE>
hipError_t heat_p(const double* c, T* F, const Tables<T>& tb, double dt, const double* kappa, double* l2_part,
                  int part_cap, int* nparts, const double* cn, double* dkappa, hipStream_t s)
{
    constexpr bool DPPR = ADS_HEAT_DPPR && MODE == kHeatPlain;
    // two rounds of resident workgroups (a partly filled last round costs
    // as much as a full one)
    static const int resident =
        resident_groups(reinterpret_cast<const void*>(&k_heat<T, P, MODE, DPPR>), kWave * kRowsWg, 0);
    AsmDims dm = base_dims(tb);
    dm.yblocks = (dm.ney + kRowsWg - 1) / kRowsWg;
    dm.zseg = segment_length(static_cast<long>(dm.xchunks) * dm.yblocks, dm.nez, 2L * resident);
    const int nseg = (dm.nez + dm.zseg - 1) / dm.zseg;
    dim3 grid(dm.xchunks * dm.yblocks * nseg);
    *nparts = static_cast<int>(grid.x);
    if (*nparts > part_cap) return hipErrorInvalidValue;
    hipLaunchKernelGGL((k_heat<T, P, MODE, DPPR>), grid, dim3(kWave * kRowsWg), 0, s, c, F, tb.d[0].emat,
                       tb.d[1].emat_sym, tb.d[2].emat_sym, kappa, static_cast<T>(-dt), dm, l2_part, cn, dkappa);
    return hipGetLastError();
}

template <typename T, int P, int MODE>
hipError_t heat_pf_p(const double* c, T* F, const Tables<T>& tb, double dt, const double* kappa, double* l2_part,
                     int part_cap, int* nparts, const double* cn, double* dkappa, hipStream_t s)
{
    static const int resident =
        resident_groups(reinterpret_cast<const void*>(&k_heat_pf<T, P, MODE>), kWave * kRowsWg, 0);
    AsmDims dm = base_dims(tb);
    dm.yblocks = (dm.ney + kRowsWg - 1) / kRowsWg;
    dm.zseg = segment_length(static_cast<long>(dm.xchunks) * dm.yblocks, dm.nez, 2L * resident);
    const int nseg = (dm.nez + dm.zseg - 1) / dm.zseg;
    dim3 grid(dm.xchunks * dm.yblocks * nseg);
    *nparts = static_cast<int>(grid.x);
    if (*nparts > part_cap) return hipErrorInvalidValue;
    hipLaunchKernelGGL((k_heat_pf<T, P, MODE>), grid, dim3(kWave * kRowsWg), 0, s, c, F, tb.d[0].emat_sym,
                       tb.d[1].emat_sym, tb.d[2].emat_sym, kappa, static_cast<T>(-dt), dm, l2_part, cn, dkappa);
    return hipGetLastError();
}

#ifndef ADS_HEAT_PF
#define ADS_HEAT_PF 1
#endif

// the heat assembly: the position-first kernel (k_heat_pf), or with
// -DADS_HEAT_PF=0 at build time the element-first one (k_heat, whose
// -DADS_HEAT_DPPR=1 instance pre-reduces the x offsets by DPP); measured at
// 256^3 elements, p = 3: 1.20, 1.65 and 2.26 ms on one box
template <typename T, int P, int MODE>
hipError_t heat_sel(const double* c, T* F, const Tables<T>& tb, double dt, const double* kappa, double* l2_part,
                    int part_cap, int* nparts, const double* cn, double* dkappa, hipStream_t s)
{
    if constexpr (ADS_HEAT_PF != 0)
        return heat_pf_p<T, P, MODE>(c, F, tb, dt, kappa, l2_part, part_cap, nparts, cn, dkappa, s);
    else
        return heat_p<T, P, MODE>(c, F, tb, dt, kappa, l2_part, part_cap, nparts, cn, dkappa, s);
}

template <typename T, int P>
hipError_t project_p(T* F, const Tables<T>& tb, int f0, double* l2_part, int part_cap, int* nparts, hipStream_t s)
{
    AsmDims dm = base_dims(tb);
    dm.yblocks = dm.ney;
    dm.zseg = segment_length(static_cast<long>(dm.xchunks) * dm.ney, dm.nez, 2048);
    dm.f0 = f0;
    const int nseg = (dm.nez + dm.zseg - 1) / dm.zseg;
    dim3 grid(dm.xchunks * dm.ney * nseg);
    *nparts = static_cast<int>(grid.x);
    if (*nparts > part_cap) return hipErrorInvalidValue;
    hipLaunchKernelGGL((k_project<T, P>), grid, dim3(kWave), 0, s, F, tb.d[0].tab, tb.d[1].tab, tb.d[2].tab,
                       tb.d[0].jac, tb.d[1].jac, tb.d[2].jac, tb.d[0].wts, tb.d[1].wts, tb.d[2].wts, tb.d[0].pts,
                       tb.d[1].pts, tb.d[2].pts, dm, l2_part);
    return hipGetLastError();
}

template <typename T, int P>
hipError_t assemble_p(int form, int f0, const double* c, T* F, const Tables<T>& tb, double dt, const double* kappa,
                      double* l2_part, int part_cap, int* nparts, hipStream_t s)
{
    if (form != kFormHeat) return project_p<T, P>(F, tb, f0, l2_part, part_cap, nparts, s);
    return heat_sel<T, P, kHeatPlain>(c, F, tb, dt, kappa, l2_part, part_cap, nparts, nullptr, nullptr, s);
}

template <int P>
hipError_t oil_p(bool tzero, const double* c, double* F, const Tables<double>& tb, const OilDev& o, const double* kq,
                 double dt, double* l2_part, double* dr_part, int part_cap, int* nparts, hipStream_t s)
{
    AsmDims dm = base_dims(tb);
    dm.yblocks = dm.ney;
    dm.zseg = segment_length(static_cast<long>(dm.xchunks) * dm.ney, dm.nez, 2048);
    const int nseg = (dm.nez + dm.zseg - 1) / dm.zseg;
    dim3 grid(dm.xchunks * dm.ney * nseg);
    *nparts = static_cast<int>(grid.x);
    if (*nparts > part_cap) return hipErrorInvalidValue;
    const size_t list_bytes = sizeof(int) * static_cast<size_t>(o.cN) * (o.cL - 1);
    if (tzero)
        hipLaunchKernelGGL((k_oil<P, true>), grid, dim3(kWave), list_bytes, s, c, F, tb.d[0].tab, tb.d[1].tab, tb.d[2].tab,
                           tb.d[0].jac, tb.d[1].jac, tb.d[2].jac, tb.d[0].wts, tb.d[1].wts, tb.d[2].wts, tb.d[0].pts,
                           tb.d[1].pts, tb.d[2].pts, dm, o, kq, dt, l2_part, dr_part);
    else
        hipLaunchKernelGGL((k_oil<P, false>), grid, dim3(kWave), 0, s, c, F, tb.d[0].tab, tb.d[1].tab, tb.d[2].tab,
                           tb.d[0].jac, tb.d[1].jac, tb.d[2].jac, tb.d[0].wts, tb.d[1].wts, tb.d[2].wts, tb.d[0].pts,
                           tb.d[1].pts, tb.d[2].pts, dm, o, kq, dt, l2_part, dr_part);
    return hipGetLastError();
}

// Output path (SaveSplinePlot + EvalSpline, src/plot.F90:97-145,
// src/basis.F90:323-369): the spline at every point of a tensor grid, one
// thread per point, from per-direction tables (first non-zero basis index
// and the p+1 values).  The sum runs in EvalSpline's order (x outer, z
// inner; b = bx*by*bz, then val += b*coeff) without fused multiply-adds.
template <int P>
__global__ __launch_bounds__(256) void k_eval_grid(const double* __restrict__ c, int sx, int sy, int sz,
                                                   const int* __restrict__ fx, const double* __restrict__ bx, int rx,
                                                   const int* __restrict__ fy, const double* __restrict__ by, int ry,
                                                   const int* __restrict__ fz, const double* __restrict__ bz, int rz,
                                                   double* __restrict__ out)
{
#pragma clang fp contract(off)
    const long i = blockIdx.x * static_cast<long>(blockDim.x) + threadIdx.x;
    const long npts = static_cast<long>(rx) * ry * rz;
    if (i >= npts) return;
    const int ix = static_cast<int>(i % rx);
    const int iy = static_cast<int>((i / rx) % ry);
    const int iz = static_cast<int>(i / (static_cast<long>(rx) * ry));
    const int x0 = fx[ix], y0 = fy[iy], z0 = fz[iz];
    double val = 0.0;
    // a direction of lower degree has zero values past its p + 1 entries;
    // its coefficient index is clamped there (weight 0 times a finite value)
    for (int a = 0; a < P; ++a)
        for (int b = 0; b < P; ++b)
            for (int d = 0; d < P; ++d) {
                const double w = bx[ix * P + a] * by[iy * P + b] * bz[iz * P + d];
                const int xi = min(x0 + a, sx - 1), yi = min(y0 + b, sy - 1), zi = min(z0 + d, sz - 1);
                const double cv = c[xi + static_cast<long>(sx) * (yi + static_cast<long>(sy) * zi)];
                val = val + w * cv;
            }
    out[i] = val;
}

}  // namespace

template <typename T>
hipError_t launch_assemble(int p, int form, int f0, const double* c, T* F, const Tables<T>& tb, double dt,
                           const double* kappa, double* l2_part, int part_cap, int* nparts, hipStream_t s)
{
    switch (p) {
    case 1: return assemble_p<T, 2>(form, f0, c, F, tb, dt, kappa, l2_part, part_cap, nparts, s);
    case 2: return assemble_p<T, 3>(form, f0, c, F, tb, dt, kappa, l2_part, part_cap, nparts, s);
    case 3: return assemble_p<T, 4>(form, f0, c, F, tb, dt, kappa, l2_part, part_cap, nparts, s);
    case 4: return assemble_p<T, 5>(form, f0, c, F, tb, dt, kappa, l2_part, part_cap, nparts, s);
    case 5: return assemble_p<T, 6>(form, f0, c, F, tb, dt, kappa, l2_part, part_cap, nparts, s);
    default: return hipErrorInvalidValue;
    }
}

template hipError_t launch_assemble<double>(int, int, int, const double*, double*, const Tables<double>&, double,
                                            const double*, double*, int, int*, hipStream_t);
template hipError_t launch_assemble<float>(int, int, int, const double*, float*, const Tables<float>&, double,
                                           const double*, double*, int, int*, hipStream_t);

hipError_t launch_eval_grid(int p, const double* c, int sx, int sy, int sz, const int* fx, const double* bx, int rx,
                            const int* fy, const double* by, int ry, const int* fz, const double* bz, int rz,
                            double* out, hipStream_t s)
{
    const long npts = static_cast<long>(rx) * ry * rz;
    const dim3 grid(static_cast<unsigned>((npts + 255) / 256));
    switch (p) {
    case 1: hipLaunchKernelGGL(k_eval_grid<2>, grid, dim3(256), 0, s, c, sx, sy, sz, fx, bx, rx, fy, by, ry, fz, bz, rz, out); break;
    case 2: hipLaunchKernelGGL(k_eval_grid<3>, grid, dim3(256), 0, s, c, sx, sy, sz, fx, bx, rx, fy, by, ry, fz, bz, rz, out); break;
    case 3: hipLaunchKernelGGL(k_eval_grid<4>, grid, dim3(256), 0, s, c, sx, sy, sz, fx, bx, rx, fy, by, ry, fz, bz, rz, out); break;
    case 4: hipLaunchKernelGGL(k_eval_grid<5>, grid, dim3(256), 0, s, c, sx, sy, sz, fx, bx, rx, fy, by, ry, fz, bz, rz, out); break;
    case 5: hipLaunchKernelGGL(k_eval_grid<6>, grid, dim3(256), 0, s, c, sx, sy, sz, fx, bx, rx, fy, by, ry, fz, bz, rz, out); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_reduce_partials(const double* part, int n, double* out, hipStream_t s)
{
    hipLaunchKernelGGL(k_reduce_partials, dim3(1), dim3(256), 0, s, part, n, out);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_heat_adjoint(int p, const double* mu, const double* c_n, T* lam_out, const Tables<T>& tb, double dt,
                               const double* kappa, double* dkappa, double* l2_part, int part_cap, int* nparts,
                               hipStream_t s)
{
    if (!dkappa)
        return launch_assemble<T>(p, kFormHeat, 0, mu, lam_out, tb, dt, kappa, l2_part, part_cap, nparts, s);
    switch (p) {
    case 1: return heat_sel<T, 2, kHeatGrad>(mu, lam_out, tb, dt, kappa, l2_part, part_cap, nparts, c_n, dkappa, s);
    case 2: return heat_sel<T, 3, kHeatGrad>(mu, lam_out, tb, dt, kappa, l2_part, part_cap, nparts, c_n, dkappa, s);
    case 3: return heat_sel<T, 4, kHeatGrad>(mu, lam_out, tb, dt, kappa, l2_part, part_cap, nparts, c_n, dkappa, s);
    case 4: return heat_sel<T, 5, kHeatGrad>(mu, lam_out, tb, dt, kappa, l2_part, part_cap, nparts, c_n, dkappa, s);
    case 5: return heat_sel<T, 6, kHeatGrad>(mu, lam_out, tb, dt, kappa, l2_part, part_cap, nparts, c_n, dkappa, s);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_heat_dual(int p, const double* c_val, const double* c_grad, double* F, const Tables<double>& tb,
                            double dt, const double* kappa, double* l2_part, int part_cap, int* nparts, hipStream_t s)
{
    switch (p) {
    case 1: return heat_sel<double, 2, kHeatDual>(c_val, F, tb, dt, kappa, l2_part, part_cap, nparts, c_grad, nullptr, s);
    case 2: return heat_sel<double, 3, kHeatDual>(c_val, F, tb, dt, kappa, l2_part, part_cap, nparts, c_grad, nullptr, s);
    case 3: return heat_sel<double, 4, kHeatDual>(c_val, F, tb, dt, kappa, l2_part, part_cap, nparts, c_grad, nullptr, s);
    case 4: return heat_sel<double, 5, kHeatDual>(c_val, F, tb, dt, kappa, l2_part, part_cap, nparts, c_grad, nullptr, s);
    case 5: return heat_sel<double, 6, kHeatDual>(c_val, F, tb, dt, kappa, l2_part, part_cap, nparts, c_grad, nullptr, s);
    default: return hipErrorInvalidValue;
    }
}

template hipError_t launch_heat_adjoint<double>(int, const double*, const double*, double*, const Tables<double>&,
                                                double, const double*, double*, double*, int, int*, hipStream_t);
template hipError_t launch_heat_adjoint<float>(int, const double*, const double*, float*, const Tables<float>&,
                                               double, const double*, double*, double*, int, int*, hipStream_t);

hipError_t launch_oil_kq(int p, const Tables<double>& tb, const OilDev& o, double* kq, hipStream_t s)
{
    const int nex = tb.d[0].nelem, ney = tb.d[1].nelem, nez = tb.d[2].nelem;
    const long q3 = static_cast<long>(p + 1) * (p + 1) * (p + 1);
    const long n = static_cast<long>(nex) * ney * nez * q3;
    const dim3 grid(static_cast<unsigned>((n + 255) / 256));
    const size_t list_bytes = sizeof(int) * static_cast<size_t>(o.cN) * (o.cL - 1);
    switch (p) {
    case 1: hipLaunchKernelGGL((k_oil_kq<2>), grid, dim3(256), list_bytes, s, tb.d[0].pts, tb.d[1].pts, tb.d[2].pts, nex, ney, nez, o, kq); break;
    case 2: hipLaunchKernelGGL((k_oil_kq<3>), grid, dim3(256), list_bytes, s, tb.d[0].pts, tb.d[1].pts, tb.d[2].pts, nex, ney, nez, o, kq); break;
    case 3: hipLaunchKernelGGL((k_oil_kq<4>), grid, dim3(256), list_bytes, s, tb.d[0].pts, tb.d[1].pts, tb.d[2].pts, nex, ney, nez, o, kq); break;
    case 4: hipLaunchKernelGGL((k_oil_kq<5>), grid, dim3(256), list_bytes, s, tb.d[0].pts, tb.d[1].pts, tb.d[2].pts, nex, ney, nez, o, kq); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_oil(int p, bool tzero, const double* c, double* F, const Tables<double>& tb, const OilDev& o,
                      const double* kq, double dt, double* l2_part, double* dr_part, int part_cap, int* nparts,
                      hipStream_t s)
{
    switch (p) {
    case 1: return oil_p<2>(tzero, c, F, tb, o, kq, dt, l2_part, dr_part, part_cap, nparts, s);
    case 2: return oil_p<3>(tzero, c, F, tb, o, kq, dt, l2_part, dr_part, part_cap, nparts, s);
    case 3: return oil_p<4>(tzero, c, F, tb, o, kq, dt, l2_part, dr_part, part_cap, nparts, s);
    case 4: return oil_p<5>(tzero, c, F, tb, o, kq, dt, l2_part, dr_part, part_cap, nparts, s);
    default: return hipErrorInvalidValue;
    }
}

}  // namespace ads
