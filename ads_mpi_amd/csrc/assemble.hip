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

template <typename T, int P, int MODE>
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
                        // lanes of one wave address distinct columns here; the
                        // LDS applies one wave's adds in program order
                        atomic_add(&fs[slot[azo]][wv + ayo][lane + axo], A(fma(s, ok, om)));
                    }
                    // keep the scheduler from computing every plane's values up front
                    __builtin_amdgcn_sched_barrier(0);
                }
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

template <typename T, int P, int MODE>
hipError_t heat_p(const double* c, T* F, const Tables<T>& tb, double dt, const double* kappa, double* l2_part,
                  int part_cap, int* nparts, const double* cn, double* dkappa, hipStream_t s)
{
    // two rounds of resident workgroups (a partly filled last round costs
    // as much as a full one)
    static const int resident =
        resident_groups(reinterpret_cast<const void*>(&k_heat<T, P, MODE>), kWave * kRowsWg, 0);
    AsmDims dm = base_dims(tb);
    dm.yblocks = (dm.ney + kRowsWg - 1) / kRowsWg;
    dm.zseg = segment_length(static_cast<long>(dm.xchunks) * dm.yblocks, dm.nez, 2L * resident);
    const int nseg = (dm.nez + dm.zseg - 1) / dm.zseg;
    dim3 grid(dm.xchunks * dm.yblocks * nseg);
    *nparts = static_cast<int>(grid.x);
    if (*nparts > part_cap) return hipErrorInvalidValue;
    hipLaunchKernelGGL((k_heat<T, P, MODE>), grid, dim3(kWave * kRowsWg), 0, s, c, F, tb.d[0].emat,
                       tb.d[1].emat_sym, tb.d[2].emat_sym, kappa, static_cast<T>(-dt), dm, l2_part, cn, dkappa);
    return hipGetLastError();
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
    return form == kFormHeat
               ? heat_p<T, P, kHeatPlain>(c, F, tb, dt, kappa, l2_part, part_cap, nparts, nullptr, nullptr, s)
               : project_p<T, P>(F, tb, f0, l2_part, part_cap, nparts, s);
}

// Output path (SaveSplinePlot + EvalSpline, src/plot.F90:97-145,
// src/basis.F90:323-369): the spline at every point of a tensor grid, one
// thread per point, from per-direction tables (first non-zero basis index
// and the p+1 values).  The sum runs in EvalSpline's order (x outer, z
// inner; b = bx*by*bz, then val += b*coeff) without fused multiply-adds.
template <int P>
__global__ __launch_bounds__(256) void k_eval_grid(const double* __restrict__ c, int sx, int sy,
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
    for (int a = 0; a < P; ++a)
        for (int b = 0; b < P; ++b)
            for (int d = 0; d < P; ++d) {
                const double w = bx[ix * P + a] * by[iy * P + b] * bz[iz * P + d];
                const double cv = c[(x0 + a) + static_cast<long>(sx) * ((y0 + b) + static_cast<long>(sy) * (z0 + d))];
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
    default: return hipErrorInvalidValue;
    }
}

template hipError_t launch_assemble<double>(int, int, int, const double*, double*, const Tables<double>&, double,
                                            const double*, double*, int, int*, hipStream_t);
template hipError_t launch_assemble<float>(int, int, int, const double*, float*, const Tables<float>&, double,
                                           const double*, double*, int, int*, hipStream_t);

hipError_t launch_eval_grid(int p, const double* c, int sx, int sy, const int* fx, const double* bx, int rx,
                            const int* fy, const double* by, int ry, const int* fz, const double* bz, int rz,
                            double* out, hipStream_t s)
{
    const long npts = static_cast<long>(rx) * ry * rz;
    const dim3 grid(static_cast<unsigned>((npts + 255) / 256));
    switch (p) {
    case 1: hipLaunchKernelGGL(k_eval_grid<2>, grid, dim3(256), 0, s, c, sx, sy, fx, bx, rx, fy, by, ry, fz, bz, rz, out); break;
    case 2: hipLaunchKernelGGL(k_eval_grid<3>, grid, dim3(256), 0, s, c, sx, sy, fx, bx, rx, fy, by, ry, fz, bz, rz, out); break;
    case 3: hipLaunchKernelGGL(k_eval_grid<4>, grid, dim3(256), 0, s, c, sx, sy, fx, bx, rx, fy, by, ry, fz, bz, rz, out); break;
    case 4: hipLaunchKernelGGL(k_eval_grid<5>, grid, dim3(256), 0, s, c, sx, sy, fx, bx, rx, fy, by, ry, fz, bz, rz, out); break;
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
    case 1: return heat_p<T, 2, kHeatGrad>(mu, lam_out, tb, dt, kappa, l2_part, part_cap, nparts, c_n, dkappa, s);
    case 2: return heat_p<T, 3, kHeatGrad>(mu, lam_out, tb, dt, kappa, l2_part, part_cap, nparts, c_n, dkappa, s);
    case 3: return heat_p<T, 4, kHeatGrad>(mu, lam_out, tb, dt, kappa, l2_part, part_cap, nparts, c_n, dkappa, s);
    case 4: return heat_p<T, 5, kHeatGrad>(mu, lam_out, tb, dt, kappa, l2_part, part_cap, nparts, c_n, dkappa, s);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_heat_dual(int p, const double* c_val, const double* c_grad, double* F, const Tables<double>& tb,
                            double dt, const double* kappa, double* l2_part, int part_cap, int* nparts, hipStream_t s)
{
    switch (p) {
    case 1: return heat_p<double, 2, kHeatDual>(c_val, F, tb, dt, kappa, l2_part, part_cap, nparts, c_grad, nullptr, s);
    case 2: return heat_p<double, 3, kHeatDual>(c_val, F, tb, dt, kappa, l2_part, part_cap, nparts, c_grad, nullptr, s);
    case 3: return heat_p<double, 4, kHeatDual>(c_val, F, tb, dt, kappa, l2_part, part_cap, nparts, c_grad, nullptr, s);
    case 4: return heat_p<double, 5, kHeatDual>(c_val, F, tb, dt, kappa, l2_part, part_cap, nparts, c_grad, nullptr, s);
    default: return hipErrorInvalidValue;
    }
}

template hipError_t launch_heat_adjoint<double>(int, const double*, const double*, double*, const Tables<double>&,
                                                double, const double*, double*, double*, int, int*, hipStream_t);
template hipError_t launch_heat_adjoint<float>(int, const double*, const double*, float*, const Tables<float>&,
                                               double, const double*, double*, double*, int, int*, hipStream_t);

}  // namespace ads
