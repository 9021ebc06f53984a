// assemble.hip -- gfx950 kernels of the RHS assembly (FormUn + Form3DRHS).
//
// k_assemble fuses FormUn (projection_engine.F90:306-441) and Form3DRHS
// (projection_engine.F90:154-284) for the heat and projection integrands
// (problems/heat/RHS_eq.F90:35-88, problems/L2Projection/RHS_eq.F90:34-71):
//
//   * a wave owns 64 consecutive elements along x at one ey and walks a
//     segment of ez; one lane per element;
//   * the coefficients of the P planes z = ez..ez+p the current element layer
//     touches sit in an LDS ring ("slab", (64+p) x P x P values), so each
//     coefficient is fetched from HBM once per wave, coalesced along x;
//   * values and gradients at the Q^3 Gauss points are sum-factorised one
//     direction at a time (x stage: per-lane tables; y/z stages: wave-uniform
//     tables in SGPRs), the integrand is applied pointwise and the transposed
//     contractions accumulate into an LDS ring of right-hand-side planes;
//   * a right-hand-side plane is complete once the element layer below it is
//     done; it is then added to F with one coalesced atomic per entry (other
//     waves own the neighbouring y rows and z segments).
//
// k_kappa_grad gives the adjoint's per-element diffusivity gradient.
#include "kernels.hpp"

#include <hip/hip_runtime.h>

namespace ads {

namespace {

constexpr int kWave = 64;

#ifndef ADS_QY_UNROLL
#define ADS_QY_UNROLL 1
#endif
#ifndef ADS_NY_PREFETCH
#define ADS_NY_PREFETCH 0
#endif

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
    int zseg;           // element layers per wave
    int f0;             // initial-state kind (projection form)
    int xchunks;        // waves along x
};

// MI355X hands consecutive workgroups to its 8 XCDs round robin.  Remapping
// the 1-D block index gives XCD k a contiguous run of (x chunk, ey, z
// segment) tiles, so the waves that add into the same right-hand-side rows
// (neighbouring ey) share one L2.  The map is a bijection for any count.
constexpr int kXcds = 8;
__device__ __forceinline__ int xcd_tile(int b, int nb)
{
    const int q = nb / kXcds, r = nb % kXcds, k = b % kXcds, s = b / kXcds;
    return k * q + min(k, r) + s;
}

// Every table pointer is a separate __restrict__ argument so the compiler
// may read the wave-uniform y/z tables with scalar loads into SGPRs.
template <typename T, int P, int FORM>
__global__ __launch_bounds__(kWave) void k_assemble(const double* __restrict__ c, T* __restrict__ F,
                                                    const T* __restrict__ tabx, const T* __restrict__ taby,
                                                    const T* __restrict__ tabz, const T* __restrict__ jacx,
                                                    const T* __restrict__ jacy, const T* __restrict__ jacz,
                                                    const T* __restrict__ Wx, const T* __restrict__ Wy,
                                                    const T* __restrict__ Wz, const T* __restrict__ ptsx,
                                                    const T* __restrict__ ptsy, const T* __restrict__ ptsz,
                                                    const double* __restrict__ kappa, T ndt, AsmDims dm,
                                                    double* __restrict__ l2_part)
{
    constexpr int Q = P, KL = P - 1, W = kWave + KL;
    __shared__ T cs[FORM == kFormHeat ? P : 1][P][W];  // coefficient planes, ring by z
    __shared__ T fs[P][P][W];                           // right-hand-side planes, ring by z

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
    const int oy = ey;  // first DOF of element e is e (open knots, checked on the host)
    const int ncol = min(W, sx - ex0);

    const T* __restrict__ Nx = tabx + static_cast<size_t>(exc) * Q * P * 2;
    const T* __restrict__ Ny = taby + static_cast<size_t>(ey) * Q * P * 2;
    const T Jxy = jacx[exc] * jacy[ey];

    for (int i = lane; i < P * P * W; i += kWave) (&fs[0][0][0])[i] = T(0);

    auto load_plane = [&](int z) {
        const int slot = z % P;
        const double* src = c + ex0 + static_cast<long>(sx) * oy + plane * z;
        for (int i = lane; i < P * W; i += kWave) {
            const int ay = i / W, col = i - ay * W;
            cs[slot][ay][col] = col < ncol ? T(src[col + static_cast<long>(sx) * ay]) : T(0);
        }
    };
    auto flush_plane = [&](int z) {
        const int slot = z % P;
        T* dst = F + ex0 + static_cast<long>(sx) * oy + plane * z;
        for (int i = lane; i < P * W; i += kWave) {
            const int ay = i / W, col = i - ay * W;
            if (col < ncol) atomic_add(&dst[col + static_cast<long>(sx) * ay], fs[slot][ay][col]);
            fs[slot][ay][col] = T(0);
        }
    };

    if constexpr (FORM == kFormHeat)
        for (int z = ez0; z < ez0 + KL; ++z) load_plane(z);
    __syncthreads();

    T l2 = T(0);
    for (int ez = ez0; ez < ez1; ++ez) {
        if constexpr (FORM == kFormHeat) {
            load_plane(ez + KL);
            __syncthreads();
        }
        const T* __restrict__ Nz = tabz + static_cast<size_t>(ez) * Q * P * 2;
        const T J = Jxy * jacz[ez];
        T gk0 = T(0);
        if constexpr (FORM == kFormHeat) {
            const T kap =
                kappa ? T(kappa[exc + static_cast<size_t>(dm.nex) * (ey + static_cast<size_t>(dm.ney) * ez)]) : T(1);
            gk0 = ndt * kap * J;
        }
        const int s0 = ez % P;
        int slot[P];
#pragma unroll
        for (int az = 0; az < P; ++az) slot[az] = (s0 + az) % P;

#pragma unroll 1
        for (int qx = 0; qx < Q; ++qx) {
            T n0x[P], n1x[P];
#pragma unroll
            for (int a = 0; a < P; ++a) {
                n0x[a] = Nx[(qx * P + a) * 2];
                n1x[a] = Nx[(qx * P + a) * 2 + 1];
            }
            // x stage: sx0 = N0x c, sx1 = N1x c   [ay][az]
            T sx0[P][P], sx1[P][P];
            if constexpr (FORM == kFormHeat) {
#pragma unroll
                for (int ay = 0; ay < P; ++ay)
#pragma unroll
                    for (int az = 0; az < P; ++az) {
                        const T* col = &cs[slot[az]][ay][lane];
                        T a0 = n0x[0] * col[0];
                        T a1 = n1x[0] * col[0];
#pragma unroll
                        for (int ax = 1; ax < P; ++ax) {
                            a0 = fma(n0x[ax], col[ax], a0);
                            a1 = fma(n1x[ax], col[ax], a1);
                        }
                        sx0[ay][az] = a0;
                        sx1[ay][az] = a1;
                    }
            }
            T t0[P][P], t1[P][P];
#pragma unroll
            for (int ay = 0; ay < P; ++ay)
#pragma unroll
                for (int az = 0; az < P; ++az) t0[ay][az] = t1[ay][az] = T(0);
            const T xq = (FORM == kFormProject) ? ptsx[static_cast<size_t>(exc) * Q + qx] : T(0);
            const T wxq = Wx[qx];
#if ADS_NY_PREFETCH
            T nyn0[P], nyn1[P];
#pragma unroll
            for (int a = 0; a < P; ++a) {
                nyn0[a] = Ny[a * 2];
                nyn1[a] = Ny[a * 2 + 1];
            }
#endif
#pragma unroll ADS_QY_UNROLL
            for (int qy = 0; qy < Q; ++qy) {
                T ny0[P], ny1[P];
#if ADS_NY_PREFETCH
                // this iteration's y tables were fetched one iteration ahead
#pragma unroll
                for (int a = 0; a < P; ++a) {
                    ny0[a] = nyn0[a];
                    ny1[a] = nyn1[a];
                }
                {
                    const int qn = qy + 1 < Q ? qy + 1 : qy;
#pragma unroll
                    for (int a = 0; a < P; ++a) {
                        nyn0[a] = Ny[(qn * P + a) * 2];
                        nyn1[a] = Ny[(qn * P + a) * 2 + 1];
                    }
                }
#else
#pragma unroll
                for (int a = 0; a < P; ++a) {
                    ny0[a] = Ny[(qy * P + a) * 2];
                    ny1[a] = Ny[(qy * P + a) * 2 + 1];
                }
#endif
                T r00[P], r01[P], r10[P];
                if constexpr (FORM == kFormHeat) {
#pragma unroll
                    for (int az = 0; az < P; ++az) {
                        T a = ny0[0] * sx0[0][az];
                        T b = ny1[0] * sx0[0][az];
                        T d = ny0[0] * sx1[0][az];
#pragma unroll
                        for (int ay = 1; ay < P; ++ay) {
                            a = fma(ny0[ay], sx0[ay][az], a);
                            b = fma(ny1[ay], sx0[ay][az], b);
                            d = fma(ny0[ay], sx1[ay][az], d);
                        }
                        r00[az] = a;
                        r01[az] = b;
                        r10[az] = d;
                    }
                }
                T h00[P], h01[P], h10[P];
#pragma unroll
                for (int az = 0; az < P; ++az) h00[az] = h01[az] = h10[az] = T(0);
                const T yq = (FORM == kFormProject) ? ptsy[static_cast<size_t>(ey) * Q + qy] : T(0);
                const T wxy = wxq * Wy[qy];
#pragma unroll
                for (int qz = 0; qz < Q; ++qz) {
                    T nz0[P], nz1[P];
#pragma unroll
                    for (int a = 0; a < P; ++a) {
                        nz0[a] = Nz[(qz * P + a) * 2];
                        nz1[a] = Nz[(qz * P + a) * 2 + 1];
                    }
                    const T wq = wxy * Wz[qz];
                    const T w = J * wq;
                    if constexpr (FORM == kFormHeat) {
                        T u = nz0[0] * r00[0];
                        T uz = nz1[0] * r00[0];
                        T uy = nz0[0] * r01[0];
                        T ux = nz0[0] * r10[0];
#pragma unroll
                        for (int az = 1; az < P; ++az) {
                            u = fma(nz0[az], r00[az], u);
                            uz = fma(nz1[az], r00[az], uz);
                            uy = fma(nz0[az], r01[az], uy);
                            ux = fma(nz0[az], r10[az], ux);
                        }
                        const T g0 = w * u;
                        l2 = fma(g0, u, l2);
                        const T gk = gk0 * wq;
                        const T gx = gk * ux, gy = gk * uy, gz = gk * uz;
#pragma unroll
                        for (int az = 0; az < P; ++az) {
                            h00[az] = fma(nz0[az], g0, fma(nz1[az], gz, h00[az]));
                            h01[az] = fma(nz0[az], gy, h01[az]);
                            h10[az] = fma(nz0[az], gx, h10[az]);
                        }
                    } else {
                        const T zq = ptsz[static_cast<size_t>(ez) * Q + qz];
                        const T fv = f0_eval<T>(dm.f0, xq, yq, zq);
                        const T g0 = w * fv;
                        l2 = fma(g0, fv, l2);
#pragma unroll
                        for (int az = 0; az < P; ++az) h00[az] = fma(nz0[az], g0, h00[az]);
                    }
                }
#pragma unroll
                for (int ay = 0; ay < P; ++ay)
#pragma unroll
                    for (int az = 0; az < P; ++az) {
                        if constexpr (FORM == kFormHeat) {
                            t0[ay][az] = fma(ny0[ay], h00[az], fma(ny1[ay], h01[az], t0[ay][az]));
                            t1[ay][az] = fma(ny0[ay], h10[az], t1[ay][az]);
                        } else {
                            t0[ay][az] = fma(ny0[ay], h00[az], t0[ay][az]);
                        }
                    }
            }
            // transposed x stage into the right-hand-side ring: LDS atomic adds
            // (lanes of one wave address overlapping columns for different ax;
            // the LDS applies one wave's adds in program order)
            if (active) {
#pragma unroll
                for (int ax = 0; ax < P; ++ax) {
#pragma unroll
                    for (int ay = 0; ay < P; ++ay)
#pragma unroll
                        for (int az = 0; az < P; ++az) {
                            T v;
                            if constexpr (FORM == kFormHeat) v = fma(n0x[ax], t0[ay][az], n1x[ax] * t1[ay][az]);
                            else v = n0x[ax] * t0[ay][az];
                            atomic_add(&fs[slot[az]][ay][lane + ax], v);
                        }
                    // keep the scheduler from computing every group's values up front
                    __builtin_amdgcn_sched_barrier(0);
                }
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

// Per-element diffusivity gradient of the adjoint:
//   dkappa[e] += -dt * mu_e^T K_e c_e,
// K_e = Kx(x)My(x)Mz + Mx(x)Ky(x)Mz + Mx(x)My(x)Kz with the element's 1-D
// matrices from the same basis tables and Gauss weights (exact for degree-2p
// integrands), applied sum-factorised and dotted with mu.
template <int P>
__global__ __launch_bounds__(kWave) void k_kappa_grad(const double* __restrict__ c, const double* __restrict__ mu,
                                                      Tables<double> tb, double ndt, double* __restrict__ dkappa)
{
    constexpr int Q = P;
    const int lane = threadIdx.x;
    const int ex = blockIdx.x * kWave + lane;
    const int ey = blockIdx.y, ez = blockIdx.z;
    const int nex = tb.d[0].nelem;
    if (ex >= nex) return;
    const int sx = tb.d[0].dofs, sy = tb.d[1].dofs;
    const int ox = tb.d[0].first_dof[ex], oy = tb.d[1].first_dof[ey], oz = tb.d[2].first_dof[ez];
    double Mx[P][P], Kx[P][P], My[P][P], Ky[P][P], Mz[P][P], Kz[P][P];
    auto elem = [&](int dd, int e, double (*M)[P], double (*K)[P]) {
        const double* N = tb.d[dd].tab + static_cast<size_t>(e) * Q * P * 2;
        const double Jd = tb.d[dd].jac[e];
#pragma unroll
        for (int a = 0; a < P; ++a)
#pragma unroll
            for (int b = 0; b < P; ++b) {
                double m = 0.0, k = 0.0;
#pragma unroll
                for (int q = 0; q < Q; ++q) {
                    const double jw = Jd * tb.d[dd].wts[q];
                    m += N[(q * P + a) * 2] * N[(q * P + b) * 2] * jw;
                    k += N[(q * P + a) * 2 + 1] * N[(q * P + b) * 2 + 1] * jw;
                }
                M[a][b] = m;
                K[a][b] = k;
            }
    };
    elem(0, ex, Mx, Kx);
    elem(1, ey, My, Ky);
    elem(2, ez, Mz, Kz);
    const size_t base = static_cast<size_t>(ox) + static_cast<size_t>(sx) * (oy + static_cast<size_t>(sy) * oz);
    double cc[P][P][P];
#pragma unroll
    for (int az = 0; az < P; ++az)
#pragma unroll
        for (int ay = 0; ay < P; ++ay)
#pragma unroll
            for (int ax = 0; ax < P; ++ax)
                cc[ax][ay][az] = c[base + ax + static_cast<size_t>(sx) * (ay + static_cast<size_t>(sy) * az)];
    double acc = 0.0;
#pragma unroll 1
    for (int a = 0; a < P; ++a) {
        double kxc[P][P], mxc[P][P];
#pragma unroll
        for (int ay = 0; ay < P; ++ay)
#pragma unroll
            for (int az = 0; az < P; ++az) {
                double s = 0.0, t = 0.0;
#pragma unroll
                for (int b = 0; b < P; ++b) {
                    s += Kx[a][b] * cc[b][ay][az];
                    t += Mx[a][b] * cc[b][ay][az];
                }
                kxc[ay][az] = s;
                mxc[ay][az] = t;
            }
#pragma unroll
        for (int by = 0; by < P; ++by) {
            double u1[P], u2[P];
#pragma unroll
            for (int az = 0; az < P; ++az) {
                double s = 0.0, t = 0.0;
#pragma unroll
                for (int ay = 0; ay < P; ++ay) {
                    s += My[by][ay] * kxc[ay][az] + Ky[by][ay] * mxc[ay][az];
                    t += My[by][ay] * mxc[ay][az];
                }
                u1[az] = s;
                u2[az] = t;
            }
#pragma unroll
            for (int bz = 0; bz < P; ++bz) {
                double s = 0.0;
#pragma unroll
                for (int az = 0; az < P; ++az) s += Mz[bz][az] * u1[az] + Kz[bz][az] * u2[az];
                acc += mu[base + a + static_cast<size_t>(sx) * (by + static_cast<size_t>(sy) * bz)] * s;
            }
        }
    }
    dkappa[ex + static_cast<size_t>(nex) * (ey + static_cast<size_t>(tb.d[1].nelem) * ez)] += ndt * acc;
}

// z-segment length: enough waves to fill 256 CUs about eight times over.
int segment_length(int xchunks, int ney, int nez)
{
    const long target = 2048;
    long nseg = (target + static_cast<long>(xchunks) * ney - 1) / (static_cast<long>(xchunks) * ney);
    if (nseg < 1) nseg = 1;
    if (nseg > nez) nseg = nez;
    return static_cast<int>((nez + nseg - 1) / nseg);
}

template <typename T, int P, int FORM>
hipError_t assemble_p(const double* c, T* F, const Tables<T>& tb, double dt, const double* kappa, int f0,
                      double* l2_part, int* nparts, hipStream_t s)
{
    const int xchunks = (tb.d[0].nelem + kWave - 1) / kWave;
    AsmDims dm;
    dm.nex = tb.d[0].nelem;
    dm.ney = tb.d[1].nelem;
    dm.nez = tb.d[2].nelem;
    dm.sx = tb.d[0].dofs;
    dm.sy = tb.d[1].dofs;
    dm.zseg = segment_length(xchunks, dm.ney, dm.nez);
    dm.f0 = f0;
    dm.xchunks = xchunks;
    dim3 grid(xchunks * dm.ney * ((dm.nez + dm.zseg - 1) / dm.zseg));
    *nparts = static_cast<int>(grid.x);
    hipLaunchKernelGGL((k_assemble<T, P, FORM>), grid, dim3(kWave), 0, s, c, F, tb.d[0].tab, tb.d[1].tab,
                       tb.d[2].tab, tb.d[0].jac, tb.d[1].jac, tb.d[2].jac, tb.d[0].wts, tb.d[1].wts, tb.d[2].wts,
                       tb.d[0].pts, tb.d[1].pts, tb.d[2].pts, kappa, static_cast<T>(-dt), dm, l2_part);
    return hipGetLastError();
}

}  // namespace

template <typename T>
hipError_t launch_assemble(int p, int form, int f0, const double* c, T* F, const Tables<T>& tb, double dt,
                           const double* kappa, double* l2_part, int* nparts, hipStream_t s)
{
    switch (p) {
    case 1:
        return form == kFormHeat ? assemble_p<T, 2, kFormHeat>(c, F, tb, dt, kappa, f0, l2_part, nparts, s)
                                 : assemble_p<T, 2, kFormProject>(c, F, tb, dt, kappa, f0, l2_part, nparts, s);
    case 2:
        return form == kFormHeat ? assemble_p<T, 3, kFormHeat>(c, F, tb, dt, kappa, f0, l2_part, nparts, s)
                                 : assemble_p<T, 3, kFormProject>(c, F, tb, dt, kappa, f0, l2_part, nparts, s);
    case 3:
        return form == kFormHeat ? assemble_p<T, 4, kFormHeat>(c, F, tb, dt, kappa, f0, l2_part, nparts, s)
                                 : assemble_p<T, 4, kFormProject>(c, F, tb, dt, kappa, f0, l2_part, nparts, s);
    case 4:
        return form == kFormHeat ? assemble_p<T, 5, kFormHeat>(c, F, tb, dt, kappa, f0, l2_part, nparts, s)
                                 : assemble_p<T, 5, kFormProject>(c, F, tb, dt, kappa, f0, l2_part, nparts, s);
    default:
        return hipErrorInvalidValue;
    }
}

template hipError_t launch_assemble<double>(int, int, int, const double*, double*, const Tables<double>&, double,
                                            const double*, double*, int*, hipStream_t);
template hipError_t launch_assemble<float>(int, int, int, const double*, float*, const Tables<float>&, double,
                                           const double*, double*, int*, hipStream_t);

hipError_t launch_reduce_partials(const double* part, int n, double* out, hipStream_t s)
{
    hipLaunchKernelGGL(k_reduce_partials, dim3(1), dim3(256), 0, s, part, n, out);
    return hipGetLastError();
}

hipError_t launch_assemble_adjoint(int p, const double* c_n, const double* mu, double* lam_out,
                                   const Tables<double>& tb, double dt, const double* kappa, double* dkappa,
                                   hipStream_t s)
{
    (void)lam_out;
    (void)kappa;
    if (!dkappa) return hipSuccess;
    dim3 grid((tb.d[0].nelem + kWave - 1) / kWave, tb.d[1].nelem, tb.d[2].nelem);
    switch (p) {
    case 1: hipLaunchKernelGGL((k_kappa_grad<2>), grid, dim3(kWave), 0, s, c_n, mu, tb, -dt, dkappa); break;
    case 2: hipLaunchKernelGGL((k_kappa_grad<3>), grid, dim3(kWave), 0, s, c_n, mu, tb, -dt, dkappa); break;
    case 3: hipLaunchKernelGGL((k_kappa_grad<4>), grid, dim3(kWave), 0, s, c_n, mu, tb, -dt, dkappa); break;
    case 4: hipLaunchKernelGGL((k_kappa_grad<5>), grid, dim3(kWave), 0, s, c_n, mu, tb, -dt, dkappa); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace ads
