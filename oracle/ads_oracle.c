/*
 * ads_oracle.c -- CPU restatement of the IGA-ADS heat-step hot path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP path, never part of it.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load
 * this library.  Built with -ffp-contract=off so every operation rounds the
 * way the reference's source order says.
 *
 * What pins each function (tests named in tests/test_oracle_*.py):
 *   orc_fill_open_knot / orc_count_spans / orc_prepare_knot
 *       tests/test_knot.pf vectors (knot1, knot2, knot3) + the flang-built
 *       reference leaves in oracle/_ref (bitwise) + closed form (float)k/(float)ne.
 *   orc_gauss_rule      closed forms + numpy leggauss (1e-15) + oracle/_ref.
 *   orc_ders_basis_funs / orc_basis_data / orc_find_span / orc_eval_spline
 *       oracle/_ref (bitwise), scipy.interpolate.BSpline, partition of unity,
 *       brute-force Cox-de Boor.
 *   orc_mkbbt_band      interior Gram rows (cardinal B-spline values) and
 *       scipy BSpline + leggauss brute force; stiffness row sums = 0.
 *   orc_dgbtrf / orc_dgbtrs   scipy.linalg.lapack dgbtrf/dgbtrs golden
 *       vectors (tests/golden) + dense numpy solve.
 *   orc_reorder_*       oracle/_ref (flang-built reorderRHS.F90), bitwise.
 *   orc_form_rhs_*      L2Projection known answer (f = 1 -> all ones,
 *       problems/L2Projection/main.F90:60-81), Kronecker identity
 *       F = (Mz(x)My(x)Mx - dt(...))c with numpy, faithful == sumfact.
 *   orc_solve_kron      dense Kronecker solve with numpy at small sizes.
 *   orc_heat_step_adjoint   central finite differences on the oracle forward.
 *   orc_uniform / orc_lognormal   the product library's generator, bitwise
 *       (same definition, independent code), and moment checks.
 * No function is "parity unpinned"; the one unpinnable item is MUMPS itself
 * (not in the image), replaced by LAPACK-form band LU as documented.
 */
#include "ads_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_MAXP 9

/* ======================================================================
 * knots -- knot_vector.F90
 * ====================================================================== */

/* FillOpenKnot (knot_vector.F90:79-92).  U(1:n+p+2) in the reference is
 * U[0..n+p+1] here.  Zeros first, then ones (which overwrite for n < p),
 * then the interior knots as a REAL(4) quotient widened to REAL(8) (:89). */
void orc_fill_open_knot(int n, int p, double* U)
{
    for (int i = 1; i <= p + 1; ++i) U[i - 1] = 0.0;
    for (int i = n + 2; i <= n + p + 2; ++i) U[i - 1] = 1.0;
    for (int i = p + 2; i <= n + 1; ++i) {
        float num = (float)(i - p - 1);
        float den = (float)(n - p + 1);
        float q = num / den;
        U[i - 1] = (double)q;
    }
}

/* CountSpans (knot_vector.F90:110-130), U(0:n+p+1). */
int orc_count_spans(int n, int p, const double* U)
{
    int nelem = 0;
    int i = p;
    while (i <= n) {
        while (i < n && U[i] == U[i + 1]) ++i;
        ++nelem;
        ++i;
    }
    return nelem;
}

/* PrepareKnot (knot_vector.F90:43-61); U must hold n+p+2 values. */
int orc_prepare_knot(int n, int p, double* U, int* nelem)
{
    if (n + p + 2 <= 0) return -1;
    orc_fill_open_knot(n, p, U);
    *nelem = orc_count_spans(n, p, U);
    return 0;
}

/* ======================================================================
 * Gauss-Legendre -- gauss.F90
 * ====================================================================== */

/* Nodes/weights for q >= 6 by Newton iteration on P_q (Legendre three-term
 * recurrence); the reference stores 16-digit literals for q = 6..10
 * (gauss.F90:109-192), which agree with these to ~1e-16. */
static void gauss_newton(int q, double* X, double* W)
{
    for (int i = 0; i < q; ++i) {
        double x = -cos(M_PI * (i + 0.75) / (q + 0.5));
        for (int it = 0; it < 100; ++it) {
            double p0 = 1.0, p1 = x;
            for (int k = 2; k <= q; ++k) {
                double p2 = ((2.0 * k - 1.0) * x * p1 - (k - 1.0) * p0) / k;
                p0 = p1;
                p1 = p2;
            }
            double dp = q * (x * p1 - p0) / (x * x - 1.0);
            double dx = p1 / dp;
            x -= dx;
            if (fabs(dx) < 1e-17) break;
        }
        double p0 = 1.0, p1 = x;
        for (int k = 2; k <= q; ++k) {
            double p2 = ((2.0 * k - 1.0) * x * p1 - (k - 1.0) * p0) / k;
            p0 = p1;
            p1 = p2;
        }
        double dp = q * (x * p1 - p0) / (x * x - 1.0);
        X[i] = x;
        W[i] = 2.0 / ((1.0 - x * x) * dp * dp);
    }
    /* symmetric placement as in the reference (X(q-1-i) = -X(i)) */
    for (int i = 0; i < q / 2; ++i) {
        X[q - 1 - i] = -X[i];
        W[q - 1 - i] = W[i];
    }
    if (q % 2) X[q / 2] = 0.0;
}

/* GaussRule (gauss.F90:213-257) with the closed forms of initialize
 * (gauss.F90:71-107).  Returns -1 for q outside 1..10, where the reference
 * silently returns zeros (gauss.F90:252-254). */
int orc_gauss_rule(int q, double* X, double* W)
{
    switch (q) {
    case 1:
        X[0] = 0.0; W[0] = 2.0;
        return 0;
    case 2:
        X[0] = -1.0 / sqrt(3.0); X[1] = -X[0];
        W[0] = 1.0; W[1] = W[0];
        return 0;
    case 3:
        X[0] = -sqrt(3.0 / 5.0); X[1] = 0.0; X[2] = -X[0];
        W[0] = 5.0 / 9.0; W[1] = 8.0 / 9.0; W[2] = W[0];
        return 0;
    case 4: {
        double s = 2.0 * sqrt(6.0 / 5.0);
        X[0] = -sqrt((3.0 + s) / 7.0);
        X[1] = -sqrt((3.0 - s) / 7.0);
        X[2] = -X[1]; X[3] = -X[0];
        W[0] = (18.0 - sqrt(30.0)) / 36.0;
        W[1] = (18.0 + sqrt(30.0)) / 36.0;
        W[2] = W[1]; W[3] = W[0];
        return 0;
    }
    case 5: {
        double s = 2.0 * sqrt(10.0 / 7.0);
        X[0] = -1.0 / 3.0 * sqrt(5.0 + s);
        X[1] = -1.0 / 3.0 * sqrt(5.0 - s);
        X[2] = 0.0; X[3] = -X[1]; X[4] = -X[0];
        W[0] = (322.0 - 13.0 * sqrt(70.0)) / 900.0;
        W[1] = (322.0 + 13.0 * sqrt(70.0)) / 900.0;
        W[2] = 128.0 / 225.0; W[3] = W[1]; W[4] = W[0];
        return 0;
    }
    default:
        if (q >= 6 && q <= 10) {
            gauss_newton(q, X, W);
            return 0;
        }
        for (int i = 0; i < q; ++i) { X[i] = 0.0; W[i] = 0.0; }
        return -1;
    }
}

/* ======================================================================
 * B-spline basis -- basis.F90
 * ====================================================================== */

/* DersBasisFuns (basis.F90:136-202): Piegl & Tiller, The NURBS Book, A2.3.
 * ders(0:p, 0:d) column-major -> ders[a + (p+1)*k]. */
void orc_ders_basis_funs(int i, double uu, int p, int d, const double* U, double* ders)
{
    double left[ORC_MAXP + 2], right[ORC_MAXP + 2];
    double ndu[ORC_MAXP + 1][ORC_MAXP + 1]; /* ndu[j][r] == Fortran ndu(j, r) */
    double a[2][ORC_MAXP + 1];

    ndu[0][0] = 1.0;
    for (int j = 1; j <= p; ++j) {
        left[j] = uu - U[i + 1 - j];
        right[j] = U[i + j] - uu;
        double saved = 0.0;
        for (int r = 0; r <= j - 1; ++r) {
            ndu[j][r] = right[r + 1] + left[j - r];
            double temp = ndu[r][j - 1] / ndu[j][r];
            ndu[r][j] = saved + right[r + 1] * temp;
            saved = left[j - r] * temp;
        }
        ndu[j][j] = saved;
    }
    for (int r = 0; r <= p; ++r) ders[r] = ndu[r][p];

    for (int r = 0; r <= p; ++r) {
        int s1 = 0, s2 = 1;
        a[0][0] = 1.0;
        for (int k = 1; k <= d; ++k) {
            double der = 0.0;
            int rk = r - k, pk = p - k;
            if (r >= k) {
                a[s2][0] = a[s1][0] / ndu[pk + 1][rk];
                der = a[s2][0] * ndu[rk][pk];
            }
            int j1 = (rk > -1) ? 1 : -rk;
            int j2 = (r - 1 <= pk) ? k - 1 : p - r;
            for (int j = j1; j <= j2; ++j) {
                a[s2][j] = (a[s1][j] - a[s1][j - 1]) / ndu[pk + 1][rk + j];
                der = der + a[s2][j] * ndu[rk + j][pk];
            }
            if (r <= pk) {
                a[s2][k] = -a[s1][k - 1] / ndu[pk + 1][r];
                der = der + a[s2][k] * ndu[r][pk];
            }
            ders[r + (p + 1) * k] = der;
            int t = s1; s1 = s2; s2 = t;
        }
    }
    int rr = p;
    for (int k = 1; k <= d; ++k) {
        for (int j = 0; j <= p; ++j) ders[j + (p + 1) * k] = ders[j + (p + 1) * k] * rr;
        rr = rr * (p - k);
    }
}

/* FindSpan (basis.F90:226-259): Piegl & Tiller A2.1. */
int orc_find_span(int n, int p, double uu, const double* U)
{
    if (uu >= U[n + 1]) return n;
    if (uu <= U[p]) return p;
    int low = p, high = n + 1;
    int span = (low + high) / 2;
    while (uu < U[span] || uu >= U[span + 1]) {
        if (uu < U[span]) high = span;
        else low = span;
        span = (low + high) / 2;
    }
    return span;
}

/* BasisData (basis.F90:74-112).  O is 0-based (first global DOF of each
 * element), as in the reference. */
int orc_basis_data(int p, int m, const double* U, int d, int q, int r,
                   int* O, double* J, double* W, double* X, double* NN)
{
    double Xg[16];
    double ders[(ORC_MAXP + 1) * 2];
    if (p > ORC_MAXP || q > 10 || d > 1) return -1;
    int ir = 0;
    for (int i = p; i <= m - p && i + 1 <= m; ++i) {
        if (U[i] != U[i + 1]) {
            if (ir < r) O[ir] = i - p;
            ++ir;
        }
    }
    if (ir != r) return -2;
    if (orc_gauss_rule(q, Xg, W) != 0) return -3;
    for (ir = 0; ir < r; ++ir) {
        int i = O[ir] + p;
        J[ir] = (U[i + 1] - U[i]) / 2.0;
        for (int iq = 0; iq < q; ++iq) X[iq + q * ir] = (Xg[iq] + 1.0) * J[ir] + U[i];
        for (int iq = 0; iq < q; ++iq) {
            double uu = X[iq + q * ir];
            orc_ders_basis_funs(i, uu, p, d, U, ders);
            for (int a = 0; a <= p; ++a)
                for (int k = 0; k <= d; ++k)
                    NN[k + (d + 1) * (a + (p + 1) * (iq + q * ir))] = ders[a + (p + 1) * k];
        }
    }
    return 0;
}

/* EvalSpline (basis.F90:323-369); coeffs(0:nx, 0:ny, 0:nz) x fastest. */
double orc_eval_spline(int d,
                       const double* Ux, int px, int nx,
                       const double* Uy, int py, int ny,
                       const double* Uz, int pz, int nz,
                       const double* coeffs, double x, double y, double z)
{
    double bx[(ORC_MAXP + 1) * 2], by[(ORC_MAXP + 1) * 2], bz[(ORC_MAXP + 1) * 2];
    int xs = orc_find_span(nx, px, x, Ux);
    int ys = orc_find_span(ny, py, y, Uy);
    int zs = orc_find_span(nz, pz, z, Uz);
    /* the reference asks for derivative order 0 and then reads column d
     * (basis.F90:350-363); only d = 0 is meaningful there. */
    orc_ders_basis_funs(xs, x, px, d, Ux, bx);
    orc_ders_basis_funs(ys, y, py, d, Uy, by);
    orc_ders_basis_funs(zs, z, pz, d, Uz, bz);
    int x0 = xs - px, y0 = ys - py, z0 = zs - pz;
    double val = 0.0;
    for (int ix = 0; ix <= px; ++ix)
        for (int iy = 0; iy <= py; ++iy)
            for (int iz = 0; iz <= pz; ++iz) {
                double b = bx[ix + (px + 1) * d] * by[iy + (py + 1) * d] * bz[iz + (pz + 1) * d];
                val = val + b * coeffs[(x0 + ix) + (nx + 1) * ((y0 + iy) + (ny + 1) * (z0 + iz))];
            }
    return val;
}

/* ======================================================================
 * 1-D matrices -- MKBBT (projection_engine.F90:63-132) into band storage
 * ====================================================================== */

/* Entries accumulate in the reference's loop order (d fastest, then c, then
 * quadrature point, then element; :103-129) into A(ia, ib), ia = O(e)+c row,
 * ib = O(e)+d column (sparse.F90:176-186 adds val to entry (x=ia, y=ib)).
 * The band layout is LAPACK's dgbtrf input (kl = ku = p, ldab >= 3p+1). */
int orc_mkbbt_band(const double* U, int p, int n, int nelem, const double mix[4],
                   double* ab, int ldab)
{
    int q = p + 1, m = n + p + 1;
    int kl = p, ku = p;
    if (ldab < 2 * kl + ku + 1) return -1;
    int* O = (int*)malloc(sizeof(int) * nelem);
    double* J = (double*)malloc(sizeof(double) * nelem);
    double W[16];
    double* X = (double*)malloc(sizeof(double) * q * nelem);
    double* NN = (double*)malloc(sizeof(double) * 2 * (p + 1) * q * nelem);
    int rc = orc_basis_data(p, m, U, 1, q, nelem, O, J, W, X, NN);
    if (rc == 0) {
        memset(ab, 0, sizeof(double) * (size_t)ldab * (n + 1));
#define NNV(k, a, iq, e) NN[(k) + 2 * ((a) + (p + 1) * ((iq) + q * (e)))]
        for (int e = 0; e < nelem; ++e)
            for (int i = 0; i < q; ++i)
                for (int c = 0; c <= p; ++c)
                    for (int d = 0; d <= p; ++d) {
                        int ia = O[e] + c, ib = O[e] + d;
                        double M = NNV(0, c, i, e) * NNV(0, d, i, e) * J[e] * W[i];
                        double K = NNV(1, c, i, e) * NNV(1, d, i, e) * J[e] * W[i];
                        double B = NNV(1, c, i, e) * NNV(0, d, i, e) * J[e] * W[i];
                        double BT = NNV(0, c, i, e) * NNV(1, d, i, e) * J[e] * W[i];
                        double val = mix[0] * M + mix[1] * K + mix[2] * B + mix[3] * BT;
                        ab[(kl + ku + ia - ib) + (size_t)ldab * ib] += val;
                    }
#undef NNV
    }
    free(O); free(J); free(X); free(NN);
    return rc;
}

/* ======================================================================
 * banded LU -- the LAPACK dgbtf2 / dgbtrs algorithms (LAPACK 3.x, the
 * published reference implementation), restated for one-based-free C.
 * dgbtrf with kl < its block size runs dgbtf2, so this is dgbtrf's result.
 * ====================================================================== */

#define AB(i, j) ab[(i) + (size_t)ldab * (j)] /* 0-based row i of band, column j */

int orc_dgbtrf(int n, int kl, int ku, double* ab, int ldab, int* ipiv)
{
    int kv = ku + kl;
    int info = 0;
    /* zero the fill-in rows of columns ku+1 .. kv-1 (0-based) */
    for (int j = ku + 1; j < (kv < n ? kv : n); ++j)
        for (int i = kv - j; i < kl; ++i) AB(i, j) = 0.0;
    int ju = 0;
    for (int j = 0; j < n; ++j) {
        if (j + kv < n)
            for (int i = 0; i < kl; ++i) AB(i, j + kv) = 0.0;
        int km = kl < n - 1 - j ? kl : n - 1 - j;
        /* idamax over AB(kv .. kv+km, j): first index of max |.| */
        int jp = 0;
        double best = fabs(AB(kv, j));
        for (int t = 1; t <= km; ++t) {
            double v = fabs(AB(kv + t, j));
            if (v > best) { best = v; jp = t; }
        }
        ipiv[j] = jp + j; /* 0-based row index */
        if (AB(kv + jp, j) != 0.0) {
            int jj = j + ku + jp;
            if (jj > n - 1) jj = n - 1;
            if (jj > ju) ju = jj;
            if (jp != 0) {
                /* swap rows j and j+jp across columns j..ju */
                for (int c = 0; c <= ju - j; ++c) {
                    double t = AB(kv + jp - c, j + c);
                    AB(kv + jp - c, j + c) = AB(kv - c, j + c);
                    AB(kv - c, j + c) = t;
                }
            }
            if (km > 0) {
                double rp = 1.0 / AB(kv, j);
                for (int t = 1; t <= km; ++t) AB(kv + t, j) *= rp;
                /* rank-1 update of the trailing block */
                for (int c = 1; c <= ju - j; ++c) {
                    double y = AB(kv - c, j + c);
                    if (y != 0.0)
                        for (int t = 1; t <= km; ++t)
                            AB(kv + t - c, j + c) -= AB(kv + t, j) * y;
                }
            }
        } else if (info == 0) {
            info = j + 1;
        }
    }
    return info;
}

void orc_dgbtrs(char trans, int n, int kl, int ku, int nrhs, const double* ab, int ldab,
                const int* ipiv, double* b, int ldb)
{
    int kd = ku + kl; /* 0-based row of the diagonal */
    int k = kl + ku;  /* superdiagonals of U */
    for (int r = 0; r < nrhs; ++r) {
        double* x = b + (size_t)ldb * r;
        if (trans == 'N' || trans == 'n') {
            /* L solve with row interchanges (dgbtrs, lnoti branch) */
            if (kl > 0) {
                for (int j = 0; j < n - 1; ++j) {
                    int lm = kl < n - 1 - j ? kl : n - 1 - j;
                    int l = ipiv[j];
                    if (l != j) { double t = x[l]; x[l] = x[j]; x[j] = t; }
                    double xj = x[j];
                    for (int t = 1; t <= lm; ++t) x[j + t] -= ab[(kd + t) + (size_t)ldab * j] * xj;
                }
            }
            /* dtbsv upper, no transpose, non-unit */
            for (int j = n - 1; j >= 0; --j) {
                if (x[j] != 0.0) {
                    x[j] = x[j] / ab[kd + (size_t)ldab * j];
                    double temp = x[j];
                    int i0 = j - k > 0 ? j - k : 0;
                    for (int i = j - 1; i >= i0; --i) x[i] -= temp * ab[(kd + i - j) + (size_t)ldab * j];
                }
            }
        } else {
            /* dtbsv upper, transpose, non-unit */
            for (int j = 0; j < n; ++j) {
                double temp = x[j];
                int i0 = j - k > 0 ? j - k : 0;
                for (int i = i0; i <= j - 1; ++i) temp -= ab[(kd + i - j) + (size_t)ldab * j] * x[i];
                x[j] = temp / ab[kd + (size_t)ldab * j];
            }
            /* L^T solve with interchanges in reverse */
            if (kl > 0) {
                for (int j = n - 2; j >= 0; --j) {
                    int lm = kl < n - 1 - j ? kl : n - 1 - j;
                    double s = 0.0;
                    for (int t = 1; t <= lm; ++t) s += x[j + t] * ab[(kd + t) + (size_t)ldab * j];
                    x[j] = x[j] - s;
                    int l = ipiv[j];
                    if (l != j) { double t = x[l]; x[l] = x[j]; x[j] = t; }
                }
            }
        }
    }
}
#undef AB

/* ======================================================================
 * reorders -- reorderRHS.F90 (one rank: ibeg = 1, iend = n + 1, s = n + 1)
 * ====================================================================== */

/* ReorderRHSForY (reorderRHS.F90:7-33): F(x | y,z) -> F2(y | x,z). */
void orc_reorder_for_y(const int s[3], const double* F, double* F2)
{
    for (int ix = 0; ix < s[0]; ++ix)
        for (int iy = 0; iy < s[1]; ++iy)
            for (int iz = 0; iz < s[2]; ++iz)
                F2[iy + (size_t)s[1] * (ix + (size_t)iz * s[0])] = F[ix + (size_t)s[0] * (iy + (size_t)iz * s[1])];
}

/* ReorderRHSForZ (reorderRHS.F90:36-63): F(y | x,z) -> F2(z | x,y). */
void orc_reorder_for_z(const int s[3], const double* F, double* F2)
{
    for (int ix = 0; ix < s[0]; ++ix)
        for (int iy = 0; iy < s[1]; ++iy)
            for (int iz = 0; iz < s[2]; ++iz)
                F2[iz + (size_t)s[2] * (ix + (size_t)iy * s[0])] = F[iy + (size_t)s[1] * (ix + (size_t)iz * s[0])];
}

/* ReorderRHSForX (reorderRHS.F90:66-95): F(z | x,y) -> F2(x | y,z). */
void orc_reorder_for_x(const int s[3], const double* F, double* F2)
{
    for (int ix = 0; ix < s[0]; ++ix)
        for (int iy = 0; iy < s[1]; ++iy)
            for (int iz = 0; iz < s[2]; ++iz)
                F2[ix + (size_t)s[0] * (iy + (size_t)iz * s[1])] = F[iz + (size_t)s[2] * (ix + (size_t)iy * s[0])];
}

/* ======================================================================
 * setup -- ADSS::initialize (ADS.F90:11-107) on one rank
 * ====================================================================== */

struct orc_setup {
    int n[3], p[3], nelem[3], q[3], dofs[3];
    double* U[3];
    int* O[3];
    double* J[3];
    double W[3][16];
    double* X[3];
    double* NN[3];
    /* factored mass matrices (sumfact mode and the adjoint reuse them) */
    double* ab[3];
    int* ipiv[3];
    int ldab[3];
};

orc_setup* orc_setup_create(const int n[3], const int p[3])
{
    orc_setup* s = (orc_setup*)calloc(1, sizeof(orc_setup));
    for (int d = 0; d < 3; ++d) {
        if (p[d] < 1 || p[d] > ORC_MAXP || n[d] < p[d]) { orc_setup_destroy(s); return NULL; }
        s->n[d] = n[d];
        s->p[d] = p[d];
        s->q[d] = p[d] + 1;
        s->dofs[d] = n[d] + 1;
        s->U[d] = (double*)malloc(sizeof(double) * (n[d] + p[d] + 2));
        orc_prepare_knot(n[d], p[d], s->U[d], &s->nelem[d]);
        int r = s->nelem[d], q = s->q[d];
        s->O[d] = (int*)malloc(sizeof(int) * r);
        s->J[d] = (double*)malloc(sizeof(double) * r);
        s->X[d] = (double*)malloc(sizeof(double) * q * r);
        s->NN[d] = (double*)malloc(sizeof(double) * 2 * (p[d] + 1) * q * r);
        if (orc_basis_data(p[d], n[d] + p[d] + 1, s->U[d], 1, q, r, s->O[d], s->J[d], s->W[d],
                           s->X[d], s->NN[d]) != 0) {
            orc_setup_destroy(s);
            return NULL;
        }
        s->ldab[d] = 3 * p[d] + 1;
        s->ab[d] = (double*)malloc(sizeof(double) * s->ldab[d] * (n[d] + 1));
        s->ipiv[d] = (int*)malloc(sizeof(int) * (n[d] + 1));
        const double mix[4] = {1.0, 0.0, 0.0, 0.0};
        orc_mkbbt_band(s->U[d], p[d], n[d], r, mix, s->ab[d], s->ldab[d]);
        if (orc_dgbtrf(n[d] + 1, p[d], p[d], s->ab[d], s->ldab[d], s->ipiv[d]) != 0) {
            orc_setup_destroy(s);
            return NULL;
        }
    }
    return s;
}

void orc_setup_destroy(orc_setup* s)
{
    if (!s) return;
    for (int d = 0; d < 3; ++d) {
        free(s->U[d]); free(s->O[d]); free(s->J[d]); free(s->X[d]); free(s->NN[d]);
        free(s->ab[d]); free(s->ipiv[d]);
    }
    free(s);
}

int orc_setup_dofs(const orc_setup* s, int dim) { return s->dofs[dim]; }
int orc_setup_nelem(const orc_setup* s, int dim) { return s->nelem[dim]; }

/* ======================================================================
 * integrands -- problems/heat/RHS_eq.F90:35-88, L2Projection RHS_eq.F90:34-71
 * ====================================================================== */

static double f0_value(int kind, double x, double y, double z)
{
    switch (kind) {
    case ORC_F0_CONST:
        return 1.0;
    case ORC_F0_SINPROD:
        return sin(M_PI * x) * sin(M_PI * y) * sin(M_PI * z);
    default: {
        /* 2*bump3d(0.05, 0.4, x, y, z): problems/heat/input_data.F90:68,
         * math.F90 bump3d/falloff/norm2 */
        double dx = x - 0.5, dy = y - 0.5, dz = z - 0.5;
        double t = sqrt(dx * dx + dy * dy + dz * dz);
        double r = 0.05 / 2.0, R = 0.4 / 2.0, fv;
        if (t < r) fv = 1.0;
        else if (t > R) fv = 0.0;
        else {
            double h = (t - r) / (R - r);
            double g = (h - 1.0) * (h + 1.0);
            fv = g * g;
        }
        return 2.0 * fv;
    }
    }
}

/* ======================================================================
 * FormUn + Form3DRHS -- projection_engine.F90:154-284 and :306-441
 * ====================================================================== */

#define NNA(s, d, k, a, iq, e) \
    (s)->NN[d][(k) + 2 * ((a) + ((s)->p[d] + 1) * ((iq) + (s)->q[d] * (e)))]

typedef struct {
    int kind;       /* 0 heat step, 1 projection */
    int f0;
    double dt;
    const double* c;
    const double* kappa;
    const double* cg;  /* coefficients of the gradient (MultiStep's dUn); NULL: c */
} orc_form;

/* Faithful element kernel: FormUn's value and gradient at every quadrature
 * point of the element (projection_engine.F90:360-437), then the per
 * (quadrature point, basis function) integrand call loop of Form3DRHS
 * (:212-259).  Returns the element's l2norm contribution. */
static double element_faithful(const orc_setup* s, const orc_form* fm, int ex, int ey, int ez,
                               double* elarr)
{
    const int px = s->p[0], py = s->p[1], pz = s->p[2];
    const int qx = s->q[0], qy = s->q[1], qz = s->q[2];
    const int sx = s->dofs[0], sy = s->dofs[1];
    double un[16 * 16 * 16], dun[3][16 * 16 * 16];
    double l2 = 0.0;
    const int ox = s->O[0][ex], oy = s->O[1][ey], oz = s->O[2][ez];
    double J = s->J[0][ex] * s->J[1][ey] * s->J[2][ez];
    double kq = fm->kappa ? fm->kappa[ex + s->nelem[0] * (ey + (size_t)s->nelem[1] * ez)] : 1.0;

    if (fm->kind == 0) {
        for (int kx = 0; kx < qx; ++kx)
            for (int ky = 0; ky < qy; ++ky)
                for (int kz = 0; kz < qz; ++kz) {
                    double Uval = 0.0, dux = 0.0, duy = 0.0, duz = 0.0;
                    for (int bx = 0; bx <= px; ++bx)
                        for (int by = 0; by <= py; ++by)
                            for (int bz = 0; bz <= pz; ++bz) {
                                size_t ind = (size_t)(ox + bx) + (size_t)sx * ((oy + by) + (size_t)sy * (oz + bz));
                                double Uc = fm->c[ind];
                                double Ug = fm->cg ? fm->cg[ind] : Uc;
                                double v = NNA(s, 0, 0, bx, kx, ex) * NNA(s, 1, 0, by, ky, ey) * NNA(s, 2, 0, bz, kz, ez);
                                double dvx = NNA(s, 0, 1, bx, kx, ex) * NNA(s, 1, 0, by, ky, ey) * NNA(s, 2, 0, bz, kz, ez);
                                double dvy = NNA(s, 0, 0, bx, kx, ex) * NNA(s, 1, 1, by, ky, ey) * NNA(s, 2, 0, bz, kz, ez);
                                double dvz = NNA(s, 0, 0, bx, kx, ex) * NNA(s, 1, 0, by, ky, ey) * NNA(s, 2, 1, bz, kz, ez);
                                Uval = Uval + Uc * v;
                                dux = dux + Ug * dvx;
                                duy = duy + Ug * dvy;
                                duz = duz + Ug * dvz;
                            }
                    int kk = kx + qx * (ky + qy * kz);
                    un[kk] = Uval; dun[0][kk] = dux; dun[1][kk] = duy; dun[2][kk] = duz;
                }
    }
    for (int i = 0; i < (px + 1) * (py + 1) * (pz + 1); ++i) elarr[i] = 0.0;
    for (int kx = 0; kx < qx; ++kx)
        for (int ky = 0; ky < qy; ++ky)
            for (int kz = 0; kz < qz; ++kz) {
                double W = s->W[0][kx] * s->W[1][ky] * s->W[2][kz];
                int kk = kx + qx * (ky + qy * kz);
                double fval = 0.0;
                if (fm->kind == 1)
                    fval = f0_value(fm->f0, s->X[0][kx + qx * ex], s->X[1][ky + qy * ey], s->X[2][kz + qz * ez]);
                for (int ax = 0; ax <= px; ++ax)
                    for (int ay = 0; ay <= py; ++ay)
                        for (int az = 0; az <= pz; ++az) {
                            double v = NNA(s, 0, 0, ax, kx, ex) * NNA(s, 1, 0, ay, ky, ey) * NNA(s, 2, 0, az, kz, ez);
                            double ret, l2t;
                            if (fm->kind == 0) {
                                /* heat, t > 0 (RHS_eq.F90:69-81), f = 0 */
                                double dvx = NNA(s, 0, 1, ax, kx, ex) * NNA(s, 1, 0, ay, ky, ey) * NNA(s, 2, 0, az, kz, ez);
                                double dvy = NNA(s, 0, 0, ax, kx, ex) * NNA(s, 1, 1, ay, ky, ey) * NNA(s, 2, 0, az, kz, ez);
                                double dvz = NNA(s, 0, 0, ax, kx, ex) * NNA(s, 1, 0, ay, ky, ey) * NNA(s, 2, 1, az, kz, ez);
                                double rhs = fm->dt * (-kq * (dun[0][kk] * dvx + dun[1][kk] * dvy + dun[2][kk] * dvz) + v * 0.0);
                                ret = J * W * (v * un[kk] + rhs);
                                l2t = J * W * v * un[kk] * un[kk];
                            } else {
                                /* t = 0 branch (RHS_eq.F90:82-86) */
                                ret = J * W * v * fval;
                                l2t = J * W * v * fval * fval;
                            }
                            elarr[ax + (px + 1) * (ay + (py + 1) * az)] += ret;
                            l2 += l2t;
                        }
            }
    return l2;
}

/* Sum-factorised element kernel: the same integral, contracted one direction
 * at a time (no per-(q, a) products).  Equal to element_faithful up to
 * rounding. */
static double element_sumfact(const orc_setup* s, const orc_form* fm, int ex, int ey, int ez,
                              double* elarr)
{
    const int px = s->p[0], py = s->p[1], pz = s->p[2];
    const int Px = px + 1, Py = py + 1, Pz = pz + 1;
    const int qx = s->q[0], qy = s->q[1], qz = s->q[2];
    const int sx = s->dofs[0], sy = s->dofs[1];
    const int ox = s->O[0][ex], oy = s->O[1][ey], oz = s->O[2][ez];
    double J = s->J[0][ex] * s->J[1][ey] * s->J[2][ez];
    double kq = fm->kappa ? fm->kappa[ex + s->nelem[0] * (ey + (size_t)s->nelem[1] * ez)] : 1.0;
    double l2 = 0.0;
    /* quadrature-point fields g[field][kx][ky][kz] */
    static const int MAXQ3 = 11 * 11 * 11;
    double u[4][11 * 11 * 11];
    (void)MAXQ3;
    if (fm->kind == 0) {
        double c[11 * 11 * 11];
        for (int az = 0; az < Pz; ++az)
            for (int ay = 0; ay < Py; ++ay)
                for (int ax = 0; ax < Px; ++ax)
                    c[ax + Px * (ay + Py * az)] =
                        fm->c[(size_t)(ox + ax) + (size_t)sx * ((oy + ay) + (size_t)sy * (oz + az))];
        /* x stage: t0 = N0x c, t1 = N1x c   [kx][ay][az] */
        double t0[11 * 11 * 11], t1[11 * 11 * 11];
        for (int az = 0; az < Pz; ++az)
            for (int ay = 0; ay < Py; ++ay)
                for (int kx = 0; kx < qx; ++kx) {
                    double a0 = 0.0, a1 = 0.0;
                    for (int ax = 0; ax < Px; ++ax) {
                        double cv = c[ax + Px * (ay + Py * az)];
                        a0 += NNA(s, 0, 0, ax, kx, ex) * cv;
                        a1 += NNA(s, 0, 1, ax, kx, ex) * cv;
                    }
                    t0[kx + qx * (ay + Py * az)] = a0;
                    t1[kx + qx * (ay + Py * az)] = a1;
                }
        /* y stage: s00 = N0y t0, s01 = N1y t0, s10 = N0y t1   [kx][ky][az] */
        double s00[11 * 11 * 11], s01[11 * 11 * 11], s10[11 * 11 * 11];
        for (int az = 0; az < Pz; ++az)
            for (int ky = 0; ky < qy; ++ky)
                for (int kx = 0; kx < qx; ++kx) {
                    double a = 0.0, b = 0.0, cc = 0.0;
                    for (int ay = 0; ay < Py; ++ay) {
                        double v0 = t0[kx + qx * (ay + Py * az)], v1 = t1[kx + qx * (ay + Py * az)];
                        a += NNA(s, 1, 0, ay, ky, ey) * v0;
                        b += NNA(s, 1, 1, ay, ky, ey) * v0;
                        cc += NNA(s, 1, 0, ay, ky, ey) * v1;
                    }
                    int i = kx + qx * (ky + qy * az);
                    s00[i] = a; s01[i] = b; s10[i] = cc;
                }
        /* z stage */
        for (int kz = 0; kz < qz; ++kz)
            for (int ky = 0; ky < qy; ++ky)
                for (int kx = 0; kx < qx; ++kx) {
                    double uu = 0.0, ux = 0.0, uy = 0.0, uz = 0.0;
                    for (int az = 0; az < Pz; ++az) {
                        int i = kx + qx * (ky + qy * az);
                        double n0 = NNA(s, 2, 0, az, kz, ez), n1 = NNA(s, 2, 1, az, kz, ez);
                        uu += n0 * s00[i];
                        uz += n1 * s00[i];
                        uy += n0 * s01[i];
                        ux += n0 * s10[i];
                    }
                    int kk = kx + qx * (ky + qy * kz);
                    u[0][kk] = uu; u[1][kk] = ux; u[2][kk] = uy; u[3][kk] = uz;
                }
    }
    /* pointwise integrand -> weighted fields g0 (against v), gx, gy, gz */
    for (int kz = 0; kz < qz; ++kz)
        for (int ky = 0; ky < qy; ++ky)
            for (int kx = 0; kx < qx; ++kx) {
                int kk = kx + qx * (ky + qy * kz);
                double w = J * (s->W[0][kx] * s->W[1][ky] * s->W[2][kz]);
                if (fm->kind == 0) {
                    double uv = u[0][kk];
                    l2 += w * uv * uv;
                    double f = -fm->dt * kq * w;
                    u[0][kk] = w * uv;
                    u[1][kk] *= f; u[2][kk] *= f; u[3][kk] *= f;
                } else {
                    double fv = f0_value(fm->f0, s->X[0][kx + qx * ex], s->X[1][ky + qy * ey], s->X[2][kz + qz * ez]);
                    l2 += w * fv * fv;
                    u[0][kk] = w * fv;
                    u[1][kk] = u[2][kk] = u[3][kk] = 0.0;
                }
            }
    /* transposed z stage: h00 = N0z^T g0 + N1z^T gz, h01 = N0z^T gy, h10 = N0z^T gx */
    double h00[11 * 11 * 11], h01[11 * 11 * 11], h10[11 * 11 * 11];
    for (int az = 0; az < Pz; ++az)
        for (int ky = 0; ky < qy; ++ky)
            for (int kx = 0; kx < qx; ++kx) {
                double a = 0.0, b = 0.0, cc = 0.0;
                for (int kz = 0; kz < qz; ++kz) {
                    int kk = kx + qx * (ky + qy * kz);
                    double n0 = NNA(s, 2, 0, az, kz, ez), n1 = NNA(s, 2, 1, az, kz, ez);
                    a += n0 * u[0][kk] + n1 * u[3][kk];
                    b += n0 * u[2][kk];
                    cc += n0 * u[1][kk];
                }
                int i = kx + qx * (ky + qy * az);
                h00[i] = a; h01[i] = b; h10[i] = cc;
            }
    /* transposed y stage: r0 = N0y^T h00 + N1y^T h01, r1 = N0y^T h10   [kx][ay][az] */
    double r0[11 * 11 * 11], r1[11 * 11 * 11];
    for (int az = 0; az < Pz; ++az)
        for (int ay = 0; ay < Py; ++ay)
            for (int kx = 0; kx < qx; ++kx) {
                double a = 0.0, b = 0.0;
                for (int ky = 0; ky < qy; ++ky) {
                    int i = kx + qx * (ky + qy * az);
                    double n0 = NNA(s, 1, 0, ay, ky, ey), n1 = NNA(s, 1, 1, ay, ky, ey);
                    a += n0 * h00[i] + n1 * h01[i];
                    b += n0 * h10[i];
                }
                r0[kx + qx * (ay + Py * az)] = a;
                r1[kx + qx * (ay + Py * az)] = b;
            }
    /* transposed x stage */
    for (int az = 0; az < Pz; ++az)
        for (int ay = 0; ay < Py; ++ay)
            for (int ax = 0; ax < Px; ++ax) {
                double a = 0.0;
                for (int kx = 0; kx < qx; ++kx) {
                    int i = kx + qx * (ay + Py * az);
                    a += NNA(s, 0, 0, ax, kx, ex) * r0[i] + NNA(s, 0, 1, ax, kx, ex) * r1[i];
                }
                elarr[ax + Px * (ay + Py * az)] = a;
            }
    return l2;
}

/* Element loop of Form3DRHS with its scatter (projection_engine.F90:197-281).
 * Deterministic: elements are split into ex-slabs of width T >= p; even slabs
 * run concurrently, then odd slabs, so no two threads touch one F entry and
 * the accumulation order does not depend on the thread count. */
static int form_rhs(const orc_setup* s, int mode, const orc_form* fm, double* F, double* l2norm,
                    int ez_begin, int ez_end)
{
    const int nex = s->nelem[0], ney = s->nelem[1], nez = s->nelem[2];
    const int px = s->p[0], py = s->p[1], pz = s->p[2];
    const int sx = s->dofs[0], sy = s->dofs[1], sz = s->dofs[2];
    if (ez_end < 0 || ez_end > nez) ez_end = nez;
    if (ez_begin < 0) ez_begin = 0;
    if (px > 10 || py > 10 || pz > 10) return -1;
    memset(F, 0, sizeof(double) * (size_t)sx * sy * sz);
    const int T = px > 4 ? px : 4;
    const int nslab = (nex + T - 1) / T;
    double* slab_l2 = (double*)calloc((size_t)nslab, sizeof(double));
    for (int phase = 0; phase < 2; ++phase) {
#pragma omp parallel for schedule(dynamic, 1)
        for (int sl = phase; sl < nslab; sl += 2) {
            double elarr[11 * 11 * 11];
            double l2 = 0.0;
            int e0 = sl * T, e1 = (sl + 1) * T < nex ? (sl + 1) * T : nex;
            for (int ex = e0; ex < e1; ++ex)
                for (int ey = 0; ey < ney; ++ey)
                    for (int ez = ez_begin; ez < ez_end; ++ez) {
                        l2 += (mode == ORC_MODE_FAITHFUL) ? element_faithful(s, fm, ex, ey, ez, elarr)
                                                          : element_sumfact(s, fm, ex, ey, ez, elarr);
                        const int ox = s->O[0][ex], oy = s->O[1][ey], oz = s->O[2][ez];
                        for (int ax = 0; ax <= px; ++ax)
                            for (int ay = 0; ay <= py; ++ay)
                                for (int az = 0; az <= pz; ++az)
                                    F[(size_t)(ox + ax) + (size_t)sx * ((oy + ay) + (size_t)sy * (oz + az))] +=
                                        elarr[ax + (px + 1) * (ay + (py + 1) * az)];
                    }
            slab_l2[sl] = l2;
        }
    }
    double tot = 0.0;
    for (int sl = 0; sl < nslab; ++sl) tot += slab_l2[sl];
    free(slab_l2);
    if (l2norm) *l2norm = tot;
    return 0;
}

int orc_form_rhs_heat(const orc_setup* s, int mode, const double* c, const double* kappa,
                      double dt, double* F, double* l2norm, int ez_begin, int ez_end)
{
    orc_form fm = {0, 0, dt, c, kappa, NULL};
    return form_rhs(s, mode, &fm, F, l2norm, ez_begin, ez_end);
}

int orc_form_rhs_project(const orc_setup* s, int mode, int f0_kind, double* F, double* l2norm)
{
    orc_form fm = {1, f0_kind, 0.0, NULL, NULL, NULL};
    return form_rhs(s, mode, &fm, F, l2norm, 0, -1);
}

/* ======================================================================
 * the three 1-D solves -- Sub_Step (ADS.F90:514-779) on one rank
 * ====================================================================== */

/* Solve every column of RHS(nrow, ncol) with the direction's matrix.  In the
 * faithful mode the 1-D matrix is rebuilt and refactored (ComputeMatrix +
 * SolveOneDirection, ADS.F90:553-561); otherwise the factors from setup are
 * reused.  Columns are solved one at a time as MUMPS job 3 does (:363-368). */
static int solve_direction_mix(const orc_setup* s, int d, int faithful, const double* mixp, double* rhs,
                               int ncol, char trans, int line_limit)
{
    int n = s->dofs[d], p = s->p[d];
    int ldab = 3 * p + 1;
    const double* ab = s->ab[d];
    const int* ipiv = s->ipiv[d];
    double* ab_own = NULL;
    int* ipiv_own = NULL;
    if (faithful) {
        ab_own = (double*)malloc(sizeof(double) * ldab * n);
        ipiv_own = (int*)malloc(sizeof(int) * n);
        const double mass[4] = {1.0, 0.0, 0.0, 0.0};
        orc_mkbbt_band(s->U[d], p, s->n[d], s->nelem[d], mixp ? mixp : mass, ab_own, ldab);
        if (orc_dgbtrf(n, p, p, ab_own, ldab, ipiv_own) != 0) {
            free(ab_own); free(ipiv_own);
            return -1;
        }
        ab = ab_own;
        ipiv = ipiv_own;
    }
    int lim = (line_limit > 0 && line_limit < ncol) ? line_limit : ncol;
#pragma omp parallel for schedule(static)
    for (int col = 0; col < lim; ++col)
        orc_dgbtrs(trans, n, p, p, 1, ab, ldab, ipiv, rhs + (size_t)n * col, n);
    free(ab_own);
    free(ipiv_own);
    return 0;
}

static int solve_direction(const orc_setup* s, int d, int faithful, double* rhs, int ncol, char trans,
                           int line_limit)
{
    return solve_direction_mix(s, d, faithful, NULL, rhs, ncol, trans, line_limit);
}

int orc_solve_kron(const orc_setup* s, int faithful, double* F, int line_limit)
{
    int sz3[3] = {s->dofs[0], s->dofs[1], s->dofs[2]};
    size_t D = (size_t)sz3[0] * sz3[1] * sz3[2];
    double* F2 = (double*)malloc(sizeof(double) * D);
    int rc = 0;
    /* x: F(x | y,z) solved in place (Gather/Scatter are identities on one rank) */
    rc |= solve_direction(s, 0, faithful, F, sz3[1] * sz3[2], 'N', line_limit);
    orc_reorder_for_y(sz3, F, F2);
    rc |= solve_direction(s, 1, faithful, F2, sz3[0] * sz3[2], 'N', line_limit);
    orc_reorder_for_z(sz3, F2, F);
    rc |= solve_direction(s, 2, faithful, F, sz3[0] * sz3[1], 'N', line_limit);
    orc_reorder_for_x(sz3, F, F2);
    memcpy(F, F2, sizeof(double) * D);
    free(F2);
    return rc;
}

int orc_solve_kron_T(const orc_setup* s, double* F)
{
    int sz3[3] = {s->dofs[0], s->dofs[1], s->dofs[2]};
    size_t D = (size_t)sz3[0] * sz3[1] * sz3[2];
    double* F2 = (double*)malloc(sizeof(double) * D);
    int rc = 0;
    rc |= solve_direction(s, 0, 0, F, sz3[1] * sz3[2], 'T', 0);
    orc_reorder_for_y(sz3, F, F2);
    rc |= solve_direction(s, 1, 0, F2, sz3[0] * sz3[2], 'T', 0);
    orc_reorder_for_z(sz3, F2, F);
    rc |= solve_direction(s, 2, 0, F, sz3[0] * sz3[1], 'T', 0);
    orc_reorder_for_x(sz3, F, F2);
    memcpy(F, F2, sizeof(double) * D);
    free(F2);
    return rc;
}

int orc_heat_step(const orc_setup* s, int mode, const double* c_in, const double* kappa,
                  double dt, double* c_out, double* l2norm)
{
    int rc = orc_form_rhs_heat(s, mode, c_in, kappa, dt, c_out, l2norm, 0, -1);
    if (rc) return rc;
    return orc_solve_kron(s, mode == ORC_MODE_FAITHFUL, c_out, 0);
}

/* Sub_Step's three solves with A = mix1 M + mix2 K + mix3 B + mix4 B^T in
 * every direction (ComputeMatrix(..., mix, ...), ADS.F90:553-561, 632, 722),
 * matrices rebuilt and factored as the reference does. */
int orc_solve_kron_mix(const orc_setup* s, const double mix[4], double* F)
{
    int sz3[3] = {s->dofs[0], s->dofs[1], s->dofs[2]};
    size_t D = (size_t)sz3[0] * sz3[1] * sz3[2];
    double* F2 = (double*)malloc(sizeof(double) * D);
    int rc = 0;
    rc |= solve_direction_mix(s, 0, 1, mix, F, sz3[1] * sz3[2], 'N', 0);
    orc_reorder_for_y(sz3, F, F2);
    rc |= solve_direction_mix(s, 1, 1, mix, F2, sz3[0] * sz3[2], 'N', 0);
    orc_reorder_for_z(sz3, F2, F);
    rc |= solve_direction_mix(s, 2, 1, mix, F, sz3[0] * sz3[1], 'N', 0);
    orc_reorder_for_x(sz3, F, F2);
    memcpy(F, F2, sizeof(double) * D);
    free(F2);
    return rc;
}

/* MultiStep (ADS.F90:391-429) with the heat integrand: substep s = 1..3
 * forms the right-hand side from the point values of c_in (Un, kept from
 * substep 1's FormUn) and the gradients of the latest coefficients (dUn of
 * FormUn(s)), then solves with mix(:, s) in x, y and z.  l2norm is the last
 * substep's (all three equal sum J W u_n^2).  mix is column-major 4 x 3. */
int orc_multistep(const orc_setup* s, const double mix[12], const double* c_in, const double* kappa, double dt,
                  double* c_out, double* l2norm)
{
    size_t D = (size_t)s->dofs[0] * s->dofs[1] * s->dofs[2];
    double* prev = (double*)malloc(sizeof(double) * D);
    memcpy(prev, c_in, sizeof(double) * D);
    int rc = 0;
    for (int sub = 0; sub < 3 && !rc; ++sub) {
        orc_form fm = {0, 0, dt, c_in, kappa, prev};
        rc = form_rhs(s, ORC_MODE_FAITHFUL, &fm, c_out, l2norm, 0, -1);
        if (!rc) rc = orc_solve_kron_mix(s, mix + 4 * sub, c_out);
        if (!rc) memcpy(prev, c_out, sizeof(double) * D);
    }
    free(prev);
    return rc;
}

/* ======================================================================
 * the oil problem -- problems/oil/RHS_eq.F90:69-91, input_data.F90:139-380,
 * src/math.F90:44-88
 * ====================================================================== */

double orc_oil_lerp(double t, double x, double y) { return (1.0 - t) * x + t * y; }

double orc_oil_falloff(double r, double R, double t)
{
    if (t < r) return 1.0;
    if (t > R) return 0.0;
    double h = (t - r) / (R - r);
    double g = (h - 1.0) * (h + 1.0);
    return g * g;
}

double orc_oil_dist_from_segment(const double p[3], const double a[3], const double b[3])
{
    double dx = b[0] - a[0], dy = b[1] - a[1], dz = b[2] - a[2];
    double cx = p[0] - a[0], cy = p[1] - a[1], cz = p[2] - a[2];
    double dot = dx * cx + dy * cy + dz * cz;
    double len2 = dx * dx + dy * dy + dz * dz;
    double proj = dot / len2;
    double xx, yy, zz;
    if (proj < 0) { xx = a[0]; yy = a[1]; zz = a[2]; }
    else if (proj > 1) { xx = b[0]; yy = b[1]; zz = b[2]; }
    else { xx = a[0] + proj * dx; yy = a[1] + proj * dy; zz = a[2] + proj * dz; }
    double ex = p[0] - xx, ey = p[1] - yy, ez = p[2] - zz;
    return ex * ex + ey * ey + ez * ez;
}

double orc_oil_dist_from_curves(const orc_oil* o, const double p[3])
{
    double fval = 1e3;
    for (int i = 0; i < o->cN; ++i)
        for (int j = 1; j < o->cL; ++j) {
            int k = i * o->cL + j;
            double a[3] = {o->cx[k - 1], o->cy[k - 1], o->cz[k - 1]};
            double b[3] = {o->cx[k], o->cy[k], o->cz[k]};
            double d = orc_oil_dist_from_segment(p, a, b);
            if (d < fval) fval = d;
        }
    return fval;
}

double orc_oil_kq(const orc_oil* o, double x, double y, double z)
{
    double p[3] = {x, y, z};
    double dist = sqrt(orc_oil_dist_from_curves(o, p));
    return orc_oil_lerp(orc_oil_falloff(0.0, 0.06, dist), o->kq_min, o->kq_max);
}

static double oil_norm3(double x, double y, double z) { return sqrt(x * x + y * y + z * z); }

double orc_oil_pumping(const orc_oil* o, double x, double y, double z)
{
    double fval = 0.0;
    for (int i = 0; i < o->npumps; ++i) {
        const double* q = o->pumps + 3 * i;
        fval = fval + o->pump_strength * orc_oil_falloff(0.0, o->radius, oil_norm3(q[0] - x, q[1] - y, q[2] - z));
    }
    return fval;
}

double orc_oil_draining(const orc_oil* o, double u, double x, double y, double z)
{
    double fval = 0.0;
    for (int i = 0; i < o->ndrains; ++i) {
        const double* q = o->drains + 3 * i;
        fval = fval + o->drain_strength * orc_oil_falloff(0.0, o->radius, oil_norm3(q[0] - x, q[1] - y, q[2] - z));
    }
    return fval * u;
}

double orc_oil_initial_state(const orc_oil* o, double x, double y, double z)
{
    double p[3] = {x, y, z};
    double dist = sqrt(orc_oil_dist_from_curves(o, p));
    /* bump3d(0.2, 0.6, x, y, z) = falloff(0.1, 0.3, |X - 0.5|) (math.F90:173-181) */
    double b = orc_oil_falloff(0.2 / 2.0, 0.6 / 2.0, oil_norm3(x - 0.5, y - 0.5, z - 0.5));
    return 0.1 * orc_oil_lerp(orc_oil_falloff(0.0, 0.1, dist), 0.0, 1.0) * b;
}

int orc_oil_kq_cache(const orc_setup* s, const orc_oil* o, double* kq)
{
    const int qx = s->q[0], qy = s->q[1], qz = s->q[2];
    const int nex = s->nelem[0], ney = s->nelem[1], nez = s->nelem[2];
#pragma omp parallel for collapse(2) schedule(static)
    for (int ez = 0; ez < nez; ++ez)
        for (int ey = 0; ey < ney; ++ey)
            for (int ex = 0; ex < nex; ++ex)
                for (int kz = 0; kz < qz; ++kz)
                    for (int ky = 0; ky < qy; ++ky)
                        for (int kx = 0; kx < qx; ++kx)
                            kq[kx + qx * (ky + qy * (kz + (size_t)qz * (ex + (size_t)nex * (ey + (size_t)ney * ez))))] =
                                orc_oil_kq(o, s->X[0][kx + qx * ex], s->X[1][ky + qy * ey], s->X[2][kz + qz * ez]);
    return 0;
}

/* FormUn + the per (point, basis function) integrand loop for one oil
 * element (as element_faithful); returns l2, adds the element's drained */
static double element_oil(const orc_setup* s, const orc_oil* o, const double* kqv, const double* c, double dt,
                          int tzero, int ex, int ey, int ez, double* elarr, double* drained)
{
    const int px = s->p[0], py = s->p[1], pz = s->p[2];
    const int qx = s->q[0], qy = s->q[1], qz = s->q[2];
    const int sx = s->dofs[0], sy = s->dofs[1];
    const int nex = s->nelem[0], ney = s->nelem[1];
    double un[11 * 11 * 11], dun[3][11 * 11 * 11];
    double l2 = 0.0, dr = 0.0;
    const int ox = s->O[0][ex], oy = s->O[1][ey], oz = s->O[2][ez];
    double J = s->J[0][ex] * s->J[1][ey] * s->J[2][ez];
    if (!tzero) {
        for (int kx = 0; kx < qx; ++kx)
            for (int ky = 0; ky < qy; ++ky)
                for (int kz = 0; kz < qz; ++kz) {
                    double Uval = 0.0, dux = 0.0, duy = 0.0, duz = 0.0;
                    for (int bx = 0; bx <= px; ++bx)
                        for (int by = 0; by <= py; ++by)
                            for (int bz = 0; bz <= pz; ++bz) {
                                double Uc = c[(size_t)(ox + bx) + (size_t)sx * ((oy + by) + (size_t)sy * (oz + bz))];
                                double v = NNA(s, 0, 0, bx, kx, ex) * NNA(s, 1, 0, by, ky, ey) * NNA(s, 2, 0, bz, kz, ez);
                                double dvx = NNA(s, 0, 1, bx, kx, ex) * NNA(s, 1, 0, by, ky, ey) * NNA(s, 2, 0, bz, kz, ez);
                                double dvy = NNA(s, 0, 0, bx, kx, ex) * NNA(s, 1, 1, by, ky, ey) * NNA(s, 2, 0, bz, kz, ez);
                                double dvz = NNA(s, 0, 0, bx, kx, ex) * NNA(s, 1, 0, by, ky, ey) * NNA(s, 2, 1, bz, kz, ez);
                                Uval = Uval + Uc * v;
                                dux = dux + Uc * dvx;
                                duy = duy + Uc * dvy;
                                duz = duz + Uc * dvz;
                            }
                    int kk = kx + qx * (ky + qy * kz);
                    un[kk] = Uval; dun[0][kk] = dux; dun[1][kk] = duy; dun[2][kk] = duz;
                }
    }
    for (int i = 0; i < (px + 1) * (py + 1) * (pz + 1); ++i) elarr[i] = 0.0;
    for (int kx = 0; kx < qx; ++kx)
        for (int ky = 0; ky < qy; ++ky)
            for (int kz = 0; kz < qz; ++kz) {
                double W = s->W[0][kx] * s->W[1][ky] * s->W[2][kz];
                int kk = kx + qx * (ky + qy * kz);
                double X = s->X[0][kx + qx * ex], Y = s->X[1][ky + qy * ey], Z = s->X[2][kz + qz * ez];
                double vpump = 0.0, vdrain = 0.0, kqval = 0.0, f0 = 0.0;
                if (tzero) {
                    f0 = orc_oil_initial_state(o, X, Y, Z);
                } else {
                    vpump = orc_oil_pumping(o, X, Y, Z);
                    kqval = kqv[kx + qx * (ky + qy * (kz + (size_t)qz * (ex + (size_t)nex * (ey + (size_t)ney * ez))))];
                    double d = orc_oil_draining(o, un[kk], X, Y, Z);
                    vdrain = d > 0.0 ? d : 0.0;
                }
                double fval = vpump - vdrain;
                for (int ax = 0; ax <= px; ++ax)
                    for (int ay = 0; ay <= py; ++ay)
                        for (int az = 0; az <= pz; ++az) {
                            double v = NNA(s, 0, 0, ax, kx, ex) * NNA(s, 1, 0, ay, ky, ey) * NNA(s, 2, 0, az, kz, ez);
                            double ret;
                            if (!tzero) {
                                double dvx = NNA(s, 0, 1, ax, kx, ex) * NNA(s, 1, 0, ay, ky, ey) * NNA(s, 2, 0, az, kz, ez);
                                double dvy = NNA(s, 0, 0, ax, kx, ex) * NNA(s, 1, 1, ay, ky, ey) * NNA(s, 2, 0, az, kz, ez);
                                double dvz = NNA(s, 0, 0, ax, kx, ex) * NNA(s, 1, 0, ay, ky, ey) * NNA(s, 2, 1, az, kz, ez);
                                double rhs = dt * (-kqval * exp(o->mi * un[kk]) *
                                                       (dun[0][kk] * dvx + dun[1][kk] * dvy + dun[2][kk] * dvz) +
                                                   v * fval);
                                ret = J * W * (v * un[kk] + rhs);
                                dr = dr + J * W * v * dt * vdrain;
                                l2 += J * W * v * un[kk] * un[kk];
                            } else {
                                ret = J * W * v * f0;
                                l2 += J * W * v * f0 * f0;
                            }
                            elarr[ax + (px + 1) * (ay + (py + 1) * az)] += ret;
                        }
            }
    *drained += dr;
    return l2;
}

/* the oil right-hand side over all elements, slab-parallel as form_rhs */
static int oil_form_rhs(const orc_setup* s, const orc_oil* o, const double* kq, const double* c, double dt, int tzero,
                        double* F, double* l2norm, double* drained)
{
    const int nex = s->nelem[0], ney = s->nelem[1], nez = s->nelem[2];
    const int px = s->p[0], py = s->p[1], pz = s->p[2];
    const int sx = s->dofs[0], sy = s->dofs[1], sz = s->dofs[2];
    if (px > 10 || py > 10 || pz > 10) return -1;
    memset(F, 0, sizeof(double) * (size_t)sx * sy * sz);
    const int T = px > 4 ? px : 4;
    const int nslab = (nex + T - 1) / T;
    double* slab = (double*)calloc((size_t)nslab * 2, sizeof(double));
    for (int phase = 0; phase < 2; ++phase) {
#pragma omp parallel for schedule(dynamic, 1)
        for (int sl = phase; sl < nslab; sl += 2) {
            double elarr[11 * 11 * 11];
            double l2 = 0.0, dr = 0.0;
            int e0 = sl * T, e1 = (sl + 1) * T < nex ? (sl + 1) * T : nex;
            for (int ex = e0; ex < e1; ++ex)
                for (int ey = 0; ey < ney; ++ey)
                    for (int ez = 0; ez < nez; ++ez) {
                        l2 += element_oil(s, o, kq, c, dt, tzero, ex, ey, ez, elarr, &dr);
                        const int ox = s->O[0][ex], oy = s->O[1][ey], oz = s->O[2][ez];
                        for (int ax = 0; ax <= px; ++ax)
                            for (int ay = 0; ay <= py; ++ay)
                                for (int az = 0; az <= pz; ++az)
                                    F[(size_t)(ox + ax) + (size_t)sx * ((oy + ay) + (size_t)sy * (oz + az))] +=
                                        elarr[ax + (px + 1) * (ay + (py + 1) * az)];
                    }
            slab[2 * sl] = l2;
            slab[2 * sl + 1] = dr;
        }
    }
    double t2 = 0.0, td = 0.0;
    for (int sl = 0; sl < nslab; ++sl) { t2 += slab[2 * sl]; td += slab[2 * sl + 1]; }
    free(slab);
    if (l2norm) *l2norm = t2;
    if (drained) *drained = td;
    return 0;
}

int orc_oil_form_rhs(const orc_setup* s, const orc_oil* o, const double* kq, const double* c, double dt, int tzero,
                     double* F, double* l2norm, double* drained)
{
    return oil_form_rhs(s, o, kq, c, dt, tzero, F, l2norm, drained);
}

int orc_oil_step(const orc_setup* s, const orc_oil* o, const double* kq, const double* c_in, double dt,
                 double* c_out, double* l2norm, double* drained)
{
    int rc = oil_form_rhs(s, o, kq, c_in, dt, 0, c_out, l2norm, drained);
    if (rc) return rc;
    return orc_solve_kron(s, 1, c_out, 0);
}

int orc_oil_project(const orc_setup* s, const orc_oil* o, double* c_out, double* l2norm)
{
    int rc = oil_form_rhs(s, o, NULL, NULL, 0.0, 1, c_out, l2norm, NULL);
    if (rc) return rc;
    return orc_solve_kron(s, 1, c_out, 0);
}

int orc_project(const orc_setup* s, int mode, int f0_kind, double* c_out, double* l2norm)
{
    int rc = orc_form_rhs_project(s, mode, f0_kind, c_out, l2norm);
    if (rc) return rc;
    return orc_solve_kron(s, mode == ORC_MODE_FAITHFUL, c_out, 0);
}

/* ======================================================================
 * adjoint of one heat step
 *   c_{n+1} = M^-1 A(kappa) c_n,  A = M - dt K(kappa) symmetric
 *   mu = M^-T lam_{n+1};  lam_n = A mu;  dL/dkappa_e += -dt sum_q JW grad c_n . grad mu
 * ====================================================================== */

int orc_heat_step_adjoint(const orc_setup* s, const double* c_n, const double* kappa, double dt,
                          const double* lam_next, double* lam_out, double* dkappa)
{
    size_t D = (size_t)s->dofs[0] * s->dofs[1] * s->dofs[2];
    double* mu = (double*)malloc(sizeof(double) * D);
    memcpy(mu, lam_next, sizeof(double) * D);
    int rc = orc_solve_kron_T(s, mu);
    if (rc) { free(mu); return rc; }
    rc = orc_form_rhs_heat(s, ORC_MODE_SUMFACT, mu, kappa, dt, lam_out, NULL, 0, -1);
    if (rc) { free(mu); return rc; }
    if (dkappa) {
        const int nex = s->nelem[0], ney = s->nelem[1], nez = s->nelem[2];
        const int sx = s->dofs[0], sy = s->dofs[1];
#pragma omp parallel for collapse(2) schedule(static)
        for (int ez = 0; ez < nez; ++ez)
            for (int ey = 0; ey < ney; ++ey)
                for (int ex = 0; ex < nex; ++ex) {
                    const int px = s->p[0], py = s->p[1], pz = s->p[2];
                    const int qx = s->q[0], qy = s->q[1], qz = s->q[2];
                    const int ox = s->O[0][ex], oy = s->O[1][ey], oz = s->O[2][ez];
                    double J = s->J[0][ex] * s->J[1][ey] * s->J[2][ez];
                    double acc = 0.0;
                    for (int kx = 0; kx < qx; ++kx)
                        for (int ky = 0; ky < qy; ++ky)
                            for (int kz = 0; kz < qz; ++kz) {
                                double gc[3] = {0, 0, 0}, gm[3] = {0, 0, 0};
                                for (int bx = 0; bx <= px; ++bx)
                                    for (int by = 0; by <= py; ++by)
                                        for (int bz = 0; bz <= pz; ++bz) {
                                            size_t ind = (size_t)(ox + bx) + (size_t)sx * ((oy + by) + (size_t)sy * (oz + bz));
                                            double dvx = NNA(s, 0, 1, bx, kx, ex) * NNA(s, 1, 0, by, ky, ey) * NNA(s, 2, 0, bz, kz, ez);
                                            double dvy = NNA(s, 0, 0, bx, kx, ex) * NNA(s, 1, 1, by, ky, ey) * NNA(s, 2, 0, bz, kz, ez);
                                            double dvz = NNA(s, 0, 0, bx, kx, ex) * NNA(s, 1, 0, by, ky, ey) * NNA(s, 2, 1, bz, kz, ez);
                                            gc[0] += c_n[ind] * dvx; gc[1] += c_n[ind] * dvy; gc[2] += c_n[ind] * dvz;
                                            gm[0] += mu[ind] * dvx; gm[1] += mu[ind] * dvy; gm[2] += mu[ind] * dvz;
                                        }
                                double W = s->W[0][kx] * s->W[1][ky] * s->W[2][kz];
                                acc += J * W * (gc[0] * gm[0] + gc[1] * gm[1] + gc[2] * gm[2]);
                            }
                    dkappa[ex + nex * (ey + (size_t)ney * ez)] += -dt * acc;
                }
    }
    free(mu);
    return 0;
}

/* ======================================================================
 * synthetic inputs: SplitMix64 counter generator
 *   u_i = (splitmix64(seed * 0xD1B54A32D192ED03 + i) >> 11) * 2^-53
 *   uniform: lo + (hi - lo) * u_i
 *   lognormal: Box-Muller with u1 = u_{2i}, u2 = u_{2i+1} on the key
 *              seed * 0xD1B54A32D192ED03 + 0x8CB92BA72F3D8DD7 + 2i(+1);
 *              z = sqrt(-2 ln(1 - u1)) cos(2 pi u2); k = exp(mu + sigma z)
 * ====================================================================== */

uint64_t orc_splitmix64(uint64_t x)
{
    uint64_t z = x + 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

static double u01(uint64_t key)
{
    return (double)(orc_splitmix64(key) >> 11) * (1.0 / 9007199254740992.0);
}

void orc_oil_curves(uint64_t seed, int cN, int cL, double* cx, double* cy, double* cz)
{
    uint64_t base = seed * 0xD1B54A32D192ED03ULL + 0x2545F4914F6CDD1DULL, k = 0;
    const double step = 0.05;
    for (int i = 0; i < cN; ++i) {
        double x[3], dx[3];
        for (int d = 0; d < 3; ++d) x[d] = u01(base + k++);
        cx[i * cL] = x[0];
        cy[i * cL] = x[1];
        cz[i * cL] = x[2];
        for (int d = 0; d < 3; ++d) dx[d] = u01(base + k++);
        for (int j = 1; j < cL; ++j) {
            double ddx[3];
            k += 3; /* random_number(t): drawn and unused (input_data.F90:153) */
            for (int d = 0; d < 3; ++d) ddx[d] = u01(base + k++);
            double ss = dx[0] * dx[0] + dx[1] * dx[1] + dx[2] * dx[2];
            for (int d = 0; d < 3; ++d) ddx[d] = 2.0 * (ddx[d] - 0.5) - 0.4 * dx[d] * ss;
            for (int d = 0; d < 3; ++d) dx[d] = dx[d] + 0.4 * ddx[d];
            cx[i * cL + j] = cx[i * cL + j - 1] + step * dx[0];
            cy[i * cL + j] = cy[i * cL + j - 1] + step * dx[1];
            cz[i * cL + j] = cz[i * cL + j - 1] + step * dx[2];
        }
    }
}

void orc_uniform(uint64_t seed, int64_t n, double lo, double hi, double* out)
{
    uint64_t base = seed * 0xD1B54A32D192ED03ULL;
    for (int64_t i = 0; i < n; ++i) out[i] = lo + (hi - lo) * u01(base + (uint64_t)i);
}

void orc_lognormal(uint64_t seed, int64_t n, double mu, double sigma, double* out)
{
    uint64_t base = seed * 0xD1B54A32D192ED03ULL + 0x8CB92BA72F3D8DD7ULL;
    for (int64_t i = 0; i < n; ++i) {
        double a = u01(base + 2 * (uint64_t)i);
        double b = u01(base + 2 * (uint64_t)i + 1);
        double z = sqrt(-2.0 * log(1.0 - a)) * cos(2.0 * M_PI * b);
        out[i] = exp(mu + sigma * z);
    }
}
