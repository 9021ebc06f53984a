// setup.cpp -- knots, quadrature, B-spline tables, 1-D matrices and banded
// LU for one direction, on the host, once per run.
//
// Reference semantics followed (macwozni/ADS_MPI):
//   knots         src/knot_vector.F90:79-92 (interior knots are REAL(4) quotients, :89)
//   elements      src/basis.F90:91-97 (first DOF of each non-empty span)
//   Gauss rule    src/gauss.F90:71-107 (closed forms for q <= 5)
//   basis tables  src/basis.F90:99-110 with DersBasisFuns (:136-202), i.e.
//                 Piegl & Tiller, The NURBS Book, algorithm A2.3
//   matrices      src/projection_engine.F90:103-129 (MKBBT accumulation order)
//   banded LU     LAPACK dgbtf2 (what dgbtrf runs for kl < its block size);
//                 replaces the MUMPS factorisation of SolveOneDirection
//                 (src/ADS.F90:352-361), including its zero-pivot failure.
#include "setup.hpp"

#include "bspline.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>

namespace ads {

namespace {

// Gauss-Legendre nodes and weights on [-1, 1].
bool gauss_rule(int q, std::vector<double>& x, std::vector<double>& w)
{
    x.assign(q, 0.0);
    w.assign(q, 0.0);
    switch (q) {
    case 1:
        x[0] = 0.0;
        w[0] = 2.0;
        return true;
    case 2:
        x[0] = -1.0 / std::sqrt(3.0);
        x[1] = -x[0];
        w[0] = w[1] = 1.0;
        return true;
    case 3:
        x[0] = -std::sqrt(3.0 / 5.0);
        x[1] = 0.0;
        x[2] = -x[0];
        w[0] = w[2] = 5.0 / 9.0;
        w[1] = 8.0 / 9.0;
        return true;
    case 4: {
        const double r = 2.0 * std::sqrt(6.0 / 5.0);
        x[0] = -std::sqrt((3.0 + r) / 7.0);
        x[1] = -std::sqrt((3.0 - r) / 7.0);
        x[2] = -x[1];
        x[3] = -x[0];
        w[0] = w[3] = (18.0 - std::sqrt(30.0)) / 36.0;
        w[1] = w[2] = (18.0 + std::sqrt(30.0)) / 36.0;
        return true;
    }
    case 5: {
        const double r = 2.0 * std::sqrt(10.0 / 7.0);
        x[0] = -1.0 / 3.0 * std::sqrt(5.0 + r);
        x[1] = -1.0 / 3.0 * std::sqrt(5.0 - r);
        x[2] = 0.0;
        x[3] = -x[1];
        x[4] = -x[0];
        w[0] = w[4] = (322.0 - 13.0 * std::sqrt(70.0)) / 900.0;
        w[1] = w[3] = (322.0 + 13.0 * std::sqrt(70.0)) / 900.0;
        w[2] = 128.0 / 225.0;
        return true;
    }
    default:
        break;
    }
    if (q < 1 || q > 10) return false;
    // Newton on the Legendre polynomial for the remaining orders.
    for (int i = 0; i < (q + 1) / 2; ++i) {
        double z = -std::cos(M_PI * (i + 0.75) / (q + 0.5));
        double dp = 1.0;
        for (int it = 0; it < 64; ++it) {
            double pm = 1.0, pc = z;
            for (int k = 2; k <= q; ++k) {
                double pn = ((2 * k - 1) * z * pc - (k - 1) * pm) / k;
                pm = pc;
                pc = pn;
            }
            dp = q * (z * pc - pm) / (z * z - 1.0);
            double step = pc / dp;
            z -= step;
            if (std::fabs(step) < 1e-17) break;
        }
        x[i] = z;
        x[q - 1 - i] = -z;
        w[i] = w[q - 1 - i] = 2.0 / ((1.0 - z * z) * dp * dp);
    }
    if (q % 2) x[q / 2] = 0.0;
    return true;
}

std::string finish_direction(Direction& d, bool open_uniform);

}  // namespace

std::string build_direction(int n, int p, Direction& d)
{
    if (p < 1 || p > kMaxDegree) return "degree must be 1.." + std::to_string(kMaxDegree);
    if (n < p) return "n must be >= p (at least one element)";
    d = Direction();
    d.n = n;
    d.p = p;
    d.q = p + 1;
    d.dofs = n + 1;

    // knots (FillOpenKnot): zeros, then ones, then the REAL(4) interior quotients
    d.knots.assign(n + p + 2, 0.0);
    for (int k = n + 1; k <= n + p + 1; ++k) d.knots[k] = 1.0;
    const float denom = static_cast<float>(n - p + 1);
    for (int k = p + 1; k <= n; ++k) {
        const float num = static_cast<float>(k - p);
        const float v = num / denom;
        d.knots[k] = static_cast<double>(v);
    }
    return finish_direction(d, true);
}

std::string build_direction_knots(const double* U, int n, int p, int nelem, Direction& d)
{
    if (p < 1 || p > kMaxDegree) return "degree must be 1.." + std::to_string(kMaxDegree);
    if (n < p) return "n must be >= p (at least one element)";
    if (!U) return "no knot vector";
    for (int k = 0; k <= n + p + 1; ++k) {
        if (!std::isfinite(U[k])) return "knot vector has a non-finite entry";
        if (k > 0 && U[k] < U[k - 1]) return "knot vector is not non-decreasing";
    }
    d = Direction();
    d.n = n;
    d.p = p;
    d.q = p + 1;
    d.dofs = n + 1;
    d.knots.assign(U, U + n + p + 2);
    const std::string e = finish_direction(d, false);
    if (!e.empty()) return e;
    if (d.nelem != nelem)
        return "nelem (" + std::to_string(nelem) + ") is not the knot vector's number of non-empty spans (" +
               std::to_string(d.nelem) + ")";
    return "";
}

namespace {

// Everything of a direction after its knots: elements, Gauss rule, basis
// tables, element matrices, 1-D band matrices and the banded LU of M.
std::string finish_direction(Direction& d, bool open_uniform)
{
    const int n = d.n, p = d.p;
    // elements: the non-empty spans [U(i), U(i+1)), i = p..n; first DOF = i - p
    for (int i = p; i <= n; ++i)
        if (d.knots[i] != d.knots[i + 1]) d.first_dof.push_back(i - p);
    d.nelem = static_cast<int>(d.first_dof.size());
    if (d.nelem < 1) return "knot vector has no non-empty span";
    if (open_uniform && d.nelem != n - p + 1) return "knot vector has repeated interior knots";

    std::vector<double> xg;
    if (!gauss_rule(d.q, xg, d.wts)) return "unsupported quadrature order";

    const int q = d.q, P = p + 1;
    d.jac.resize(d.nelem);
    d.pts.resize(static_cast<size_t>(d.nelem) * q);
    d.tab.resize(static_cast<size_t>(d.nelem) * q * P * 2);
    double vals[(kMaxDegree + 1) * 2];
    for (int e = 0; e < d.nelem; ++e) {
        const int span = d.first_dof[e] + p;
        const double J = (d.knots[span + 1] - d.knots[span]) / 2.0;
        d.jac[e] = J;
        for (int iq = 0; iq < q; ++iq) {
            const double u = (xg[iq] + 1.0) * J + d.knots[span];
            d.pts[static_cast<size_t>(e) * q + iq] = u;
            basis_and_derivative(span, u, p, d.knots.data(), vals);
            for (int a = 0; a < P; ++a)
                for (int k = 0; k < 2; ++k)
                    d.tab[((static_cast<size_t>(e) * q + iq) * P + a) * 2 + k] = vals[a * 2 + k];
        }
    }

    // element matrices
    const int nt = P * (P + 1) / 2;
    d.emat.assign(static_cast<size_t>(d.nelem) * P * 2 * P, 0.0);
    d.emat_sym.assign(static_cast<size_t>(d.nelem) * 2 * nt, 0.0);
    for (int e = 0; e < d.nelem; ++e)
        for (int a = 0; a < P; ++a)
            for (int b = 0; b < P; ++b) {
                double m = 0.0, k = 0.0;
                for (int iq = 0; iq < q; ++iq) {
                    const double* t = &d.tab[(static_cast<size_t>(e) * q + iq) * P * 2];
                    const double jw = d.jac[e] * d.wts[iq];
                    m += (t[a * 2] * t[b * 2]) * jw;
                    k += (t[a * 2 + 1] * t[b * 2 + 1]) * jw;
                }
                d.emat[((static_cast<size_t>(e) * P + a) * 2 + 0) * P + b] = m;
                d.emat[((static_cast<size_t>(e) * P + a) * 2 + 1) * P + b] = k;
                if (a <= b) {
                    const int t = a * P - a * (a - 1) / 2 + (b - a);
                    d.emat_sym[(static_cast<size_t>(e) * 2 + 0) * nt + t] = m;
                    d.emat_sym[(static_cast<size_t>(e) * 2 + 1) * nt + t] = k;
                }
            }

    // 1-D mass, stiffness and advection matrices, band form (half-bandwidth p)
    const int bw = 2 * p + 1;
    d.mass_band.assign(static_cast<size_t>(d.dofs) * bw, 0.0);
    d.stiff_band.assign(static_cast<size_t>(d.dofs) * bw, 0.0);
    d.adv_band.assign(static_cast<size_t>(d.dofs) * bw, 0.0);
    for (int e = 0; e < d.nelem; ++e)
        for (int iq = 0; iq < q; ++iq)
            for (int c = 0; c < P; ++c)
                for (int dd = 0; dd < P; ++dd) {
                    const double* t = &d.tab[(static_cast<size_t>(e) * q + iq) * P * 2];
                    const int row = d.first_dof[e] + c, col = d.first_dof[e] + dd;
                    const size_t at = static_cast<size_t>(col) * bw + (p + row - col);
                    d.mass_band[at] += t[c * 2] * t[dd * 2] * d.jac[e] * d.wts[iq];
                    d.stiff_band[at] += t[c * 2 + 1] * t[dd * 2 + 1] * d.jac[e] * d.wts[iq];
                    d.adv_band[at] += t[c * 2 + 1] * t[dd * 2] * d.jac[e] * d.wts[iq];
                }

    d.mass_rows.assign(static_cast<size_t>(d.dofs) * bw, 0.0);
    d.stiff_rows.assign(static_cast<size_t>(d.dofs) * bw, 0.0);
    for (int i = 0; i < d.dofs; ++i)
        for (int t = 0; t < bw; ++t) {
            const int j = i - p + t;
            if (j < 0 || j >= d.dofs) continue;
            d.mass_rows[static_cast<size_t>(i) * bw + t] = d.mass_band[static_cast<size_t>(j) * bw + (p + i - j)];
            d.stiff_rows[static_cast<size_t>(i) * bw + t] = d.stiff_band[static_cast<size_t>(j) * bw + (p + i - j)];
        }

    // banded LU of the mass matrix
    return band_lu(d.dofs, p, d.mass_band, d.lu, d.piv);
}

}  // namespace

std::string band_lu(int N, int p, const std::vector<double>& band, std::vector<double>& lu, std::vector<int>& piv)
{
    // partial pivoting (dgbtf2), kl = ku = p, ldab = 3p+1
    const int bw = 2 * p + 1;
    const int kl = p, ku = p, kv = kl + ku, ldab = 2 * kl + ku + 1;
    lu.assign(static_cast<size_t>(N) * ldab, 0.0);
    piv.assign(N, 0);
    auto AB = [&](int r, int j) -> double& { return lu[static_cast<size_t>(j) * ldab + r]; };
    for (int j = 0; j < N; ++j)
        for (int i = std::max(0, j - ku); i <= std::min(N - 1, j + kl); ++i)
            AB(kv + i - j, j) = band[static_cast<size_t>(j) * bw + (p + i - j)];
    int reach = 0;  // last column touched by row interchanges so far
    for (int j = 0; j < N; ++j) {
        const int below = std::min(kl, N - 1 - j);
        int best = 0;
        double bestv = std::fabs(AB(kv, j));
        for (int t = 1; t <= below; ++t)
            if (std::fabs(AB(kv + t, j)) > bestv) {
                bestv = std::fabs(AB(kv + t, j));
                best = t;
            }
        piv[j] = j + best;
        if (AB(kv + best, j) == 0.0)
            return "singular 1-D matrix (zero pivot in column " + std::to_string(j + 1) + ")";
        reach = std::max(reach, std::min(j + ku + best, N - 1));
        if (best != 0)
            for (int c = 0; c <= reach - j; ++c) std::swap(AB(kv + best - c, j + c), AB(kv - c, j + c));
        if (below > 0) {
            const double inv = 1.0 / AB(kv, j);
            for (int t = 1; t <= below; ++t) AB(kv + t, j) *= inv;
            for (int c = 1; c <= reach - j; ++c) {
                const double y = AB(kv - c, j + c);
                if (y == 0.0) continue;
                for (int t = 1; t <= below; ++t) AB(kv + t - c, j + c) -= AB(kv + t, j) * y;
            }
        }
    }
    return std::string();
}

std::string pad_direction(const Direction& d, int pmax, Direction& out)
{
    if (pmax < d.p || pmax > kMaxDegree) return "cannot widen degree " + std::to_string(d.p) + " to " +
                                                 std::to_string(pmax);
    const int p = d.p, P = p + 1, Q = d.q, PM = pmax + 1, QM = pmax + 1, ne = d.nelem, N = d.dofs;
    out = Direction();
    out.n = d.n;
    out.p = pmax;
    out.q = QM;
    out.nelem = ne;
    out.dofs = N;
    out.knots = d.knots;
    out.first_dof = d.first_dof;
    out.jac = d.jac;
    // extra quadrature points: zero weight, at the element's midpoint
    out.wts.assign(QM, 0.0);
    for (int iq = 0; iq < Q; ++iq) out.wts[iq] = d.wts[iq];
    out.pts.assign(static_cast<size_t>(ne) * QM, 0.0);
    out.tab.assign(static_cast<size_t>(ne) * QM * PM * 2, 0.0);
    for (int e = 0; e < ne; ++e) {
        const int span = d.first_dof[e] + p;
        const double mid = 0.5 * (d.knots[span] + d.knots[span + 1]);
        for (int iq = 0; iq < QM; ++iq) {
            out.pts[static_cast<size_t>(e) * QM + iq] = iq < Q ? d.pts[static_cast<size_t>(e) * Q + iq] : mid;
            if (iq >= Q) continue;
            for (int a = 0; a < P; ++a)
                for (int k = 0; k < 2; ++k)
                    out.tab[((static_cast<size_t>(e) * QM + iq) * PM + a) * 2 + k] =
                        d.tab[((static_cast<size_t>(e) * Q + iq) * P + a) * 2 + k];
        }
    }
    // element matrices: zero rows and columns past p
    const int nt = PM * (PM + 1) / 2;
    out.emat.assign(static_cast<size_t>(ne) * PM * 2 * PM, 0.0);
    out.emat_sym.assign(static_cast<size_t>(ne) * 2 * nt, 0.0);
    for (int e = 0; e < ne; ++e)
        for (int a = 0; a < P; ++a)
            for (int k = 0; k < 2; ++k)
                for (int b = 0; b < P; ++b) {
                    const double v = d.emat[((static_cast<size_t>(e) * P + a) * 2 + k) * P + b];
                    out.emat[((static_cast<size_t>(e) * PM + a) * 2 + k) * PM + b] = v;
                    if (a <= b) {
                        const int t = a * PM - a * (a - 1) / 2 + (b - a);
                        out.emat_sym[(static_cast<size_t>(e) * 2 + k) * nt + t] = v;
                    }
                }
    // band matrices of half-bandwidth pmax (the extra diagonals are zero)
    const int bw = 2 * p + 1, bwm = 2 * pmax + 1;
    auto widen = [&](const std::vector<double>& src, std::vector<double>& dst) {
        dst.assign(static_cast<size_t>(N) * bwm, 0.0);
        for (int j = 0; j < N; ++j)
            for (int i = std::max(0, j - p); i <= std::min(N - 1, j + p); ++i)
                dst[static_cast<size_t>(j) * bwm + (pmax + i - j)] = src[static_cast<size_t>(j) * bw + (p + i - j)];
    };
    widen(d.mass_band, out.mass_band);
    widen(d.stiff_band, out.stiff_band);
    widen(d.adv_band, out.adv_band);
    out.mass_rows.assign(static_cast<size_t>(N) * bwm, 0.0);
    out.stiff_rows.assign(static_cast<size_t>(N) * bwm, 0.0);
    for (int i = 0; i < N; ++i)
        for (int t = 0; t < bwm; ++t) {
            const int j = i - pmax + t;
            if (j < 0 || j >= N) continue;
            out.mass_rows[static_cast<size_t>(i) * bwm + t] = out.mass_band[static_cast<size_t>(j) * bwm + (pmax + i - j)];
            out.stiff_rows[static_cast<size_t>(i) * bwm + t] =
                out.stiff_band[static_cast<size_t>(j) * bwm + (pmax + i - j)];
        }
    // the same factors with kl = ku = pmax: the extra multipliers are zero
    return band_lu(N, pmax, out.mass_band, out.lu, out.piv);
}

void mix_band(const Direction& d, const double mix[4], std::vector<double>& band)
{
    const int p = d.p, bw = 2 * p + 1, N = d.dofs;
    band.assign(static_cast<size_t>(N) * bw, 0.0);
    for (int j = 0; j < N; ++j)
        for (int i = std::max(0, j - p); i <= std::min(N - 1, j + p); ++i) {
            const size_t at = static_cast<size_t>(j) * bw + (p + i - j);
            const size_t tr = static_cast<size_t>(i) * bw + (p + j - i);  // (j, i): B^T(i, j) = B(j, i)
            band[at] = mix[0] * d.mass_band[at] + mix[1] * d.stiff_band[at] + mix[2] * d.adv_band[at] +
                       mix[3] * d.adv_band[tr];
        }
}

void dense_matrix(const Direction& d, const double mix[4], double* out)
{
    const int N = d.dofs, P = d.p + 1, q = d.q;
    std::memset(out, 0, sizeof(double) * static_cast<size_t>(N) * N);
    for (int e = 0; e < d.nelem; ++e)
        for (int iq = 0; iq < q; ++iq)
            for (int c = 0; c < P; ++c)
                for (int dd = 0; dd < P; ++dd) {
                    const double* t = &d.tab[(static_cast<size_t>(e) * q + iq) * P * 2];
                    const double jw = d.jac[e] * d.wts[iq];
                    const double M = t[c * 2] * t[dd * 2] * jw;
                    const double K = t[c * 2 + 1] * t[dd * 2 + 1] * jw;
                    const double B = t[c * 2 + 1] * t[dd * 2] * jw;
                    const double BT = t[c * 2] * t[dd * 2 + 1] * jw;
                    const int row = d.first_dof[e] + c, col = d.first_dof[e] + dd;
                    out[static_cast<size_t>(row) * N + col] += mix[0] * M + mix[1] * K + mix[2] * B + mix[3] * BT;
                }
}

std::string seg_lu_tables(const Direction& d, double tol, double tol_w, int wmax, int seg_max, std::vector<float>& tab,
                          int& w, int& lo, int& hi)
{
    const int N = d.dofs, p = d.p, nt = 2 * p + 1;
    // M = L U without interchanges (M is symmetric positive definite), in
    // FP64 over M's band: A(i, j) at a[i*nt + p + j - i]
    std::vector<double> a(static_cast<size_t>(N) * nt, 0.0);
    auto A = [&](int i, int j) -> double& { return a[static_cast<size_t>(i) * nt + p + j - i]; };
    for (int j = 0; j < N; ++j)
        for (int i = std::max(0, j - p); i <= std::min(N - 1, j + p); ++i)
            A(i, j) = d.mass_band[static_cast<size_t>(j) * nt + p + i - j];
    for (int k = 0; k < N; ++k) {
        if (!(A(k, k) > 0.0)) return "the mass matrix is not positive definite";
        for (int i = k + 1; i <= std::min(N - 1, k + p); ++i) {
            const double l = A(i, k) / A(k, k);
            A(i, k) = l;
            for (int j = k + 1; j <= std::min(N - 1, k + p); ++j) A(i, j) -= l * A(k, j);
        }
    }
    auto at = [&](int i, int j) { return A(i, j); };
    auto row = [&](int j, std::vector<double>& r) {
        r.assign(nt, 0.0);
        for (int s = 1; s <= p; ++s)
            if (j - s >= 0) r[s - 1] = at(j, j - s);
        r[p] = 1.0 / at(j, j);
        for (int s = 1; s <= p; ++s)
            if (j + s < N) r[p + s] = at(j, j + s);
    };
    tab.assign(static_cast<size_t>(N + 1) * nt, 0.0f);
    std::vector<double> r, m;
    for (int j = 0; j < N; ++j) {
        row(j, r);
        for (int t = 0; t < nt; ++t) tab[static_cast<size_t>(j) * nt + t] = static_cast<float>(r[t]);
    }
    const int mid = N / 2;
    row(mid, m);
    for (int t = 0; t < nt; ++t) tab[static_cast<size_t>(N) * nt + t] = static_cast<float>(m[t]);
    double sl = 0.0, su = 0.0;
    for (int s = 1; s <= p; ++s) {
        sl = std::max(sl, std::fabs(m[s - 1]));
        su = std::max(su, std::fabs(m[p + s]));
    }
    auto matches = [&](int j) {
        row(j, r);
        for (int s = 1; s <= p; ++s)
            if (std::fabs(r[s - 1] - m[s - 1]) > tol * sl || std::fabs(r[p + s] - m[p + s]) > tol * su) return false;
        return std::fabs(r[p] - m[p]) <= tol * std::fabs(m[p]);
    };
    lo = mid;
    while (lo > 0 && matches(lo - 1)) --lo;
    hi = mid + 1;
    while (hi < N && matches(hi)) ++hi;
    // decay of the interior recurrences from a unit state
    int need = 0;
    for (int pass = 0; pass < 2; ++pass) {
        const int len = 4 * wmax + 4 * p;
        std::vector<double> v(len, 0.0);
        v[0] = 1.0;
        int last = 0;
        for (int k = 1; k < len; ++k) {
            double s = 0.0;
            for (int t = 1; t <= p; ++t)
                if (k - t >= 0) s += (pass == 0 ? m[t - 1] : m[p + t] * m[p]) * v[k - t];
            v[k] = -s;
            if (std::fabs(v[k]) > tol_w) last = k;
        }
        need = std::max(need, last + 1);
    }
    w = N <= seg_max ? 16 : (need + 15) / 16 * 16;
    if (w > wmax) return "the factors' recurrences decay too slowly (" + std::to_string(need) + " entries)";
    return "";
}

void sweep_tables(const Direction& d, std::vector<double>& fwd, std::vector<int>& fwd_piv,
                  std::vector<double>& bwd)
{
    sweep_tables(d.dofs, d.p, d.lu, d.piv, fwd, fwd_piv, bwd);
}

void sweep_tables(int N, int p, const std::vector<double>& lu, const std::vector<int>& piv, std::vector<double>& fwd,
                  std::vector<int>& fwd_piv, std::vector<double>& bwd)
{
    const int kl = p, ku = p, kv = kl + ku, ldab = 2 * kl + ku + 1;
    const int K = kl + ku;
    fwd.assign(static_cast<size_t>(N) * p, 0.0);
    fwd_piv.assign(N, 0);
    bwd.assign(static_cast<size_t>(N) * (K + 1), 0.0);
    for (int j = 0; j < N; ++j) {
        fwd_piv[j] = piv[j] - j;
        for (int t = 1; t <= p; ++t)
            if (j + t < N) fwd[static_cast<size_t>(j) * p + t - 1] = lu[static_cast<size_t>(j) * ldab + kv + t];
        bwd[static_cast<size_t>(j) * (K + 1)] = 1.0 / lu[static_cast<size_t>(j) * ldab + kv];
        for (int t = 1; t <= K; ++t)
            if (j - t >= 0) bwd[static_cast<size_t>(j) * (K + 1) + t] = lu[static_cast<size_t>(j) * ldab + kv - t];
    }
}


void grid_tables(const Direction& d, int res, double lo, double hi, std::vector<int>& first,
                 std::vector<double>& vals)
{
    const int P = d.p + 1;
    first.assign(res, 0);
    vals.assign(static_cast<size_t>(res) * P, 0.0);
    for (int i = 0; i < res; ++i)
        grid_point(d.n, d.p, d.knots.data(), res, lo, hi, i, &first[i], &vals[static_cast<size_t>(i) * P]);
}

uint64_t splitmix64(uint64_t x)
{
    x += 0x9E3779B97F4A7C15ULL;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ULL;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBULL;
    return x ^ (x >> 31);
}

static inline double unit_from(uint64_t key)
{
    return static_cast<double>(splitmix64(key) >> 11) * 0x1.0p-53;
}

void oil_curves(uint64_t seed, int cN, int cL, double* cx, double* cy, double* cz)
{
    const uint64_t base = seed * 0xD1B54A32D192ED03ULL + 0x2545F4914F6CDD1DULL;
    uint64_t k = 0;
    const double step = 0.05;
    for (int i = 0; i < cN; ++i) {
        double x[3], dx[3];
        for (int d = 0; d < 3; ++d) x[d] = unit_from(base + k++);
        cx[i * cL] = x[0];
        cy[i * cL] = x[1];
        cz[i * cL] = x[2];
        for (int d = 0; d < 3; ++d) dx[d] = unit_from(base + k++);
        for (int j = 1; j < cL; ++j) {
            double ddx[3];
            k += 3;  // random_number(t) is drawn and not used (input_data.F90:153)
            for (int d = 0; d < 3; ++d) ddx[d] = unit_from(base + k++);
            const double ss = dx[0] * dx[0] + dx[1] * dx[1] + dx[2] * dx[2];
            for (int d = 0; d < 3; ++d) ddx[d] = 2.0 * (ddx[d] - 0.5) - 0.4 * dx[d] * ss;
            for (int d = 0; d < 3; ++d) dx[d] = dx[d] + 0.4 * ddx[d];
            cx[i * cL + j] = cx[i * cL + j - 1] + step * dx[0];
            cy[i * cL + j] = cy[i * cL + j - 1] + step * dx[1];
            cz[i * cL + j] = cz[i * cL + j - 1] + step * dx[2];
        }
    }
}

void fill_uniform(uint64_t seed, int64_t n, double lo, double hi, double* out)
{
    const uint64_t base = seed * 0xD1B54A32D192ED03ULL;
    for (int64_t i = 0; i < n; ++i) out[i] = lo + (hi - lo) * unit_from(base + static_cast<uint64_t>(i));
}

void fill_lognormal(uint64_t seed, int64_t n, double mu, double sigma, double* out)
{
    const uint64_t base = seed * 0xD1B54A32D192ED03ULL + 0x8CB92BA72F3D8DD7ULL;
    for (int64_t i = 0; i < n; ++i) {
        const double a = unit_from(base + 2 * static_cast<uint64_t>(i));
        const double b = unit_from(base + 2 * static_cast<uint64_t>(i) + 1);
        const double z = std::sqrt(-2.0 * std::log(1.0 - a)) * std::cos(2.0 * M_PI * b);
        out[i] = std::exp(mu + sigma * z);
    }
}

}  // namespace ads
