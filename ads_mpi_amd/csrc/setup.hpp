// setup.hpp -- host-side construction of one direction's discretisation:
// knots, Gauss rule, B-spline tables, 1-D matrices and their banded LU.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace ads {

// One direction of the tensor-product space (one third of ADS_setup,
// src/Setup.F90:29-112, on one rank).
struct Direction {
    int n = 0;      // number of basis functions minus one
    int p = 0;      // degree
    int q = 0;      // quadrature points per element = p + 1
    int nelem = 0;  // elements (non-empty knot spans)
    int dofs = 0;   // n + 1

    std::vector<double> knots;   // U(0:n+p+1)
    std::vector<int> first_dof;  // O(e), 0-based
    std::vector<double> jac;     // J(e)
    std::vector<double> wts;     // W(q)
    std::vector<double> pts;     // X(iq, e) at [e*q + iq]
    std::vector<double> tab;     // NN(k, a, iq, e) at [((e*q + iq)*(p+1) + a)*2 + k]

    // 1-D element matrices by Gauss quadrature over the element's basis
    // values: Me(a, b) = sum_q N_a N_b J W_q, Ke(a, b) = sum_q N'_a N'_b J W_q
    // (bitwise symmetric).  emat holds rows, [((e*P + a)*2 + k)*P + b] with
    // k = 0 for Me and 1 for Ke; emat_sym the packed upper triangles,
    // [(e*2 + k)*P(P+1)/2 + t], t = a*P - a(a-1)/2 + (b - a) for a <= b.
    std::vector<double> emat;
    std::vector<double> emat_sym;

    // 1-D matrices in band form, row-major by column: band[j*(2p+1) + (p + i - j)] = A(i, j)
    std::vector<double> mass_band;
    std::vector<double> stiff_band;
    std::vector<double> adv_band;  // B(i, j) = sum N_i' N_j J W (MKBBT's B, projection_engine.F90:125)
    // the same matrices by row: rows[i*(2p+1) + t] = A(i, i-p+t), zero outside
    std::vector<double> mass_rows;
    std::vector<double> stiff_rows;

    // LU factors of the mass matrix (partial pivoting, LAPACK dgbtrf layout:
    // ldab = 3p+1, AB(kl+ku+i-j, j) at lu[j*ldab + kl+ku+i-j]) and pivots (0-based rows).
    std::vector<double> lu;
    std::vector<int> piv;
};

// Builds the direction; returns an empty string or an error message.
std::string build_direction(int n, int p, Direction& d);
// The same on a given knot vector U(0:n+p+1) (non-decreasing) with nelem
// non-empty spans in [U(p), U(n+1)] (BasisData and MKBBT on the caller's
// knots, src/basis.F90:74-112, src/projection_engine.F90:63-132).
std::string build_direction_knots(const double* U, int n, int p, int nelem, Direction& d);

// The direction with its tables widened to degree pmax >= d.p, for kernels
// compiled for one degree in x, y and z: extra quadrature points of zero
// weight (at the element's midpoint), extra basis functions with zero values
// and derivatives, element and band matrices with zero rows, columns and
// diagonals past d.p, and the LU of the widened band (the same factors: its
// extra multipliers are zero).  Every quantity the kernels form is the same.
std::string pad_direction(const Direction& d, int pmax, Direction& out);

// Banded LU with partial pivoting (dgbtf2 semantics): band in the
// column-major layout of mass_band (half-bandwidth p), result in the dgbtrf
// layout of Direction::lu and 0-based pivot rows.  Returns an error message
// for a zero pivot.
std::string band_lu(int N, int p, const std::vector<double>& band, std::vector<double>& lu, std::vector<int>& piv);

// mix(1)M + mix(2)K + mix(3)B + mix(4)B^T in the layout of mass_band
// (ComputeMatrix -> MKBBT, projection_engine.F90:531-544, 63-132).
void mix_band(const Direction& d, const double mix[4], std::vector<double>& band);

// Dense (n+1)x(n+1) row-major mix(1)M + mix(2)K + mix(3)B + mix(4)B^T.
void dense_matrix(const Direction& d, const double mix[4], double* out);

// Sweep tables derived from the LU factors (see sweep kernels):
//   fwd[j*p + t-1]      = L(j+t, j) multiplier, t = 1..p (0 past the end)
//   fwd_piv[j]          = ipiv(j) - j in 0..p
//   bwd[j*(2p+1) + 0]   = 1 / U(j, j)
//   bwd[j*(2p+1) + t]   = U(j-t, j), t = 1..2p (0 before the start)
void sweep_tables(const Direction& d, std::vector<double>& fwd, std::vector<int>& fwd_piv,
                  std::vector<double>& bwd);
// The same from any factors in the layout of Direction::lu / piv.
void sweep_tables(int N, int p, const std::vector<double>& lu, const std::vector<int>& piv, std::vector<double>& fwd,
                  std::vector<int>& fwd_piv, std::vector<double>& bwd);

// Row tables of the mass matrix's LU factors without row interchanges (M is
// symmetric positive definite; factored here in FP64 from mass_band) for the
// segmented FP32 solve: tab[j*(2p+1) + t] = L(j, j-1-t) for t < p (L unit
// lower), 1/U(j, j) at t = p, U(j, j+t-p) for t > p; zero outside the
// matrix; then row n, the middle row, which every row in [lo, hi) matches to
// tol relative (per part: L, 1/U(j,j), U).  The recurrences y_j = b_j -
// sum L(j, j-s) y_(j-s) and x_j = (y_j - sum U(j, j+s) x_(j+s)) / U(j, j)
// forget their state geometrically; w (a multiple of 16, at most wmax) is a
// distance after which a unit state has decayed below tol_w in both, so a
// solve started w entries away from zero state matches the exact one to
// about tol_w relative.  A line of at most seg_max entries is one segment,
// solved from its true ends (exact for any w): its w is 16.  Returns an
// error message for a non-positive pivot or when no such w exists.
std::string seg_lu_tables(const Direction& d, double tol, double tol_w, int wmax, int seg_max, std::vector<float>& tab,
                          int& w, int& lo, int& hi);

// Tables of the output path (SaveSplinePlot, src/plot.F90:97-145): for the
// res points u_i = (1 - t) lo + t hi, t = i / (res - 1) (math.F90:44-51),
// first[i] = span - p and vals[i*(p+1) + a] = the p+1 non-zero B-splines at
// u_i (DersBasisFuns with d = 0, as EvalSpline, basis.F90:323-369).
void grid_tables(const Direction& d, int res, double lo, double hi, std::vector<int>& first,
                 std::vector<double>& vals);

// Synthetic input generator shared with the host API (SplitMix64 counters).
uint64_t splitmix64(uint64_t x);
void fill_uniform(uint64_t seed, int64_t n, double lo, double hi, double* out);
// oil curves: the reference's random walk (problems/oil/input_data.F90:139-165)
// with uniforms from SplitMix64(seed) in the same call order
void oil_curves(uint64_t seed, int cN, int cL, double* cx, double* cy, double* cz);
void fill_lognormal(uint64_t seed, int64_t n, double mu, double sigma, double* out);

}  // namespace ads
