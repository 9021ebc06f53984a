// bspline.hpp -- B-spline evaluation shared by the host set-up and device
// kernels, so that both round identically (no contraction into fused
// multiply-adds on either side).
//
//   find_span            FindSpan, Piegl & Tiller A2.1 (src/basis.F90:226-259)
//   basis_and_derivative DersBasisFuns with d = 1, A2.3 (src/basis.F90:136-202)
//   grid_point           one point of SaveSplinePlot's grid (src/plot.F90:97-145:
//                        lerp of math.F90:44-51) with its span and B-spline values
//                        (EvalSpline, src/basis.F90:323-369)
#pragma once

#include <hip/hip_runtime.h>

namespace ads {

constexpr int kMaxDegree = 9;

// The knot span of u in U(0:n+p+1): u >= U(n+1) -> n, u <= U(p) -> p.
__host__ __device__ inline int find_span(int n, int p, double u, const double* U)
{
    if (u >= U[n + 1]) return n;
    if (u <= U[p]) return p;
    int low = p, high = n + 1;
    int mid = (low + high) / 2;
    while (u < U[mid] || u >= U[mid + 1]) {
        if (u < U[mid]) high = mid;
        else low = mid;
        mid = (low + high) / 2;
    }
    return mid;
}

// The p+1 non-zero B-splines at u in span `span` and their first
// derivatives: out[2a] = N_a(u), out[2a+1] = N_a'(u), a = 0..p.
__host__ __device__ inline void basis_and_derivative(int span, double u, int p, const double* U, double* out)
{
#pragma clang fp contract(off)
    double left[kMaxDegree + 1], right[kMaxDegree + 1];
    double ndu[kMaxDegree + 1][kMaxDegree + 1];  // ndu[j][r]: lower = knot diffs, upper = values
    ndu[0][0] = 1.0;
    for (int j = 1; j <= p; ++j) {
        left[j] = u - U[span + 1 - j];
        right[j] = U[span + j] - u;
        double carry = 0.0;
        for (int r = 0; r < j; ++r) {
            ndu[j][r] = right[r + 1] + left[j - r];
            const double t = ndu[r][j - 1] / ndu[j][r];
            ndu[r][j] = carry + right[r + 1] * t;
            carry = left[j - r] * t;
        }
        ndu[j][j] = carry;
    }
    for (int r = 0; r <= p; ++r) {
        out[r * 2 + 0] = ndu[r][p];
        // first derivative (k = 1 of A2.3)
        double a_prev0 = 1.0, a_cur0 = 0.0, a_curk = 0.0;
        double der = 0.0;
        const int rk = r - 1, pk = p - 1;
        if (r >= 1) {
            a_cur0 = a_prev0 / ndu[pk + 1][rk];
            der = a_cur0 * ndu[rk][pk];
        }
        // j loop of A2.3 is empty for k = 1 (j1 = 1 > j2 = 0 or j2 = p - r <= 0)
        if (r <= pk) {
            a_curk = -a_prev0 / ndu[pk + 1][r];
            der = der + a_curk * ndu[r][pk];
        }
        out[r * 2 + 1] = der * p;
    }
}

// Point i of res points u_i = (1 - t) lo + t hi, t = i / (res - 1): its first
// basis index (span - p) goes to *first and its p+1 B-spline values to
// vals[0..p].
__host__ __device__ inline void grid_point(int n, int p, const double* U, int res, double lo, double hi, int i,
                                           int* first, double* vals)
{
#pragma clang fp contract(off)
    const double t = i / static_cast<double>(res - 1);
    // each product rounded (no fused multiply-add), as lerp evaluates it
    const double a = (1.0 - t) * lo;
    const double b = t * hi;
    const double u = a + b;
    const int span = find_span(n, p, u, U);
    double bv[(kMaxDegree + 1) * 2];
    basis_and_derivative(span, u, p, U, bv);
    *first = span - p;
    for (int k = 0; k <= p; ++k) vals[k] = bv[k * 2];
}

}  // namespace ads
