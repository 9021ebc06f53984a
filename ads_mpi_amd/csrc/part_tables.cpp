// part_tables.cpp -- host tables of the partitioned banded solve (part.hpp):
// the no-interchange LU of the 1-D mass matrix in FP64 and the homogeneous
// responses of both recurrences per partition.
#include <algorithm>
#include <cmath>

#include "part.hpp"

namespace ads {

std::string part_tables(const Direction& d, int smax, PartHost& t)
{
    const int n = d.dofs, p = d.p, bw = 2 * p + 1;
    t = PartHost();
    t.n = n;
    t.p = p;
    // partitions of exactly smax entries (the last one padded past the line
    // with zero rows, which the recurrences carry as exact zeros), so the
    // kernels run a fixed number of steps; smax >= 2p
    if (smax < 2 * p) return "partitions shorter than twice the degree";
    const int K = (n + smax - 1) / smax;
    t.K = K;
    t.S = smax;
    const int rows = K * t.S + p + 1;

    // M = L U without interchanges on a dense band copy: a[i][i-j+p] = M(i, j)
    // (mass_band holds column j at [j*bw + p + i - j])
    std::vector<double> a(static_cast<size_t>(n) * bw, 0.0);
    for (int j = 0; j < n; ++j)
        for (int i = std::max(0, j - p); i <= std::min(n - 1, j + p); ++i)
            a[static_cast<size_t>(i) * bw + (j - i + p)] = d.mass_band[static_cast<size_t>(j) * bw + (p + i - j)];
    auto A = [&](int i, int j) -> double& { return a[static_cast<size_t>(i) * bw + (j - i + p)]; };
    t.lrow.assign(static_cast<size_t>(rows) * p, 0.0);
    t.urow.assign(static_cast<size_t>(rows) * (p + 1), 0.0);
    for (int k = 0; k < n; ++k) {
        const double piv = A(k, k);
        if (!(piv > 0.0)) return "non-positive pivot in the mass matrix factorisation";
        for (int i = k + 1; i <= std::min(n - 1, k + p); ++i) {
            const double l = A(i, k) / piv;
            t.lrow[static_cast<size_t>(i) * p + (i - k - 1)] = l;
            for (int j = k + 1; j <= std::min(n - 1, k + p); ++j) A(i, j) -= l * A(k, j);
        }
        t.urow[static_cast<size_t>(k) * (p + 1)] = 1.0 / piv;
        for (int s = 1; s <= p && k + s < n; ++s) t.urow[static_cast<size_t>(k) * (p + 1) + s] = A(k, k + s);
    }

    // responses per partition
    t.gsp.assign(static_cast<size_t>(rows) * p, 0.0);
    t.hsp.assign(static_cast<size_t>(rows) * p, 0.0);
    std::vector<double> w(static_cast<size_t>(n) + 2 * p + 2);
    for (int k = 0; k < K; ++k) {
        const int lo = k * t.S, hi = std::min(n, lo + t.S);
        if (lo >= hi) break;
        for (int i = 0; i < p; ++i) {
            // forward: y(lo - 1 - i) = 1, other states 0; w[m + p] holds y(m)
            std::fill(w.begin(), w.end(), 0.0);
            w[static_cast<size_t>(lo - 1 - i + p)] = 1.0;
            for (int j = lo; j < hi; ++j) {
                double y = 0.0;
                for (int s = 0; s < p; ++s) y -= t.lrow[static_cast<size_t>(j) * p + s] * w[static_cast<size_t>(j - 1 - s + p)];
                w[static_cast<size_t>(j + p)] = y;
                t.gsp[static_cast<size_t>(j) * p + i] = y;
            }
            // backward: x(hi + i) = 1; w[m] holds x(m)
            std::fill(w.begin(), w.end(), 0.0);
            w[static_cast<size_t>(hi + i)] = 1.0;
            for (int j = hi - 1; j >= lo; --j) {
                double x = 0.0;
                for (int s = 1; s <= p; ++s)  // U(j, j+s) is zero for j+s >= n
                    x -= t.urow[static_cast<size_t>(j) * (p + 1) + s] * w[static_cast<size_t>(j + s)];
                x *= t.urow[static_cast<size_t>(j) * (p + 1)];
                w[static_cast<size_t>(j)] = x;
                t.hsp[static_cast<size_t>(j) * p + i] = x;
            }
        }
    }
    return "";
}

}  // namespace ads
