// part.hpp -- the partitioned banded solve: tables (host) and launch (device).
//
// A 1-D mass matrix M (symmetric positive definite, half-bandwidth p) is
// factored without row interchanges, M = L U, so a solve is the two
// recurrences
//     forward   y_j = b_j - sum_{t=1..p} L(j, j-t) y_(j-t)
//     backward  x_j = (y_j - sum_{t=1..p} U(j, j+t) x_(j+t)) / U(j, j).
// A line of n entries is cut into K partitions [a_k, b_k) that run in
// parallel, each from zero state.  Both recurrences are linear, so the true
// values differ from the local ones by the response to the true state at the
// partition's edge:
//     y_j = y_loc_j + sum_i g_i(j) s_k[i],   s_k[i] = y_(a_k - 1 - i)
//     x_j = x_loc_j + sum_i h_i(j) t_k[i],   t_k[i] = x_(b_k + i)
// where g_i (h_i) is the homogeneous forward (backward) recurrence started
// from a unit state i at the partition's lower (upper) edge.  The edge states
// follow from short scans over the partitions,
//     s_(k+1)[r] = y_loc_(b_k - 1 - r) + sum_i g_i(b_k - 1 - r) s_k[i]
//     t_(k-1)[r] = x_loc_(a_k + r)     + sum_i h_i(a_k + r) t_k[i],
// so the solve is exact (no truncation): it differs from the sequential
// substitution by rounding only.  The kernels take g and h only at the edge
// rows (for the scans); the corrections sum_i g_i(j) s_k[i] and
// sum_i h_i(j) t_k[i] are formed by running the homogeneous recurrences
// again from the true states, with the factor rows already in registers.
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "setup.hpp"

namespace ads {

// Host tables of one direction (FP64), rows j = 0 .. K*S + p (zero beyond n-1):
//   lrow[j*p + t]     = L(j, j-1-t)           t = 0..p-1
//   urow[j*(p+1)]     = 1 / U(j, j),  urow[j*(p+1) + t] = U(j, j+t), t = 1..p
//   gsp[j*p + i]      = g_i(j) of j's partition
//   hsp[j*p + i]      = h_i(j) of j's partition
struct PartHost {
    int n = 0, p = 0, K = 1, S = 0;
    std::vector<double> lrow, urow, gsp, hsp;
};

// Partitions of smax >= 2p entries, the last one padded past the line's end
// with zero rows.  Returns an error message for a non-positive pivot.
std::string part_tables(const Direction& d, int smax, PartHost& t);

// Device copies (one allocation per array, owned by the caller).
template <typename T>
struct PartDev {
    const T* lrow;
    const T* urow;
    const T* gsp;
    const T* hsp;
    int n, K, S;
};

// Entries per partition (a lane keeps its partition in registers) and most
// partitions per line (one wave each, so a workgroup of at most 1024
// threads).  Lines of up to kPartMaxK * kPartSeg entries take partitions of
// kPartSeg; longer ones, up to kPartMaxK * kPartSegLong, partitions of
// kPartSegLong (their kernels do not double-buffer, which leaves the
// registers for the longer partition); longer lines use the LU sweeps.
constexpr int kPartSeg = 24;
constexpr int kPartSegLong = 32;
constexpr int kPartMaxK = 16;
constexpr int kPartMaxLine = kPartMaxK * kPartSegLong;

// the partition length for a line of n entries (0: too long for the kernels)
inline int part_seg(int n)
{
    if (n <= kPartMaxK * kPartSeg) return kPartSeg;
    if (n <= kPartMaxLine) return kPartSegLong;
    return 0;
}

// The kernels address the 64 lines of a group from its first line with
// 32-bit byte offsets.  The y lines of a group span at most 64/nx + 1
// planes: offsets below (nx ny + 64 (nx + ny)) elements, which must stay
// below 4 GiB.
inline bool part_planes_fit(long nx, long ny, long elem_bytes)
{
    return (nx * ny + 64 * (nx + ny)) * elem_bytes < (1L << 32);
}

// F <- (Mz (x) My (x) Mx)^-1 F in place, direction by direction (x rows,
// then y and z lines), with the partitioned solve (M symmetric: also the
// transposed solve).
template <typename T>
hipError_t launch_part_solve(int p, T* F, const PartDev<T> pd[3], hipStream_t s);

}  // namespace ads
