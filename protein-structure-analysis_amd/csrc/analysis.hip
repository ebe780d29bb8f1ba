// analysis.hip -- gfx950 kernel of the ensemble metrics of protein_analysis
// (compute_mean_dist_matrix / compute_contact_freq_matrix,
// src/analysis_ext.cpp:85-167): for every residue pair i < j, the sum over
// snapshots of the Calpha-Calpha distance and the count of snapshots with
// distance^2 < cutoff^2, then sum / n_snaps and count / n_snaps into both
// (i, j) and (j, i).
//
// Layout: coordinates (n_snaps, n_ca, 3) float64 as the reference takes them.
// One workgroup owns a tile of 4 rows i x 64 columns j: a wavefront per row
// (its coordinates are wave-uniform loads), a lane per column (the 64 lanes
// read one contiguous 1.5 KB strip per snapshot, shared by the tile's four
// wavefronts through the L1 cache).  Each lane keeps its pair's sum and count
// in registers over all snapshots, in snapshot order, with the distance
// computed as the reference writes it and without FMA contraction, so the
// results equal the serial reference loop bit for bit; kUnroll snapshots'
// distances are in flight at once.  The work is FP64-ALU bound (the
// correctly rounded square root dominates).
#include "alma_internal.h"

#include <cmath>

namespace alma {

namespace {

// tile and unroll from a sweep at 200-1,500 Calpha (rows per workgroup 1-4,
// unroll 4-16, two rows per lane): within 10 % of each other, this the best
constexpr int kRows = 4;      // rows i per workgroup (one per wavefront)
constexpr int kCols = 64;     // columns j per workgroup (one per lane)
constexpr int kUnroll = 4;    // snapshots in flight per lane
constexpr int kAnBlock = 64 * kRows;

// distance^2 of atoms a, b of one snapshot, as the reference writes it
__device__ __forceinline__ double dist2(const double* p, int a, int b) {
#pragma clang fp contract(off)
    const double dx = p[3 * a] - p[3 * b], dy = p[3 * a + 1] - p[3 * b + 1], dz = p[3 * a + 2] - p[3 * b + 2];
    return dx * dx + dy * dy + dz * dz;
}

template <bool kMean, bool kFreq>
__global__ void __launch_bounds__(kAnBlock)
    k_ensemble(const double* __restrict__ xyz, int n_snaps, int n, double cut2, const int2* __restrict__ tiles,
               double* __restrict__ mean, double* __restrict__ freq) {
    const int2 t = tiles[blockIdx.x];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    // the row is wave-uniform: its coordinates come through scalar loads
    const int i = __builtin_amdgcn_readfirstlane(t.x * kRows + wave);
    const int j = t.y * kCols + lane;
    if (i >= n || j >= n || j < i) return;
    if (i == j) {
        if (kMean) mean[(size_t)i * n + i] = 0.0;
        if (kFreq) freq[(size_t)i * n + i] = 0.0;
        return;
    }
    const size_t stride = (size_t)n * 3;
    double sum = 0.0;
    int cnt = 0;
    int s = 0;
    // the distances of kUnroll snapshots are independent; the sum takes
    // them in snapshot order
    for (; s + kUnroll <= n_snaps; s += kUnroll) {
        double d2[kUnroll];
#pragma unroll
        for (int u = 0; u < kUnroll; ++u) d2[u] = dist2(xyz + (size_t)(s + u) * stride, i, j);
#pragma unroll
        for (int u = 0; u < kUnroll; ++u) {
            if (kMean) sum += sqrt(d2[u]);
            if (kFreq) cnt += d2[u] < cut2 ? 1 : 0;
        }
    }
    for (; s < n_snaps; ++s) {
        const double d2 = dist2(xyz + (size_t)s * stride, i, j);
        if (kMean) sum += sqrt(d2);
        if (kFreq) cnt += d2 < cut2 ? 1 : 0;
    }
    const double inv = 1.0 / (double)n_snaps;
    if (kMean) {
        const double v = sum * inv;
        mean[(size_t)i * n + j] = v;
        mean[(size_t)j * n + i] = v;
    }
    if (kFreq) {
        const double v = (double)cnt * inv;
        freq[(size_t)i * n + j] = v;
        freq[(size_t)j * n + i] = v;
    }
}

// atoms idx[0..m) of a (n, 4) coordinate array -> (m, 3)
__global__ void k_gather_xyz(const double* __restrict__ pos4, const int* __restrict__ idx, int m,
                             double* __restrict__ out) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= 3 * m) return;
    const int a = k / 3, d = k - 3 * a;
    out[k] = pos4[4 * (size_t)idx[a] + d];
}

}  // namespace

hipError_t launch_gather_xyz(const double* pos4, const int* idx, int m, double* out, hipStream_t st) {
    if (m <= 0) return hipSuccess;
    k_gather_xyz<<<(3 * m + 255) / 256, 256, 0, st>>>(pos4, idx, m, out);
    return hipGetLastError();
}

// Tiles (row block, column block) holding a pair i < j or a diagonal
// element: every matrix element is written by exactly one tile.
void ensemble_tiles(int n, std::vector<int2>& tiles) {
    tiles.clear();
    const int nbi = (n + kRows - 1) / kRows, nbj = (n + kCols - 1) / kCols;
    for (int bj = 0; bj < nbj; ++bj)
        for (int bi = 0; bi < nbi; ++bi)
            if (std::min(n - 1, bj * kCols + kCols - 1) >= bi * kRows) tiles.push_back(make_int2(bi, bj));
}

hipError_t launch_ensemble(const double* xyz, int n_snaps, int n, double cut2, const int2* tiles, int n_tiles,
                           double* mean, double* freq, hipStream_t st) {
    if (n_tiles <= 0) return hipSuccess;
    if (mean && freq)
        k_ensemble<true, true><<<n_tiles, kAnBlock, 0, st>>>(xyz, n_snaps, n, cut2, tiles, mean, freq);
    else if (mean)
        k_ensemble<true, false><<<n_tiles, kAnBlock, 0, st>>>(xyz, n_snaps, n, cut2, tiles, mean, freq);
    else if (freq)
        k_ensemble<false, true><<<n_tiles, kAnBlock, 0, st>>>(xyz, n_snaps, n, cut2, tiles, mean, freq);
    return hipGetLastError();
}

}  // namespace alma
