// Long-run observables beyond the Observer's sums: the pair-distance
// histogram (radial distribution function; the reference has none), the
// chain bond / angle / end-to-end histograms (the quantities
// analysis/bond_angle/h2e.cpp:64-104 writes per chain) and a 3-D bead-count
// map (analysis/trjanalysis/density_map.hpp:13-87).
//
// Integer counts from FP64 geometry on the master positions.  Every bin is
// decided by comparisons against edges the host computes in FP64 (squared
// distance edges; cosine edges compared without a square root), so a count
// is exact: a host evaluation of the same expressions gives the same bins.
#include <hip/hip_runtime.h>
#include <math.h>

#include "common.hpp"
#include "state.hpp"

namespace mdpd {

// largest k in [0, nb) with e2[k] <= x (e2 ascending, e2[0] <= x < e2[nb])
__device__ __forceinline__ int bin_of(const double* __restrict__ e2, int nb, double x) {
  int lo = 0, hi = nb - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (e2[mid] <= x) lo = mid; else hi = mid - 1;
  }
  return lo;
}

__device__ __forceinline__ void min_image(const Replica& R, double v[3]) {
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    if (v[a] > 0.5 * R.box[a]) v[a] -= R.box[a];
    else if (v[a] < -0.5 * R.box[a]) v[a] += R.box[a];
  }
}

// ---------------------------------------------------------------- RDF
// Ordered pairs i != j with r^2 in [e2[k], e2[k+1]) counted in bin k, per
// replica.  Candidates from the 27 cells around the bead's cell (rmax <= 1
// <= the cell edge), screened in FP32 with a margin, decided in FP64.  A
// block counts in LDS for the replica of its first bead and flushes once.
constexpr int kMaxRdfBins = 1024;

__global__ void __launch_bounds__(256) k_rdf(DevState d, const double* __restrict__ e2, int nb,
                                             float cut2f,
                                             unsigned long long* __restrict__ hist) {
#pragma clang fp contract(off)
  __shared__ uint32_t lh[kMaxRdfBins];
  for (int k = threadIdx.x; k < nb; k += blockDim.x) lh[k] = 0u;
  __syncthreads();
  const uint32_t s0 = blockIdx.x * blockDim.x;
  const int r0 = d.n_rep > 1 ? find_rep(d.rep_off, d.n_rep, s0 < dev_n(d) ? s0 : dev_n(d) - 1) : 0;
  const uint32_t s = s0 + threadIdx.x;
  if (s < dev_n(d)) {
    const int r = d.n_rep > 1 ? find_rep(d.rep_off, d.n_rep, s) : 0;
    const Replica& R = d.reps[r];
    const float4 pi = d.spos[s];
    const uint32_t c = d.scell[s] - R.cell_off;
    const int nx = R.nc[0], ny = R.nc[1], nz = R.nc[2];
    const int cx = (int)(c % (uint32_t)nx), cy = (int)((c / (uint32_t)nx) % (uint32_t)ny),
              cz = (int)(c / (uint32_t)(nx * ny));
    const double top = e2[nb];
    for (int oz = -1; oz <= 1; ++oz)
      for (int oy = -1; oy <= 1; ++oy)
        for (int ox = -1; ox <= 1; ++ox) {
          const int qx = (cx + ox + nx) % nx, qy = (cy + oy + ny) % ny, qz = (cz + oz + nz) % nz;
          const uint32_t q = R.cell_off + (uint32_t)((qz * ny + qy) * nx + qx);
          const uint32_t Xp = __float_as_uint(pi.x) +
                              fix_wrap(cx + ox < 0 ? 0u : (cx + ox >= nx ? 2u : 1u), (uint32_t)nx << 24);
          const float hs = R.hf[0] * kFixStep, by = pi.y - (float)oy * R.hf[1],
                      bz = pi.z - (float)oz * R.hf[2];
          for (uint32_t t = d.cell_start[q]; t < d.cell_start[q + 1]; ++t) {
            if (t == s) continue;
            const float4 pj = d.spos[t];
            const float dx = fix_dx(Xp, __float_as_uint(pj.x), hs), dy = by - pj.y, dz = bz - pj.z;
            if (dx * dx + dy * dy + dz * dz >= cut2f) continue;
            double v[3];
#pragma unroll
            for (int a = 0; a < 3; ++a) v[a] = d.pos[a * d.cap + s] - d.pos[a * d.cap + t];
            min_image(R, v);
            const double r2 = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
            if (r2 >= top || r2 < e2[0]) continue;
            const int k = bin_of(e2, nb, r2);
            if (r == r0) atomicAdd(&lh[k], 1u);
            else atomicAdd(&hist[(size_t)r * nb + k], 1ull);
          }
        }
  }
  __syncthreads();
  for (int k = threadIdx.x; k < nb; k += blockDim.x)
    if (lh[k]) atomicAdd(&hist[(size_t)r0 * nb + k], (unsigned long long)lh[k]);
}

void launch_rdf(const DevState& d, const double* e2, int nb, double rmax,
                unsigned long long* hist, hipStream_t st) {
  if (!d.nl) return;
  // FP32 screen: the squared cutoff plus a margin far above the FP32 error
  const float cut2f = (float)(rmax * rmax) + 1e-4f;
  hipLaunchKernelGGL(k_rdf, dim3(div_up(d.nl, 256)), dim3(256), 0, st, d, e2, nb, cut2f, hist);
}

// ---------------------------------------------------------------- chains
// Per chain (topo rows, amp-major): rows of out[rep][4][nb]:
//   0  first bond |r1 - r0| (head to the next bead), bins of bmax / nb
//   1  the other bonds |r_{u+1} - r_u|
//   2  cos of the angle between consecutive bond vectors, bins of 2 / nb
//      over [-1, 1]
//   3  end-to-end |r3 - r0| (minimum image of the direct difference), bins
//      of 3 bmax / nb
// Bonds and end-to-end against squared edges eb2 / ee2; the cosine bin k
// holds dot / (|b1| |b2|) in [c_k, c_{k+1}), c_k = -1 + 2k / nb, decided on
// dot and the squared lengths without a square root.
__device__ __forceinline__ bool cos_ge(double dot, double l2, double ck) {
  // dot / sqrt(l2) >= ck
  if (ck >= 0.0) return dot >= 0.0 && dot * dot >= ck * ck * l2;
  return dot >= 0.0 || dot * dot <= ck * ck * l2;
}

__global__ void __launch_bounds__(256) k_chain_hist(DevState d, const double* __restrict__ eb2,
                                                    const double* __restrict__ ee2, int nb,
                                                    unsigned long long* __restrict__ hist) {
#pragma clang fp contract(off)
  const uint32_t a = blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= d.n_amp) return;
  const int r = d.n_rep > 1 ? (int)d.amp_rep[a] : 0;
  const Replica& R = d.reps[r];
  unsigned long long* h = hist + (size_t)r * 4 * nb;
  double x[4][3];
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    const uint32_t sl = d.topo[4 * a + u];
#pragma unroll
    for (int c = 0; c < 3; ++c) x[u][c] = d.pos[c * d.cap + sl];
  }
  double b[3][3], l2[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
#pragma unroll
    for (int c = 0; c < 3; ++c) b[k][c] = x[k + 1][c] - x[k][c];
    min_image(R, b[k]);
    l2[k] = b[k][0] * b[k][0] + b[k][1] * b[k][1] + b[k][2] * b[k][2];
    if (l2[k] >= eb2[0] && l2[k] < eb2[nb]) atomicAdd(&h[(k == 0 ? 0 : 1) * nb + bin_of(eb2, nb, l2[k])], 1ull);
  }
#pragma unroll
  for (int k = 0; k < 2; ++k) {
    const double dot = b[k][0] * b[k + 1][0] + b[k][1] * b[k + 1][1] + b[k][2] * b[k + 1][2];
    const double l = l2[k] * l2[k + 1];
    if (l == 0.0) continue;
    // largest bin whose lower edge the cosine reaches
    int lo = 0, hi = nb - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (cos_ge(dot, l, -1.0 + 2.0 * mid / nb)) lo = mid; else hi = mid - 1;
    }
    atomicAdd(&h[2 * nb + lo], 1ull);
  }
  double e[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) e[c] = x[3][c] - x[0][c];
  min_image(R, e);
  const double le = e[0] * e[0] + e[1] * e[1] + e[2] * e[2];
  if (le >= ee2[0] && le < ee2[nb]) atomicAdd(&h[3 * nb + bin_of(ee2, nb, le)], 1ull);
}

void launch_chain_hist(const DevState& d, const double* eb2, const double* ee2, int nb,
                       unsigned long long* hist, hipStream_t st) {
  if (!d.n_amp) return;
  hipLaunchKernelGGL(k_chain_hist, dim3(div_up(d.n_amp, 256)), dim3(256), 0, st, d, eb2, ee2,
                     nb, hist);
}

// ---------------------------------------------------------------- count map
// Beads whose type bit is set in type_mask, counted per voxel of a g[0] x
// g[1] x g[2] grid of each replica's box (voxel edge box / g), voxel index
// (int)(x / edge) with the last voxel also taking x == box
// (density_map.hpp:13-22); out[rep][g2][g1][g0].
__global__ void __launch_bounds__(256) k_count_map(DevState d, int g0, int g1, int g2,
                                                   uint32_t type_mask,
                                                   unsigned int* __restrict__ map) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= dev_n(d)) return;
  if (!((type_mask >> (d.prop[i] & 31u)) & 1u)) return;
  const int r = d.n_rep > 1 ? find_rep(d.rep_off, d.n_rep, i) : 0;
  const Replica& R = d.reps[r];
  const int g[3] = {g0, g1, g2};
  int id[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const double edge = R.box[a] / (double)g[a];
    int k = (int)(d.pos[a * d.cap + i] / edge);
    k = k < 0 ? 0 : (k >= g[a] ? g[a] - 1 : k);
    id[a] = k;
  }
  atomicAdd(&map[(size_t)r * g0 * g1 * g2 + (size_t)(id[2] * g1 + id[1]) * g0 + id[0]], 1u);
}

void launch_count_map(const DevState& d, int g0, int g1, int g2, uint32_t type_mask,
                      unsigned int* map, hipStream_t st) {
  if (!d.nl) return;
  hipLaunchKernelGGL(k_count_map, dim3(div_up(d.nl, 256)), dim3(256), 0, st, d, g0, g1, g2,
                     type_mask, map);
}

}  // namespace mdpd
