// Initial configurations on the device: the bead geometry of the
// reference's ConfigMaker (src/config_maker.cpp), one thread per bead.
//
// The reference walks amphiphiles in order with one MT19937 stream seeded
// from the wall clock.  Here every amphiphile, solvent bead and velocity
// owns a keyed Saru stream, so a bead's position depends only on its index
// and the seed, and no thread waits for another:
//   amphiphile a:    Saru(a, kTagAmp, seed)  base point, then direction
//   solvent bead k:  Saru(k, kTagSol, seed)  three uniforms per attempt
//   velocity of i:   Saru(i, kTagVel, seed)  two uniforms per component
// A uniform takes 26 bits of each of two draws (52 bits, k); "open" is
// (k + 1/2) 2^-52 in (0, 1) in the role of genrand_real3, "res" is
// k 2^-52 in [0, 1) in the role of genrand_res53.  Both are exact in FP64.
//
// Sphere and cylinder rings: the host lays the rings out in the reference's
// order (outer leaflet, inner leaflet, the half-offset fill of
// MakeSphSheet :196-205 / MakeCylindSheet :251-259) with their polar-angle
// sines and cosines, first amphiphile and count; an amphiphile finds its
// ring by binary search and takes phi = d_phi * (its index on the ring).
//
// FP64 in the reference's operation order, without contraction.
#include <hip/hip_runtime.h>

#include "common.hpp"
#include "config.hpp"
#include "saru.hpp"

namespace mdpd {

namespace {

__host__ __device__ __forceinline__ double cfg_uni(Saru& s, bool open) {
  const uint32_t a = saru_u32(s) >> 6, b = saru_u32(s) >> 6;
  const double k = (double)a * 67108864.0 + (double)b;
  return (open ? k + 0.5 : k) * (1.0 / 4503599627370496.0);
}

__device__ __forceinline__ double wrap(double x, double L) {
#pragma clang fp contract(off)
  // ApplyPBC, config_maker.cpp:74-77
  return x - floor(x / L) * L;
}

__device__ __forceinline__ uint32_t find_ring(const CfgRing* __restrict__ r, uint32_t nr, uint32_t a) {
  uint32_t lo = 0, hi = nr - 1;
  while (lo < hi) {
    const uint32_t mid = (lo + hi + 1) >> 1;
    if (r[mid].start <= a) lo = mid; else hi = mid - 1;
  }
  return lo;
}

// Neumaier-compensated FP64 sums: per block a fixed shuffle tree, then one
// block finishes the block partials in a fixed order, so the result does
// not depend on scheduling.
struct CSum {
  double s, c;
  __device__ __forceinline__ void add(double x) {
#pragma clang fp contract(off)
    const double t = s + x;
    if (fabs(s) >= fabs(x)) c += (s - t) + x;
    else c += (x - t) + s;
    s = t;
  }
  __device__ __forceinline__ void merge(const CSum& o) {
#pragma clang fp contract(off)
    add(o.s);
    c += o.c;
  }
};

// the block's Q sums to out[q * 2 + {0, 1}] (256 threads)
template <int Q>
__device__ __forceinline__ void block_sum(CSum (&v)[Q], double* __restrict__ out) {
  __shared__ double sh[Q][2][4];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    for (int o = 32; o > 0; o >>= 1) {
      CSum u{__shfl_down(v[q].s, o, 64), __shfl_down(v[q].c, o, 64)};
      if (lane < o) v[q].merge(u);
    }
    if (lane == 0) {
      sh[q][0][w] = v[q].s;
      sh[q][1][w] = v[q].c;
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      CSum t{sh[q][0][0], sh[q][1][0]};
      for (int k = 1; k < 4; ++k) t.merge(CSum{sh[q][0][k], sh[q][1][k]});
      out[2 * q] = t.s;
      out[2 * q + 1] = t.c;
    }
  }
}

__device__ __forceinline__ void bead(const CfgLaunch& L, uint32_t i, CSum (&vs)[3]) {
#pragma clang fp contract(off)
  double p[3];
  uint32_t prop, amp, unit;
  const uint32_t amp_beads = L.amp_num * L.all_unit;
  if (i < amp_beads) {
    const uint32_t a = i / L.all_unit;
    unit = i - a * L.all_unit;
    amp = a;
    prop = unit < L.head_unit ? 0u : 1u;  // Hyphil / Hyphob (SetAmphilPartPos :79-94)
    double base[3], nv[3] = {0.0, 0.0, 0.0};
    if (L.mode == MDPD_CFG_FLAT || L.mode == MDPD_CFG_ASYM_FLAT) {
      // MakeFlatSheet :103-138: base uniform in the box, its y replaced by
      // the leaflet's head height; the chain points to the mid plane
      Saru s = saru_seed(a, kTagAmp, L.seed);
      for (int c = 0; c < 3; ++c) base[c] = L.box[c] * cfg_uni(s, true);
      bool up = true;
      if (L.mode == MDPD_CFG_FLAT) {
        up = (a & 1u) == 0u;  // flap toggled after every amphiphile
      } else {
        // flap toggled after placing amphiphile m-1 whenever m % 100 equals
        // rat or 0 (m >= 1): count the toggles among m = 1..a
        uint32_t t = a / 100u;
        if (L.rat >= 1u && L.rat <= 99u && a >= L.rat) t += (a - L.rat) / 100u + 1u;
        up = (t & 1u) == 0u;
      }
      const double sign = up ? 1.0 : -1.0;
      base[1] = 0.5 * L.box[1] + L.span * sign;
      nv[1] = -sign;
    } else if (L.mode == MDPD_CFG_RANDOM) {
      // MakeRandomConfig :264-280 (directions in the positive octant)
      Saru s = saru_seed(a, kTagAmp, L.seed);
      for (int c = 0; c < 3; ++c) base[c] = L.box[c] * cfg_uni(s, false);
      for (int c = 0; c < 3; ++c) nv[c] = cfg_uni(s, false);
      const double nrm = sqrt((nv[0] * nv[0] + nv[1] * nv[1]) + nv[2] * nv[2]);
      for (int c = 0; c < 3; ++c) nv[c] = nv[c] / nrm;
    } else {
      // MakeSphLine :139-161 on the ring holding amphiphile a
      const CfgRing& r = L.rings[find_ring(L.rings, L.n_rings, a)];
      const double phi = r.d_phi * (double)(int32_t)(a - r.start);
      double sp, cp;
      sincos(phi, &sp, &cp);
      nv[0] = r.sign * r.sin_t * cp;
      nv[1] = r.sign * r.sin_t * sp;
      nv[2] = r.sign * r.cos_t;
      const double f = r.rad - r.sign * L.q_thick;
      const double cen[3] = {L.center[0], L.center[1], r.cz};
      for (int c = 0; c < 3; ++c) base[c] = (r.sign * nv[c]) * f + cen[c];
    }
    const double step = L.bond_leng * (double)unit;
    for (int c = 0; c < 3; ++c) p[c] = wrap(base[c] + step * nv[c], L.box[c]);
  } else {
    // solvent: MakeCuboidSolvent :282-293 / MakeSphericalSolvent :295-321
    const uint32_t k = i - amp_beads;
    prop = 2u;
    amp = unit = MDPD_UNIT_NONE;
    Saru s = saru_seed(k, kTagSol, L.seed);
    if (L.sol_kind == kSolCuboid) {
      for (int c = 0; c < 3; ++c) p[c] = (L.sol_hi[c] - L.sol_lo[c]) * cfg_uni(s, true) + L.sol_lo[c];
    } else {
      bool ok = false;
      for (uint32_t t = 0; t < kSolMaxAttempts && !ok; ++t) {
        for (int c = 0; c < 3; ++c) p[c] = L.box[c] * cfg_uni(s, true);
        double d[3];
        for (int c = 0; c < 3; ++c) d[c] = p[c] - 0.5 * L.box[c];
        double dist = (d[0] * d[0] + d[1] * d[1]) + d[2] * d[2];
        if (L.sol_kind == kSolCylinder) dist -= d[L.sol_axis] * d[L.sol_axis];
        dist = sqrt(dist);
        ok = dist < L.sol_rad;
      }
      if (!ok) atomicOr(L.flags, kCfgFlagSolvent);
    }
  }
  // GenVeloc :60-72 before the drift removal: NormalRand(0, sqrt(T)) per
  // component, mean + sd sqrt(-2 ln u1) cos(2 pi u2)
  double v[3];
  {
    Saru s = saru_seed(i, kTagVel, L.seed);
    for (int c = 0; c < 3; ++c) {
      const double u1 = cfg_uni(s, true);
      const double u2 = cfg_uni(s, true);
      v[c] = 0.0 + L.sd * sqrt(-2.0 * log(u1)) * cos(2.0 * M_PI * u2);
    }
  }
  bool bad = false;
  for (int c = 0; c < 3; ++c) {
    // CheckParticleConfigIsValid :425-431
    if (!(p[c] <= L.box[c] && p[c] >= 0.0)) bad = true;
    L.out.pos[(size_t)i * L.out.bead_stride + (size_t)c * L.out.comp_stride] = p[c];
    L.out.vel[(size_t)i * L.out.bead_stride + (size_t)c * L.out.comp_stride] = v[c];
  }
  if (bad) atomicOr(L.flags, kCfgFlagOutOfBox);
  L.out.id[i] = i;
  L.out.prop[i] = prop;
  L.out.amp_id[i] = amp;
  L.out.unit[i] = unit;
  for (int c = 0; c < 3; ++c) vs[c].add(v[c]);
}

// one thread per bead; also the block's sums of v (part[block][3][2])
__global__ void __launch_bounds__(256) k_cfg_beads(CfgLaunch L, double* __restrict__ part) {
#pragma clang fp contract(off)
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  CSum vs[3] = {{0.0, 0.0}, {0.0, 0.0}, {0.0, 0.0}};
  if (i < L.n) bead(L, i, vs);
  block_sum<3>(vs, part + (size_t)blockIdx.x * 6);
}


// out[q] = sum over nb block partials (one block of 256 threads)
template <int Q>
__global__ void __launch_bounds__(256) k_cfg_finish_sum(const double* __restrict__ part, uint32_t nb,
                                                        double* __restrict__ out) {
#pragma clang fp contract(off)
  CSum v[Q];
  for (int q = 0; q < Q; ++q) v[q] = CSum{0.0, 0.0};
  for (uint32_t b = threadIdx.x; b < nb; b += 256)
    for (int q = 0; q < Q; ++q) v[q].merge(CSum{part[((size_t)b * Q + q) * 2], part[((size_t)b * Q + q) * 2 + 1]});
  __shared__ double res[2 * Q];
  block_sum<Q>(v, res);
  __syncthreads();
  if (threadIdx.x < Q) out[threadIdx.x] = res[2 * threadIdx.x] + res[2 * threadIdx.x + 1];
}

// RemoveCMDrift :51-58 (cmvel = sum / n; vel -= cmvel), vel_buf = vel, and
// the block's sum of v.v for the temperature check (part[block][1][2])
__global__ void __launch_bounds__(256) k_cfg_remove_drift(CfgOut o, uint32_t n,
                                                          const double* __restrict__ sum,
                                                          double* __restrict__ part) {
#pragma clang fp contract(off)
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  CSum kin[1] = {{0.0, 0.0}};
  if (i < n) {
    double v[3];
    for (int c = 0; c < 3; ++c) {
      const size_t k = (size_t)i * o.bead_stride + (size_t)c * o.comp_stride;
      v[c] = o.vel[k] - sum[c] / (double)n;
      o.vel[k] = v[c];
      if (o.vbuf) o.vbuf[k] = v[c];
    }
    kin[0].add((v[0] * v[0] + v[1] * v[1]) + v[2] * v[2]);
  }
  block_sum<1>(kin, part + (size_t)blockIdx.x * 2);
}

// a context's chain map and topology for 4-bead amphiphiles in generation
// order (bead 4a + u is unit u of amphiphile a, glob_topol_[4a + u])
__global__ void __launch_bounds__(256) k_cfg_chains(uint32_t n, uint32_t amp_num, double cf_b,
                                                    uint32_t* __restrict__ chain,
                                                    uint32_t* __restrict__ orig,
                                                    uint32_t* __restrict__ topo,
                                                    double* __restrict__ cf_bend,
                                                    uint32_t* __restrict__ amp_rep) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const bool in_chain = i < 4u * amp_num;
  chain[i] = in_chain ? ((i >> 2) << 2 | (i & 3u)) : 0xffffffffu;
  orig[i] = i;
  if (in_chain) topo[i] = i;
  if (i < amp_num) {
    cf_bend[i] = cf_b;
    amp_rep[i] = 0u;
  }
}

inline dim3 grid256(uint32_t n) { return dim3(div_up(n, 256u)); }

}  // namespace

void launch_cfg_generate(const CfgLaunch& L, double* scratch, hipStream_t st) {
  if (!L.n) return;
  const uint32_t nb = div_up(L.n, 256u);
  double* res = scratch;       // [0..2] sum v, [3] sum v.v
  double* part = scratch + 8;  // [nb][3][2]
  hipLaunchKernelGGL(k_cfg_beads, dim3(nb), dim3(256), 0, st, L, part);
  hipLaunchKernelGGL(k_cfg_finish_sum<3>, dim3(1), dim3(256), 0, st, part, nb, res);
  hipLaunchKernelGGL(k_cfg_remove_drift, dim3(nb), dim3(256), 0, st, L.out, L.n, res, part);
  hipLaunchKernelGGL(k_cfg_finish_sum<1>, dim3(1), dim3(256), 0, st, part, nb, res + 3);
}

void launch_cfg_chains(uint32_t n, uint32_t amp_num, double cf_b, uint32_t* chain, uint32_t* orig,
                       uint32_t* topo, double* cf_bend, uint32_t* amp_rep, hipStream_t st) {
  if (!n) return;
  hipLaunchKernelGGL(k_cfg_chains, grid256(n), dim3(256), 0, st, n, amp_num, cf_b, chain, orig, topo,
                     cf_bend, amp_rep);
}

}  // namespace mdpd
