// Non-bonded pair passes: the density pass (CalcDensity,
// nbdf_calculator.hpp:11-61), which also records the pair list, and the
// force pass (CalcForceEpEpDPD, nbdf_calculator.hpp:63-145).
//
// One thread per bead in cell order, gathering over the 27 cells around its
// own; no Newton-3 and no force atomics: every bead sums its own pairs, as
// the reference does.  Positions are FP32 offsets inside the bead's cell, so
// a pair displacement is (l_i - o*h) - l_j with o the stencil offset,
// accurate to a few 1e-7; pairs whose r^2 lies within kGuard of a cutoff are
// decided again in FP64 from the master positions with the oracle's
// minimum-image expression, so the pair sets are those of the FP64 oracle.
//
// The density pass writes, per 64-bead tile (one wavefront), the in-range
// neighbours (slot | stencil code << 26) in chunks of four: chunk c of lane
// l sits at [tile][c][l] (16 bytes), so every store is a whole 16-byte
// chunk and the force pass reads a wave's chunk row as one contiguous 1 KiB
// block.  The force pass walks that list, so the Saru/log/cos/sqrt work runs
// only on pairs inside the cutoff; in the step it ends with the bead's own
// bonded chain terms and the kick, so the acceleration never makes a round
// trip through memory between the two.
#include <hip/hip_runtime.h>
#include <math.h>

#include "common.hpp"
#include "saru.hpp"
#include "state.hpp"
#include "bonded.hpp"
#include "step_dev.hpp"

namespace mdpd {

__device__ __forceinline__ double fp64_r2(const DevState& d, const Replica& R, uint32_t i,
                                          uint32_t j) {
#pragma clang fp contract(off)
  double x[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    double v = d.pos[a * d.cap + i] - d.pos[a * d.cap + j];
    if (v > 0.5 * R.box[a]) v -= R.box[a];
    else if (v < -0.5 * R.box[a]) v += R.box[a];
    x[a] = v;
  }
  return x[0] * x[0] + x[1] * x[1] + x[2] * x[2];
}

// displacement pos_i - pos_j, minimum image, FP64
__device__ __forceinline__ void fp64_dvec(const DevState& d, const Replica& R, uint32_t i,
                                          uint32_t j, double v[3]) {
#pragma clang fp contract(off)
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    double x = d.pos[a * d.cap + i] - d.pos[a * d.cap + j];
    if (x > 0.5 * R.box[a]) x -= R.box[a];
    else if (x < -0.5 * R.box[a]) x += R.box[a];
    v[a] = x;
  }
}

// force pass: pairs closer than sqrt(kClose2) take their displacement in FP64
constexpr float kClose2 = 1.0e-3f;

// Pair decision shared by every pass: returns (r^2 < cut2) as the FP64
// oracle decides it.  r2 is the FP32 estimate.
__device__ __forceinline__ bool decide(float r2, float cut2f, double cut2, const DevState& d,
                                       const Replica& R, uint32_t s, uint32_t t) {
  if (fabsf(r2 - cut2f) > kGuard) return r2 < cut2f;
  return fp64_r2(d, R, s, t) < cut2;
}

// element t of a slot-indexed array through a 32-bit byte offset (slots are
// below 2^26): the load takes the scalar-base + vector-offset form
__device__ __forceinline__ float4 ld_slot(const float4* __restrict__ base, uint32_t t) {
  return *reinterpret_cast<const float4*>(reinterpret_cast<const char*>(base) + (t << 4));
}

// the bead's densities: one float with one species, else (rho_0, rho_1,
// rho_2, 0) as one 16-byte record, so the force pass gathers a neighbour's
// three densities with one load
template <int ND>
__device__ __forceinline__ void store_dens(const DevState& d, uint32_t s, const float (&rho)[ND]) {
  if (ND == 1) d.sdens[s] = rho[0];
  else reinterpret_cast<float4*>(d.sdens)[s] = make_float4(rho[0], rho[1 % ND], rho[2 % ND], 0.f);
}

__device__ __forceinline__ const uint4* tile_list(const DevState& d, uint32_t s) {
  return reinterpret_cast<const uint4*>(d.nlist) + (size_t)(s >> 6) * ((uint32_t)d.W >> 2) * 64u +
         (s & 63u);
}

// Skin-list build of skin mode (rebuild steps only): every bead but the
// bead itself with FP32 r^2 < vcut2 goes to the bead's skin list, written
// like the pair list ([tile][chunk][lane] chunks of four entries).  vcut2
// carries a margin over (1 + skin)^2 that absorbs the FP32 error, and cells
// are at least 1 + skin wide, so no pair that can come inside the cutoff
// before the next rebuild is left out.
template <bool MULTI>
__global__ void __launch_bounds__(256) k_skin_build(DevState d) {
  if (*d.rebuild == 0u) return;
  const uint32_t s = xcd_block(blockIdx.x, live_tiles(d, blockDim.x)) * blockDim.x + threadIdx.x;
  if (s >= dev_n(d)) return;
  const int r = MULTI ? find_rep(d.rep_off, d.n_rep, s) : 0;
  const Replica& R = d.reps[r];
  const int nx = R.nc[0], ny = R.nc[1], nz = R.nc[2];
  const uint32_t cell_off = R.cell_off;
  const float hx = R.hf[0], hy = R.hf[1], hz = R.hf[2];
  const float cut_hi = d.vcut2;
  const float4 pi = d.spos[s];
  const uint32_t c = d.scell[s] - cell_off;
  const int cx = (int)(c % (uint32_t)nx);
  const int cy = (int)((c / (uint32_t)nx) % (uint32_t)ny);
  const int cz = (int)(c / (uint32_t)(nx * ny));
  const uint32_t W = (uint32_t)d.Wv;
  uint4* lst = reinterpret_cast<uint4*>(d.vlist) + (size_t)(s >> 6) * (W >> 2) * 64u + (s & 63u);
  // the bead's fixed-point x and the periodic correction of x
  const uint32_t Xi = __float_as_uint(pi.x), Wx = (uint32_t)nx << 24;
  const float hs = hx * kFixStep;
  uint32_t cnt = 0, e0 = 0, e1 = 0, e2 = 0;
  const bool wrap = !(cx >= 1 && cx <= nx - 2);
  for (int oz = -1; oz <= 1; ++oz) {
    int qz = cz + oz;
    qz += (qz < 0) ? nz : 0;
    qz -= (qz >= nz) ? nz : 0;
    const float bz = pi.z - (float)oz * hz;
    for (int oy = -1; oy <= 1; ++oy) {
      int qy = cy + oy;
      qy += (qy < 0) ? ny : 0;
      qy -= (qy >= ny) ? ny : 0;
      const float by = pi.y - (float)oy * hy;
      const uint32_t* S = d.cell_start + cell_off + (uint32_t)((qz * ny + qy) * nx);
      const uint32_t cyz = (uint32_t)(((oy + 1) << 2) | ((oz + 1) << 4)) << kCodeShift;
      for (int pc = 0; pc < (wrap ? 2 : 1); ++pc) {
        // run [ta, tb) over the row's three cells; the piece across the
        // periodic x boundary (if any) takes the box correction
        uint32_t ta, tb, fx = 1u;
        if (!wrap) { ta = S[cx - 1]; tb = S[cx + 2]; }
        else if (cx == 0) {
          if (pc == 0) { ta = S[nx - 1]; tb = S[nx]; fx = 0u; }
          else { ta = S[0]; tb = S[2]; }
        } else {
          if (pc == 0) { ta = S[nx - 2]; tb = S[nx]; }
          else { ta = S[0]; tb = S[1]; fx = 2u; }
        }
        const uint32_t Xp = Xi + fix_wrap(fx, Wx), C = cyz | (fx << kCodeShift);
        for (uint32_t t = ta; t < tb; ++t) {
          const float4 pj = d.spos[t];
          const float dx = fix_dx(Xp, __float_as_uint(pj.x), hs), dy = by - pj.y, dz = bz - pj.z;
          const float r2 = dx * dx + dy * dy + dz * dz;
          if (r2 >= cut_hi || t == s) continue;
          const uint32_t e = t | C;
          if ((cnt & 3u) == 3u) {
            if (cnt < W) lst[(size_t)(cnt >> 2) * 64u] = make_uint4(e0, e1, e2, e);
          } else {
            e0 = e1;
            e1 = e2;
            e2 = e;
          }
          ++cnt;
        }
      }
    }
  }
  // the last (cnt % 4) entries sit at the right end of the shift register
  const uint32_t tail = cnt & 3u;
  if (tail != 0u && cnt < W) {
    const uint4 ch = tail == 1u ? make_uint4(e2, 0u, 0u, 0u)
                   : tail == 2u ? make_uint4(e1, e2, 0u, 0u)
                                : make_uint4(e0, e1, e2, 0u);
    lst[(size_t)(cnt >> 2) * 64u] = ch;
  }
  d.vcount[s] = cnt;
  if (cnt > W) atomicOr(d.flags, FLAG_NLIST_OVERFLOW);
}

// Density pass, queued form (the every-step pass).  The candidate scan
// only forms the FP32 r^2; a candidate inside cut_hi goes into the lane's
// queue in LDS as (entry, r^2), kDensQ deep.  Once some lane of the wave has
// fewer than kScanStep free places (the candidates of one scan iteration),
// and after the scan, every active lane drains its own queue in order: the
// guard-band decisions, the density weight and the
// pair-list write run there with most lanes busy, rather than inside the
// scan with only the lanes whose candidate is in range (about one in six).
// Each lane still meets its in-range candidates in scan order, so the FP32
// sums and the pair list are those of a plain per-candidate loop.  The
// scan takes four candidates per iteration (four gathers in flight per
// wave; 80 VGPRs, 6 waves per SIMD).  Ten entries of 8 bytes per lane are
// 5 KB of LDS per wave, which costs no wave.
constexpr uint32_t kDensQ = 10;
constexpr uint32_t kScanStep = 4;  // candidates per scan iteration

// The queued r^2 word.  A kept candidate has r^2 < 1 + kGuard < 2 (rd <= 1),
// so the float's two top bits (sign and the exponent's top bit) are zero;
// with several density species they carry the candidate's type (.w & 3 of
// its FP32 copy), which the drain then needs no gather for.
template <int ND>
__device__ __forceinline__ uint32_t queue_r2(float r2, float w) {
  if (ND == 1) return __float_as_uint(r2);
  return __float_as_uint(r2) | (__float_as_uint(w) << 30);
}

template <int ND, bool MULTI, bool RDLE1, int BS>
__global__ void __launch_bounds__(BS) k_density_q(DevState d) {
  __shared__ uint2 dens_queue[kDensQ * BS];  // [kDensQ][BS]
  if (halted(d)) return;
  const uint32_t s = xcd_block(blockIdx.x, live_tiles(d, BS)) * BS + threadIdx.x;
  if (s >= dev_n(d)) return;
  const int r = MULTI ? find_rep(d.rep_off, d.n_rep, s) : 0;
  const Replica& R = d.reps[r];
  const int nx = R.nc[0], ny = R.nc[1], nz = R.nc[2];
  const uint32_t cell_off = R.cell_off;
  const float hx = R.hf[0], hy = R.hf[1], hz = R.hf[2];
  const float rd2f = R.rd2f, inv_rd = R.inv_rd, crho = R.crho;
  const float cut_hi = fmaxf(1.0f, rd2f) + kGuard;
  const float4 pi = d.spos[s];
  const uint32_t c = d.scell[s] - cell_off;
  const int cx = (int)(c % (uint32_t)nx);
  const int cy = (int)((c / (uint32_t)nx) % (uint32_t)ny);
  const int cz = (int)(c / (uint32_t)(nx * ny));
  const uint32_t W = (uint32_t)d.W;
  uint4* lst = const_cast<uint4*>(tile_list(d, s));
  // the lane's queue entries sit BS entries apart from byte offset qi;
  // [qi, qe) is in use, and past qhi fewer than two places are free
  constexpr uint32_t qs = BS * (uint32_t)sizeof(uint2);
  const uint32_t qi = threadIdx.x * (uint32_t)sizeof(uint2);
  const uint32_t qhi = qi + (kDensQ - kScanStep) * qs;
  uint32_t qe = qi;
  char* const qbase = reinterpret_cast<char*>(dens_queue);
  float rho[ND];
#pragma unroll
  for (int k = 0; k < ND; ++k) rho[k] = 0.f;
  uint32_t cnt = 0, e0 = 0, e1 = 0, e2 = 0;
  // the queued candidates, in order: decisions, weights, pair-list entries
  auto drain = [&]() {
    for (uint32_t k = qi; k < qe; k += qs) {
      const uint2 en = *reinterpret_cast<const uint2*>(qbase + k);
      const uint32_t e = en.x, t = e & kIndexMask;
      float r2 = __uint_as_float(ND == 1 ? en.y : (en.y & 0x3FFFFFFFu));
      // outside the guard bands the FP32 r^2 decides; the clamp of w at 0
      // is then the rd test, and with RDLE1 no cutoff compare is needed
      bool in_rc = RDLE1 ? true : (r2 < 1.0f);
      const bool near = (fabsf(r2 - 1.0f) <= kGuard) | (fabsf(r2 - rd2f) <= kGuard) |
                        (r2 < 1e-6f);
      if (near) {
        if (t == s) continue;  // the bead itself
        const double r2d = fp64_r2(d, R, s, t);
        in_rc = r2d < 1.0;
        if (!(r2d < R.rd2 && r2d != 0.0)) r2 = 2.0f * rd2f + 1.0f;
      }
      const float rr = __builtin_amdgcn_sqrtf(r2);
      const float w = fmaxf(1.0f - rr * inv_rd, 0.0f);
      const float v = crho * w * w;
      if (ND == 1) rho[0] += v;
      else {
        const uint32_t tj = en.y >> 30;  // the candidate's type, from the scan
#pragma unroll
        for (int k2 = 0; k2 < ND; ++k2) rho[k2] += (tj == (uint32_t)k2) ? v : 0.f;
      }
      if (in_rc) {
        // three-deep shift register; every fourth entry the window plus the
        // entry that fell out go out as one 16-byte chunk (past W entries
        // the last chunk is overwritten and the overflow flag fails the call)
        const uint32_t out = e0;
        e0 = e1;
        e1 = e2;
        e2 = e;
        if ((cnt & 3u) == 3u)
          lst[(size_t)min(cnt >> 2, (W >> 2) - 1u) * 64u] = make_uint4(out, e0, e1, e2);
        ++cnt;
      }
    }
    qe = qi;
  };
  // the bead's fixed-point x and the periodic correction of x
  const uint32_t Xi = __float_as_uint(pi.x), Wx = (uint32_t)nx << 24;
  const float hs = hx * kFixStep;
  const bool wrap = !(cx >= 1 && cx <= nx - 2);
  for (int oz = -1; oz <= 1; ++oz) {
    int qz = cz + oz;
    qz += (qz < 0) ? nz : 0;
    qz -= (qz >= nz) ? nz : 0;
    const float bz = pi.z - (float)oz * hz;
    for (int oy = -1; oy <= 1; ++oy) {
      int qy = cy + oy;
      qy += (qy < 0) ? ny : 0;
      qy -= (qy >= ny) ? ny : 0;
      const float by = pi.y - (float)oy * hy;
      const uint32_t rowq = cell_off + (uint32_t)((qz * ny + qy) * nx);
      const uint32_t* S = d.cell_start + rowq;
      const uint32_t cyz = (uint32_t)(((oy + 1) << 2) | ((oz + 1) << 4)) << kCodeShift;
      for (int pc = 0; pc < (wrap ? 2 : 1); ++pc) {
        // run [ta, tb) over the row's three cells; the piece across the
        // periodic x boundary (if any) takes the box correction
        uint32_t ta, tb, fx = 1u;
        if (!wrap) { ta = S[cx - 1]; tb = S[cx + 2]; }
        else if (cx == 0) {
          if (pc == 0) { ta = S[nx - 1]; tb = S[nx]; fx = 0u; }
          else { ta = S[0]; tb = S[2]; }
        } else {
          if (pc == 0) { ta = S[nx - 2]; tb = S[nx]; }
          else { ta = S[0]; tb = S[1]; fx = 2u; }
        }
        const uint32_t Xp = Xi + fix_wrap(fx, Wx), C = cyz | (fx << kCodeShift);
        // kScanStep candidates per iteration, as many gathers in flight per
        // wave; past the run's end the slot repeats the run's last one and
        // is not queued.  The queue keeps room for kScanStep more (drained
        // when fewer places are free).
        for (uint32_t t = ta; t < tb; t += kScanStep) {
          uint32_t tq[kScanStep];
          float4 pq[kScanStep];
#pragma unroll
          for (uint32_t q = 0; q < kScanStep; ++q) {
            tq[q] = min(t + q, tb - 1u);
            pq[q] = ld_slot(d.spos, tq[q]);
          }
#pragma unroll
          for (uint32_t q = 0; q < kScanStep; ++q) {
            const float dx = fix_dx(Xp, __float_as_uint(pq[q].x), hs), dy = by - pq[q].y,
                        dz = bz - pq[q].z;
            const float r2 = dx * dx + dy * dy + dz * dz;
            // written always (the queue has room), kept only in range and
            // inside the run: no exec-mask branch per candidate
            *reinterpret_cast<uint2*>(qbase + qe) = make_uint2(tq[q] | C, queue_r2<ND>(r2, pq[q].w));
            qe += (r2 < cut_hi && t + q < tb) ? qs : 0u;
          }
          if (__ballot(qe > qhi) != 0ull) drain();
        }
      }
    }
  }
  drain();
  const uint32_t tail = cnt & 3u;
  if (tail != 0u && cnt < W) {
    const uint4 ch = tail == 1u ? make_uint4(e2, 0u, 0u, 0u)
                   : tail == 2u ? make_uint4(e1, e2, 0u, 0u)
                                : make_uint4(e0, e1, e2, 0u);
    lst[(size_t)(cnt >> 2) * 64u] = ch;
  }
  store_dens<ND>(d, s, rho);
  if (ND == 1) reinterpret_cast<float*>(&d.sb[s].vel)[3] = rho[0];
  d.ncount[s] = cnt;
  if (cnt > W) atomicOr(d.flags, FLAG_NLIST_OVERFLOW);
}

template <int ND, bool MULTI>
static void dens_launch(const DevState& d, int block, hipStream_t st) {
  // the queue lives in a static LDS array sized by the block
  const dim3 g(div_up(d.nl, block));
  if (d.rd_le_1) {
    if (block == 64) hipLaunchKernelGGL((k_density_q<ND, MULTI, true, 64>), g, dim3(64), 0, st, d);
    else if (block == 128) hipLaunchKernelGGL((k_density_q<ND, MULTI, true, 128>), g, dim3(128), 0, st, d);
    else hipLaunchKernelGGL((k_density_q<ND, MULTI, true, 256>), g, dim3(256), 0, st, d);
  } else {
    if (block == 64) hipLaunchKernelGGL((k_density_q<ND, MULTI, false, 64>), g, dim3(64), 0, st, d);
    else if (block == 128) hipLaunchKernelGGL((k_density_q<ND, MULTI, false, 128>), g, dim3(128), 0, st, d);
    else hipLaunchKernelGGL((k_density_q<ND, MULTI, false, 256>), g, dim3(256), 0, st, d);
  }
}

void launch_skin_build(const DevState& d, hipStream_t st) {
  if (!d.nl) return;
  const int block = 128;
  if (d.n_rep > 1)
    hipLaunchKernelGGL((k_skin_build<true>), dim3(div_up(d.nl, block)), dim3(block), 0, st, d);
  else
    hipLaunchKernelGGL((k_skin_build<false>), dim3(div_up(d.nl, block)), dim3(block), 0, st, d);
}

// Density pass of skin mode: the same decisions and sums as k_density_q, over
// the skin list instead of the 27 cells (positions relative to the cells of
// the last build, with the list's stencil codes), writing the same pair
// list for the force pass.
template <int ND, bool MULTI>
__global__ void __launch_bounds__(256) k_density_skin(DevState d) {
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    // every rebuild kernel of this step has run: count the build and clear
    // the word for the next step's displacement test
    if (d.rebuild[0]) d.rebuild[1] += 1u;
    d.rebuild[0] = 0u;
  }
  const uint32_t s = xcd_block(blockIdx.x, live_tiles(d, blockDim.x)) * blockDim.x + threadIdx.x;
  if (s >= dev_n(d)) return;
  const int r = MULTI ? find_rep(d.rep_off, d.n_rep, s) : 0;
  const Replica& R = d.reps[r];
  const float hx = R.hf[0], hy = R.hf[1], hz = R.hf[2];
  const float rd2f = R.rd2f, inv_rd = R.inv_rd, crho = R.crho;
  const float cut_hi = fmaxf(1.0f, rd2f) + kGuard;
  const float4 pi = d.spos[s];
  const uint32_t Xi = __float_as_uint(pi.x), Wx = (uint32_t)R.nc[0] << 24;
  const float hs = hx * kFixStep;
  const uint32_t W = (uint32_t)d.W, Wv = (uint32_t)d.Wv;
  uint4* lst = const_cast<uint4*>(tile_list(d, s));
  const uint4* vl = reinterpret_cast<const uint4*>(d.vlist) + (size_t)(s >> 6) * (Wv >> 2) * 64u +
                    (s & 63u);
  uint32_t vc = d.vcount[s];
  vc = vc < Wv ? vc : Wv;
  float rho[ND];
#pragma unroll
  for (int k = 0; k < ND; ++k) rho[k] = 0.f;
  uint32_t cnt = 0, e0 = 0, e1 = 0, e2 = 0;
  auto cand = [&](uint32_t e) {
    const uint32_t t = e & kIndexMask;
    const uint32_t code = e >> kCodeShift;
    const float4 pj = d.spos[t];
    const float dx = fix_dx(Xi + fix_wrap(code & 3u, Wx), __float_as_uint(pj.x), hs);
    const float dy = (pi.y - (float)((int)((code >> 2) & 3u) - 1) * hy) - pj.y;
    const float dz = (pi.z - (float)((int)(code >> 4) - 1) * hz) - pj.z;
    const float r2 = dx * dx + dy * dy + dz * dz;
    if (r2 >= cut_hi) return;
    bool in_rc = r2 < 1.0f, in_rd = r2 < rd2f;
    if (fabsf(r2 - 1.0f) <= kGuard || fabsf(r2 - rd2f) <= kGuard || r2 < 1e-6f) {
      const double r2d = fp64_r2(d, R, s, t);
      in_rc = r2d < 1.0;
      in_rd = r2d < R.rd2 && r2d != 0.0;
    }
    if (in_rd) {
      const float rr = __builtin_amdgcn_sqrtf(r2);
      const float w = fmaxf(1.0f - rr * inv_rd, 0.0f);
      const float v = crho * w * w;
      if (ND == 1) rho[0] += v;
      else {
        const uint32_t tj = __float_as_uint(pj.w) & 3u;
#pragma unroll
        for (int q = 0; q < ND; ++q) rho[q] += (tj == (uint32_t)q) ? v : 0.f;
      }
    }
    if (in_rc) {
      if ((cnt & 3u) == 3u) {
        if (cnt < W) lst[(size_t)(cnt >> 2) * 64u] = make_uint4(e0, e1, e2, e);
      } else {
        e0 = e1;
        e1 = e2;
        e2 = e;
      }
      ++cnt;
    }
  };
  for (uint32_t k = 0; k < vc; k += 4) {
    const uint4 ch = vl[(size_t)(k >> 2) * 64u];
    cand(ch.x);
    if (k + 1 < vc) cand(ch.y);
    if (k + 2 < vc) cand(ch.z);
    if (k + 3 < vc) cand(ch.w);
  }
  const uint32_t tail = cnt & 3u;
  if (tail != 0u && cnt < W) {
    const uint4 ch = tail == 1u ? make_uint4(e2, 0u, 0u, 0u)
                   : tail == 2u ? make_uint4(e1, e2, 0u, 0u)
                                : make_uint4(e0, e1, e2, 0u);
    lst[(size_t)(cnt >> 2) * 64u] = ch;
  }
  store_dens<ND>(d, s, rho);
  if (ND == 1) reinterpret_cast<float*>(&d.sb[s].vel)[3] = rho[0];
  d.ncount[s] = cnt;
  if (cnt > W) atomicOr(d.flags, FLAG_NLIST_OVERFLOW);
}

void launch_density_skin(const DevState& d, int nd, int block, hipStream_t st) {
  if (!d.nl) return;
  const dim3 g(div_up(d.nl, block)), b(block);
  const bool multi = d.n_rep > 1;
  if (nd == 1) {
    if (multi) hipLaunchKernelGGL((k_density_skin<1, true>), g, b, 0, st, d);
    else hipLaunchKernelGGL((k_density_skin<1, false>), g, b, 0, st, d);
  } else {
    if (multi) hipLaunchKernelGGL((k_density_skin<3, true>), g, b, 0, st, d);
    else hipLaunchKernelGGL((k_density_skin<3, false>), g, b, 0, st, d);
  }
}

void launch_density(const DevState& d, int /*ns*/, int nd, int block, hipStream_t st) {
  if (!d.nl) return;
  const bool multi = d.n_rep > 1;
  if (nd == 1) multi ? dens_launch<1, true>(d, block, st) : dens_launch<1, false>(d, block, st);
  else multi ? dens_launch<3, true>(d, block, st) : dens_launch<3, false>(d, block, st);
}

// ---------------------------------------------------------------- force
// Near the cutoff the FP32 value of 1 - r can decide whether a bead's force
// meets 1e-5 of its pair-force sum.  The FP32 displacement carries about
// 1e-7 of absolute error in r (kDrEst bounds it with a margin of 2.5), and a
// pair's force term moves by its coefficient times that error over the
// derivative of its weight: relative to the term, 1e-7 / (1 - r) under the
// linear weights and 1e-7 / (2 (1 - r)) under sqrt(1 - r), above 4e-6 only
// within 0.025 of the cutoff.  There (the band: 7% of the pairs) a bead
// with few other neighbours can miss the bar: under the linear envelope the
// random force sigma' xi (1 - r) (sigma' = sigma / sqrt(dt), about 42 at
// dt = 0.005) is off by sigma' |xi| 1e-7 absolute, more than 1e-5 of the
// forces of a vapour bead with one neighbour.  The pair loop sums, over the
// band pairs, a bound of |dF/dr| (the random, conservative and dissipative
// coefficients times the derivatives of their weights) and, over all pairs,
// |F_ij|; a bead whose band error estimate kDrEst * sum |dF/dr| exceeds half
// of 1e-5 * max(sum |F_ij|, 1e-3) walks its list again with 1 - r of every
// band pair from the FP64 master positions.  Only isolated beads (vapour, a
// few neighbours) fail the test; in a liquid the band error is orders of
// magnitude below the bar.
constexpr float kNearCut = 2.5e-2f;
constexpr float kDrEst = 2.5e-7f;

// 1 - r of the pair (s, t) from the FP64 master positions, minimum image,
// for the FP32 r = rr: 1 - r = (1 - r^2) / (1 + r), where only the
// cancelling 1 - r^2 needs FP64
__device__ __forceinline__ float fp64_one_minus_r(const DevState& d, const Replica& R,
                                                  const double* pid, uint32_t t, float rr) {
  double v2 = 0.0;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    double x = pid[a] - d.pos[a * d.cap + t];
    if (x > 0.5 * R.box[a]) x -= R.box[a];
    else if (x < -0.5 * R.box[a]) x += R.box[a];
    v2 += x * x;
  }
  return fmaxf((float)(1.0 - v2) / (1.0f + rr), 0.0f);
}

template <int NS>
__device__ __forceinline__ float sel(const float (&v)[NS], uint32_t t) {
  if (NS == 1) return v[0];
  return t == 0u ? v[0] : (t == 1u ? v[1] : v[2]);
}

// Pair coefficients of one system by type pair (3 ti + tj):
// [0] = (a0, sigma, gamma, mb[0]), [1] = (c1, mb[1], mb[2], 0).
__device__ __forceinline__ void fill_pair_table(const Replica& R, float4* tab) {
  const uint32_t k = threadIdx.x;
  if (k < 9u) {
    const uint32_t a = k / 3u, b = k % 3u;
    tab[2 * k] = make_float4(R.a0[a][b], R.sigma[a][b], R.gamma[a][b], R.mb[a][b][0]);
    tab[2 * k + 1] = make_float4(R.c1[a][b], R.mb[a][b][1], R.mb[a][b][2], 0.0f);
  }
}

// keys (step mode only): the epilogue also bins the bead's position after
// the next step's drift (the same FP64 drift as k_drift, from the kicked
// velocity and the acceleration just stored) into d.key and the cell
// histogram, which the next step's scan reads instead of a key pass.
template <int NS, int ND, int ENV, bool MULTI, bool VIRIAL>
__global__ void __launch_bounds__(256) k_force(DevState d, double* __restrict__ vir_bead,
                                               int conv, int keys) {
  if (halted(d)) return;
  // single system: the coefficients come from a block-shared table indexed
  // by the type pair instead of per-lane selects
  __shared__ float4 ptab[18];
  // single system: per stencil code, the periodic x correction of the
  // fixed-point x (as raw bits) and the y, z image offsets o h (o in
  // {-1, 0, 1}, exact in FP32), so pi - o h rounds once
  __shared__ float4 otab[64];
  if (!MULTI) {
    if (NS > 1) fill_pair_table(d.reps[0], ptab);
    if (threadIdx.x < 64u) {
      const uint32_t k = threadIdx.x;
      const Replica& R0 = d.reps[0];
      otab[k] = make_float4(__uint_as_float(fix_wrap(k & 3u, (uint32_t)R0.nc[0] << 24)),
                            (float)((int)((k >> 2) & 3u) - 1) * R0.hf[1],
                            (float)((int)(k >> 4) - 1) * R0.hf[2], 0.0f);
    }
    __syncthreads();
  }
  const uint32_t s = xcd_block(blockIdx.x, live_tiles(d, blockDim.x)) * blockDim.x + threadIdx.x;
  const bool active = s < dev_n(d);
  // with keys every lane of the wave reaches the aggregated histogram
  // increment at the end
  if (!active && !keys) return;
  uint32_t key = 0;
  const int r = active ? (MULTI ? find_rep(d.rep_off, d.n_rep, s) : 0) : 0;
  if (active) {
  const Replica& R = d.reps[r];
  const uint32_t seed3 = R.pair_seed + *d.d_time;
  const float4 pi = d.sb[s].pos;
  const float4 vi = d.sb[s].vel;
  const uint32_t wi = __float_as_uint(pi.w);
  const uint32_t idi = wi >> 2, ti = NS == 1 ? 0u : (wi & 3u), ti3 = 3u * ti;
  float rho_i[ND];
  if (ND == 1) {
    rho_i[0] = d.sdens[s];
  } else {
    const float4 ri = reinterpret_cast<const float4*>(d.sdens)[s];
    rho_i[0] = ri.x;
    rho_i[1 % ND] = ri.y;
    rho_i[2 % ND] = ri.z;
  }
  float A0[NS], C1[NS], SG[NS], GM[NS], MB[ND][NS];
#pragma unroll
  for (int q = 0; q < NS; ++q) {
    A0[q] = R.a0[ti][q];
    C1[q] = R.c1[ti][q];
    SG[q] = R.sigma[ti][q];
    GM[q] = R.gamma[ti][q];
#pragma unroll
    for (int k = 0; k < ND; ++k) MB[k][q] = R.mb[ti][q][k];
  }
  const float hx = R.hf[0], hy = R.hf[1], hz = R.hf[2];
  const uint32_t Xi = __float_as_uint(pi.x), Wx = (uint32_t)R.nc[0] << 24;
  const float hs = hx * kFixStep;
  const float arc = R.arc, inv_rdm = R.inv_rdm;
  const bool rdm1 = R.rdm_is_one != 0;
  const bool has_c1 = R.has_c1 != 0;
  const uint32_t W = (uint32_t)d.W;
  uint32_t cnt = d.ncount[s];
  cnt = cnt < W ? cnt : W;
  const uint4* lst = tile_list(d, s);
  float fx = 0.f, fy = 0.f, fz = 0.f, px = 0.f, py = 0.f, pz = 0.f;
  // band error estimate (sum of |dF/dr| over band pairs) and sum of |F_ij|
  float band_e = 0.f, fsum = 0.f;
  double pid[3];  // the bead's FP64 master position (exact walk only)
  // exact: 1 - r of band pairs from the FP64 master positions
  // pj, vj: the neighbour's record sb[e & kIndexMask]
  auto pair = [&](uint32_t e, const float4 pj, const float4 vj, bool exact) {
    const uint32_t t = e & kIndexMask;
    const uint32_t code = e >> kCodeShift;
    float dx, dy, dz;
    if (!MULTI) {
      const float4 o = otab[code];
      dx = fix_dx(Xi + __float_as_uint(o.x), __float_as_uint(pj.x), hs);
      dy = (pi.y - o.y) - pj.y;
      dz = (pi.z - o.z) - pj.z;
    } else {
      dx = fix_dx(Xi + fix_wrap(code & 3u, Wx), __float_as_uint(pj.x), hs);
      dy = (pi.y - (float)((int)((code >> 2) & 3u) - 1) * hy) - pj.y;
      dz = (pi.z - (float)((int)(code >> 4) - 1) * hz) - pj.z;
    }
    float r2 = dx * dx + dy * dy + dz * dz;
    if (r2 < kClose2) {
      // a near-coincident pair: the FP32 coordinates' absolute error (a few
      // 1e-7) is too coarse for the direction of so short a displacement;
      // take it from the FP64 master positions (rare)
      uint32_t tt = t;
      asm volatile("" : "+v"(tt));
      double v[3];
      fp64_dvec(d, R, s, tt, v);
      dx = (float)v[0];
      dy = (float)v[1];
      dz = (float)v[2];
      r2 = dx * dx + dy * dy + dz * dz;
    }
    const float rinv = __builtin_amdgcn_rsqf(r2);
    const float rr = r2 * rinv;
    const uint32_t wj = __float_as_uint(pj.w);
    const uint32_t idj = wj >> 2, tj = NS == 1 ? 0u : (wj & 3u);
    Saru sr = saru_seed(idi < idj ? idi : idj, idi < idj ? idj : idi, seed3);
    const float xi = saru_nrml_f32(sr);
    float wc = fmaxf(1.0f - rr, 0.0f);
    const bool band = wc < kNearCut;
    if (exact && band) wc = fp64_one_minus_r(d, R, pid, t, rr);
    float wR, wD;
    if (ENV == 0) { wR = __builtin_amdgcn_sqrtf(wc); wD = wc; }
    else { wR = wc; wD = wc * wc; }
    float a0, c1, sg, gm, mbk[ND];
    if (!MULTI && NS > 1) {
      const float4 q0 = ptab[2 * (ti3 + tj)];
      a0 = q0.x;
      sg = q0.y;
      gm = q0.z;
      mbk[0] = q0.w;
      if (ND > 1 || has_c1) {
        const float4 q1 = ptab[2 * (ti3 + tj) + 1];
        c1 = q1.x;
        if (ND > 1) {
          mbk[1 % ND] = q1.y;
          mbk[2 % ND] = q1.z;
        }
      } else {
        c1 = 0.0f;
      }
    } else {
      a0 = sel<NS>(A0, tj);
      c1 = sel<NS>(C1, tj);
      sg = sel<NS>(SG, tj);
      gm = sel<NS>(GM, tj);
#pragma unroll
      for (int k = 0; k < ND; ++k) mbk[k] = sel<NS>(MB[k], tj);
    }
    const float cons_c = has_c1 ? a0 + c1 * fmaxf(rr - arc, 0.0f) : a0;
    float mbs = 0.f;
    if (ND == 1) {
      mbs = mbk[0] * (rho_i[0] + vj.w);
    } else {
      {
        const float4 dj = reinterpret_cast<const float4*>(d.sdens)[t];
        const float rj[3] = {dj.x, dj.y, dj.z};
#pragma unroll
        for (int k = 0; k < ND; ++k) mbs += mbk[k] * (rho_i[k] + rj[k]);
      }
    }
    // with rdm = 1, 1 - r * 1 is 1 - r bit for bit, so only the reference
    // envelope (whose wc may come from FP64 near the cutoff) needs wc itself
    const float wm = (ENV == 0 && rdm1) ? wc : fmaxf(1.0f - rr * inv_rdm, 0.0f);
    const float rv = dx * (vi.x - vj.x) + dy * (vi.y - vj.y) + dz * (vi.z - vj.z);
    const float sgx = sg * xi, gmr = gm * rv * rinv;
    const float fs = (cons_c * wc + mbs * wm + sgx * wR - gmr * wD) * rinv;
    if (!exact) {
      // |dwR/dr| = 1 / (2 sqrt(1 - r)) or 1, |dwD/dr| = 1 or 2 (1 - r),
      // |dwm/dr| = 1 / rdm where wm > 0 (in the band only with rdm near 1)
      const float kR = ENV == 0 ? 0.5f * __builtin_amdgcn_rsqf(wc) : 1.0f;
      const float kD = ENV == 0 ? 1.0f : 2.0f * wc;
      const float de = fabsf(sgx) * kR + fabsf(gmr) * kD + fabsf(cons_c) +
                       fabsf(mbs) * (wm > 0.0f ? inv_rdm : 0.0f);
      band_e += band ? de : 0.0f;
      fsum += fabsf(fs) * rr;
    }
    fx += fs * dx;
    fy += fs * dy;
    fz += fs * dz;
    if (VIRIAL) {
      px += fs * dx * dx;
      py += fs * dy * dy;
      pz += fs * dz * dz;
    }
  };
  // the list is read once here (the rare exact walk below reads it again):
  // in large systems streaming loads keep its 4 B per pair from displacing
  // neighbour records in L2 (1M beads: force 157.5 -> 151.0 us; the 198k
  // slab and the 24k fluid, whose data fit in the caches, are better
  // without).  Each chunk's four records are gathered before its first pair
  // (four gathers in flight per lane; the unused entries of a list's last
  // chunk are 0, slot 0, so every load is in bounds) and the next chunk of
  // the list is loaded meanwhile: 95 VGPRs and 5 waves per SIMD instead of
  // 70 and 7, the pass waits less on its gathers.
  const bool stream = d.nl >= kStreamMin;
  auto ldl = [&](uint32_t k) -> uint4 {
    const uint4* q = lst + (size_t)(k >> 2) * 64u;
    return stream ? ld_stream_u4(q) : *q;
  };
  uint4 nxt = cnt ? ldl(0) : make_uint4(0u, 0u, 0u, 0u);
  for (uint32_t k = 0; k < cnt; k += 4) {
    const uint4 ch = nxt;
    if (k + 4 < cnt) nxt = ldl(k + 4);
    const SBead b0 = d.sb[ch.x & kIndexMask], b1 = d.sb[ch.y & kIndexMask],
                b2 = d.sb[ch.z & kIndexMask], b3 = d.sb[ch.w & kIndexMask];
    pair(ch.x, b0.pos, b0.vel, false);
    if (k + 1 < cnt) pair(ch.y, b1.pos, b1.vel, false);
    if (k + 2 < cnt) pair(ch.z, b2.pos, b2.vel, false);
    if (k + 3 < cnt) pair(ch.w, b3.pos, b3.vel, false);
  }
  if (kDrEst * band_e > 0.5e-5f * fmaxf(fsum, 1.0e-3f)) {
    // rare: the same walk with the band's 1 - r from FP64
#pragma unroll
    for (int a = 0; a < 3; ++a) pid[a] = d.pos[a * d.cap + s];
    fx = fy = fz = px = py = pz = 0.f;
    for (uint32_t k = 0; k < cnt; k += 4) {
      const uint4 ch = lst[(size_t)(k >> 2) * 64u];
      const uint32_t e4[4] = {ch.x, ch.y, ch.z, ch.w};
      for (uint32_t q = 0; q < 4 && k + q < cnt; ++q) {
        const SBead& b = d.sb[e4[q] & kIndexMask];
        pair(e4[q], b.pos, b.vel, true);
      }
    }
  }
  if (VIRIAL) {
    d.press[s] = px;
    d.press[d.cap + s] = py;
    d.press[2 * d.cap + s] = pz;
  }
  if (!isfinite(fx + fy + fz)) atomicOr(d.flags, FLAG_NONFINITE);
  if (vir_bead) {
    // ForceBonded::CalcListedForce and kick for this bead, from the FP32
    // acceleration just summed (the same value the stand-alone pass reads)
    float af[3] = {fx, fy, fz};
    double vk[3];
    bonded_kick_bead(d, R, s, af, VIRIAL, true, vir_bead, vk);
    if (keys) {
      // the next step's drifted position (drift_and_predict,
      // driftkick.hpp:5-30) and its cell, as k_drift computes them
      double pn[3];
      bool fin = true;
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        pn[a] = drift_axis(R, a, d.pos[a * d.cap + s], vk[a], (double)af[a], 0.0).p;
        fin = fin && isfinite(pn[a]) && isfinite(vk[a]);
      }
      if (!fin) atomicOr(d.flags, FLAG_NONFINITE);
      key = cell_of(R, pn);
      d.key[s] = key;
    }
    if (conv && d.unit[s] == 0u && d.prop[s] == (uint32_t)R.conv_from && R.conv_p > 0.0) {
      // the step's head conversion for this bead (same draw as
      // k_head_conv); only the master type changes, and the pair passes
      // read the FP32 copy, which the next rebuild rewrites
      Saru q = saru_seed(d.amp_id[s], *d.d_time, R.chem_seed);
      if (saru_uni_f64(q) <= R.conv_p) d.prop[s] = (uint32_t)R.conv_to;
    }
  } else {
    d.acc[s] = fx;
    d.acc[d.cap + s] = fy;
    d.acc[2 * d.cap + s] = fz;
  }
  }  // active
  if (keys) (void)agg_inc(d.cell_cnt, key, active);
}

template <int NS, int ND, int ENV>
static void force_launch3(const DevState& d, bool virial, int block, hipStream_t st, double* vb,
                          int cv, int ky) {
  const dim3 g(div_up(d.nl, block)), b(block);
  const bool multi = d.n_rep > 1;
  if (multi) {
    if (virial) hipLaunchKernelGGL((k_force<NS, ND, ENV, true, true>), g, b, 0, st, d, vb, cv, ky);
    else hipLaunchKernelGGL((k_force<NS, ND, ENV, true, false>), g, b, 0, st, d, vb, cv, ky);
  } else {
    if (virial) hipLaunchKernelGGL((k_force<NS, ND, ENV, false, true>), g, b, 0, st, d, vb, cv, ky);
    else hipLaunchKernelGGL((k_force<NS, ND, ENV, false, false>), g, b, 0, st, d, vb, cv, ky);
  }
}

template <int NS, int ND>
static void force_launch2(const DevState& d, int env, bool virial, int block, hipStream_t st,
                          double* vb, int cv, int ky) {
  if (env == 0) force_launch3<NS, ND, 0>(d, virial, block, st, vb, cv, ky);
  else force_launch3<NS, ND, 1>(d, virial, block, st, vb, cv, ky);
}

void launch_force(const DevState& d, int ns, int nd, int env, bool virial, int block,
                  hipStream_t st, bool fuse, bool conv, bool keys) {
  if (!d.nl) return;
  double* vb = fuse ? bead_vir_buffer(d) : nullptr;
  const int cv = (fuse && conv) ? 1 : 0;
  const int ky = (fuse && keys) ? 1 : 0;
  if (nd == 3) force_launch2<3, 3>(d, env, virial, block, st, vb, cv, ky);
  else if (ns == 1) force_launch2<1, 1>(d, env, virial, block, st, vb, cv, ky);
  else force_launch2<3, 1>(d, env, virial, block, st, vb, cv, ky);
}

// ---------------------------------------------------------------- pair listing
// Every ordered pair with r^2 < cutoff^2 as (upload slot i, upload slot j),
// for the neighbour-set checks; same decision code as the density pass.
__global__ void __launch_bounds__(256) k_pairs_list(DevState d, float cut2f, double cut2,
                                                    uint32_t* __restrict__ oi,
                                                    uint32_t* __restrict__ oj,
                                                    unsigned long long* __restrict__ counter,
                                                    uint64_t cap) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= dev_n(d)) return;
  const int r = d.n_rep > 1 ? find_rep(d.rep_off, d.n_rep, s) : 0;
  const Replica& R = d.reps[r];
  const float4 pi = d.spos[s];
  const uint32_t c = d.scell[s] - R.cell_off;
  const int nx = R.nc[0], ny = R.nc[1], nz = R.nc[2];
  const int cx = (int)(c % (uint32_t)nx), cy = (int)((c / (uint32_t)nx) % (uint32_t)ny),
            cz = (int)(c / (uint32_t)(nx * ny));
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
          const float r2 = dx * dx + dy * dy + dz * dz;
          if (r2 >= cut2f + kGuard) continue;
          if (decide(r2, cut2f, cut2, d, R, s, t)) {
            const unsigned long long k = atomicAdd(counter, 1ull);
            if (k < cap) {
              oi[k] = d.orig[s];
              oj[k] = d.orig[t];
            }
          }
        }
      }
}

void launch_pairs_list(const DevState& d, double cutoff, uint32_t* out_i, uint32_t* out_j,
                       unsigned long long* counter, uint64_t cap, hipStream_t st) {
  if (!d.nl) return;
  hipLaunchKernelGGL(k_pairs_list, dim3(div_up(d.nl, 256)), dim3(256), 0, st, d,
                     (float)(cutoff * cutoff), cutoff * cutoff, out_i, out_j, counter, cap);
}

}  // namespace mdpd
