// Shared definitions of the engine's kernels and host code.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mdpd {

constexpr int kMaxRep = 1024;        // replicas per context
constexpr int kWave = 64;            // CDNA wavefront
constexpr uint32_t kCodeShift = 26;  // pair-list entry: j | stencil code << 26 (2 bits per axis)
constexpr uint32_t kIndexMask = (1u << kCodeShift) - 1u;

// x of the FP32 copies (spos.x, sb.pos.x, as raw bits): a fixed-point
// coordinate in units of h_x / 2^24, the cell index in the bits from 24 up
// (modulo 2^32), so a pair's x displacement is one exact integer difference
// whichever of the three x-neighbour cells the partner is in.  Across the
// periodic x boundary the difference is corrected by the box in those
// units, nx << 24 (modulo 2^32).  Pair-list codes carry that correction in
// their x field: 0 = + (nx << 24), 1 = none, 2 = - (nx << 24); their y and
// z fields are the neighbour cell's offset -1, 0, +1 (as 0, 1, 2).
constexpr double kFixUnits = 16777216.0;                  // 2^24 units per cell
constexpr float kFixStep = 5.9604644775390625e-8f;        // 2^-24
__device__ __forceinline__ uint32_t fix_x(double l, uint32_t cx, double inv_h) {
  return (cx << 24) + (uint32_t)(int32_t)rint(l * inv_h * kFixUnits);
}
// x displacement of the pair (xi + off) - xj in length units (hs = h_x 2^-24)
__device__ __forceinline__ float fix_dx(uint32_t xi_off, uint32_t xj, float hs) {
  return (float)(int32_t)(xi_off - xj) * hs;
}
// the x correction of a pair-list code's x field (0, 1, 2)
__device__ __forceinline__ uint32_t fix_wrap(uint32_t field, uint32_t wx) {
  return field == 0u ? wx : (field == 2u ? 0u - wx : 0u);
}

// error flags (device word, sticky until read)
enum : uint32_t {
  FLAG_NLIST_OVERFLOW = 1u,
  FLAG_NONFINITE = 2u,
  FLAG_OUT_OF_BOX = 4u,
  FLAG_CHEM_CAPACITY = 8u,  // a chemistry event would have exceeded the capacity
};

// Per-replica constants, device resident.  FP32 copies of the pair
// coefficients feed the pair passes; FP64 values feed integration and the
// guard-band re-decisions.
struct Replica {
  double box[3];
  double inv_box[3];
  double h[3];          // cell edge = box / nc
  double inv_h[3];
  double dt;
  double rd2;           // density cutoff^2 (FP64 decisions)
  double conv_p;
  double cf_s, ibond;   // bond spring, 1/bond_leng
  int32_t nc[3];
  uint32_t cell_off;    // first global cell of this replica
  uint32_t bead_off;    // first bead slot
  uint32_t n_beads;
  uint32_t n_cells;
  uint32_t pair_seed;
  uint32_t chem_seed;
  int32_t conv_from, conv_to;
  int32_t n_species, n_dens, envelope;
  float hf[3];
  float rd, inv_rd, crho, inv_rdm, arc;
  float rd2f;
  int32_t rdm_is_one;
  int32_t has_c1;      // any c1 != 0 (the c1 * max(r - arc, 0) term is live)
  float a0[3][3], c1[3][3], sigma[3][3], gamma[3][3];
  float mb[3][3][3];
};

// Guard band on r^2 (FP32 local-coordinate differences carry |error| of a
// few 1e-7); inside it the pair decision is redone in FP64 from the master
// positions with the oracle's minimum-image expression.
constexpr float kGuard = 2.0e-5f;

__host__ __device__ inline uint32_t div_up(uint32_t a, uint32_t b) { return (a + b - 1) / b; }

// Systems of at least this many beads take streaming (non-temporal) loads
// and stores for the step's last-use data, so that the pair passes keep
// their neighbour data in L2; smaller systems fit in the caches and would
// only lose the next step's reuse (measured: faster at 1M and 2.5M beads,
// slower at 24k and 198k).
constexpr uint32_t kStreamMin = 1u << 19;

// a 16-byte streaming (non-temporal) load: one global_load_dwordx4 ... nt
__device__ __forceinline__ uint4 ld_stream_u4(const uint4* p) {
  typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
  const u32x4 v = __builtin_nontemporal_load(reinterpret_cast<const u32x4*>(p));
  return make_uint4(v.x, v.y, v.z, v.w);
}
template <bool ST, class T>
__device__ __forceinline__ T ld_hint(const T* p) {
  if constexpr (ST) return __builtin_nontemporal_load(p);
  else return *p;
}
template <bool ST, class T>
__device__ __forceinline__ void st_hint(T* p, T v) {
  if constexpr (ST) __builtin_nontemporal_store(v, p);
  else *p = v;
}

// replica of bead slot i (bead_off ascending, n_rep entries + sentinel)
__device__ __forceinline__ int find_rep(const uint32_t* __restrict__ off, int n_rep, uint32_t i) {
  int lo = 0, hi = n_rep - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (off[mid] <= i) lo = mid; else hi = mid - 1;
  }
  return lo;
}

// XCD-aware block order: the dispatcher hands consecutive blocks to the 8
// XCDs round robin; this gives each XCD a contiguous run of tiles so that
// neighbouring tiles share that XCD's L2.
// nb = the tiles that hold beads; the grid may be larger (grow mode launches
// over the capacity): its blocks past those the tiles need get tile nb, which
// holds no bead.  With a grid of exactly nb blocks every block gets a tile.
__device__ __forceinline__ uint32_t xcd_block(uint32_t b, uint32_t nb) {
  constexpr uint32_t X = 8;
  const uint32_t per = nb / X, rem = nb % X;
  const uint32_t x = b % X, k = b / X;
  // XCDs 0..rem-1 hold per+1 tiles, the others per
  const uint32_t base = x < rem ? x * (per + 1) : rem * (per + 1) + (x - rem) * per;
  const uint32_t cnt = x < rem ? per + 1 : per;
  return (k < cnt) ? base + k : nb;
}

}  // namespace mdpd
