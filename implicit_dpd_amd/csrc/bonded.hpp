// Bonded chain forces (ForceBonded, bdf_calculator.hpp:11-243) and the kick
// (driftkick.hpp:32-40) as device functions, shared by the stand-alone
// bonded+kick kernel and the force pass that ends with them.
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>

#include "common.hpp"
#include "state.hpp"

namespace mdpd {

// ForceBonded::MinImage (bdf_calculator.hpp:55-59)
__device__ __forceinline__ void bd_min_image(const Replica& R, double d[3]) {
#pragma clang fp contract(off)
  for (int a = 0; a < 3; ++a) d[a] -= R.box[a] * round(d[a] * R.inv_box[a]);
}

// Bonded terms of chain a (ForceBonded::CalcBondBendGlobalCell<4>,
// bdf_calculator.hpp:153-189, with MinImage :55-59, the bond :95-115 and
// the bend :117-150), FP64 in the reference's order: F[u] receives its terms
// in exactly the sequence of the reference loop.  The four bead positions
// come from the chain-major copy chainpos[4a + u].
__device__ __forceinline__ void chain_terms(const DevState& d, const Replica& R, uint32_t a,
                                            double F[4][3], double vir[3]) {
#pragma clang fp contract(off)
  double p[4][3];
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    const double4 q = d.chainpos[4 * (size_t)a + u];
    p[u][0] = q.x;
    p[u][1] = q.y;
    p[u][2] = q.z;
  }
  const double cf_b = d.cf_bend[a];
#pragma unroll
  for (int u = 0; u < 4; ++u)
#pragma unroll
    for (int c = 0; c < 3; ++c) F[u][c] = 0.0;
  double dr[3][3], inv_dr[3];
  vir[0] = vir[1] = vir[2] = 0.0;
#pragma unroll
  for (int u = 1; u < 4; ++u) {
#pragma unroll
    for (int c = 0; c < 3; ++c) dr[u - 1][c] = p[u][c] - p[u - 1][c];
    bd_min_image(R, dr[u - 1]);
    const double dist2 = dr[u - 1][0] * dr[u - 1][0] + dr[u - 1][1] * dr[u - 1][1] +
                         dr[u - 1][2] * dr[u - 1][2];
    inv_dr[u - 1] = 1.0 / sqrt(dist2);
    const double cf_bond = R.cf_s * (inv_dr[u - 1] - R.ibond);
    double Fb[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) Fb[c] = cf_bond * dr[u - 1][c];
#pragma unroll
    for (int c = 0; c < 3; ++c) vir[c] += 2.0 * dr[u - 1][c] * Fb[c];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      F[u - 1][c] -= Fb[c];
      F[u][c] += Fb[c];
    }
    if (u >= 2) {
      const double* d0 = dr[u - 2];
      const double* d1 = dr[u - 1];
      const double inv_dr_prod = inv_dr[u - 2] * inv_dr[u - 1];
      const double inv_dist0 = inv_dr[u - 2] * inv_dr[u - 2];
      const double inv_dist1 = inv_dr[u - 1] * inv_dr[u - 1];
      const double in_prod = d0[0] * d1[0] + d0[1] * d1[1] + d0[2] * d1[2];
      const double cf_bd = cf_b * inv_dr_prod;
      const double crs0 = in_prod * inv_dist0, crs1 = in_prod * inv_dist1;
      double T0[3], T1[3];
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        T0[c] = cf_bd * (d1[c] - crs0 * d0[c]);
        T1[c] = cf_bd * (d0[c] - crs1 * d1[c]);
      }
#pragma unroll
      for (int c = 0; c < 3; ++c) vir[c] += 2.0 * (d0[c] * T0[c] + d1[c] * T1[c]);
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        F[u - 2][c] -= T0[c];
        F[u - 1][c] += T0[c] - T1[c];
        F[u][c] += T1[c];
      }
    }
  }
}

// Bead s: adds the bonded force of its own chain to its FP32 acceleration
// af (the value the force pass stored), stores it, and kicks.  The chain
// virial is counted once, by the chain's first bead, into vir_bead[3s..].
// vout (may be null): the kicked velocity.
__device__ __forceinline__ void bonded_kick_bead(const DevState& d, const Replica& R, uint32_t s,
                                                 float af[3], bool virial, bool kick,
                                                 double* __restrict__ vir_bead,
                                                 double* vout = nullptr) {
#pragma clang fp contract(off)
  const uint32_t cap = d.cap;
  const uint32_t ch = d.chain[s];
  double vir[3] = {0.0, 0.0, 0.0};
  if (ch != 0xffffffffu) {
    const uint32_t u = ch & 3u;
    double F[4][3];
    chain_terms(d, R, ch >> 2, F, vir);
#pragma unroll
    for (int c = 0; c < 3; ++c) af[c] = (float)((double)af[c] + F[u][c]);
    if (u != 0) vir[0] = vir[1] = vir[2] = 0.0;
  }
#pragma unroll
  for (int c = 0; c < 3; ++c) d.acc[c * cap + s] = af[c];
  if (virial)
#pragma unroll
    for (int c = 0; c < 3; ++c) vir_bead[3 * (size_t)s + c] = vir[c];
  if (kick) {
    const double dt = R.dt;
#pragma unroll
    for (int c = 0; c < 3; ++c) {  // vel_buf's last use in the step: streaming load
      const double vb = d.n >= kStreamMin ? ld_hint<true>(&d.vbuf[c * cap + s]) : d.vbuf[c * cap + s];
      const double v = vb + dt * 0.5 * (double)af[c];
      d.vel[c * cap + s] = v;
      if (vout) vout[c] = v;
    }
  }
}

}  // namespace mdpd
