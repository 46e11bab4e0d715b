// Device helpers of the step shared by the step kernels (k_step.hip) and
// the force pass's epilogue (k_pair.hip): the cell of a position, the
// wave-aggregated histogram increment, and one axis of the drift.  Every
// kernel that drifts or bins a bead uses these, so they agree bit for bit.
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>

#include "common.hpp"

namespace mdpd {

__device__ __forceinline__ uint32_t cell_of(const Replica& R, const double p[3]) {
  int c[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    int x = (int)floor(p[a] * R.inv_h[a]);
    x = x < 0 ? 0 : x;
    x = x >= R.nc[a] ? R.nc[a] - 1 : x;
    c[a] = x;
  }
  return R.cell_off + ((uint32_t)c[2] * (uint32_t)R.nc[1] + (uint32_t)c[1]) * (uint32_t)R.nc[0] +
         (uint32_t)c[0];
}

// Wave-aggregated "slot = base[key]++": lanes of a contiguous run holding
// the same key (bead storage is in cell order, so runs are the rule) share a
// single atomic issued by the run's first lane; members get consecutive
// slots in lane order.  Inactive lanes break runs and take no slot.
__device__ __forceinline__ uint32_t agg_inc(uint32_t* base, uint32_t key, bool active) {
  const int lane = threadIdx.x & 63;
  const uint32_t kprev = __shfl_up(key, 1, 64);
  const int aprev = __shfl_up(active ? 1 : 0, 1, 64);
  const bool leader = active && (lane == 0 || !aprev || kprev != key);
  const unsigned long long stop = __ballot(leader || !active);
  const unsigned long long upto = (lane == 63) ? ~0ull : ((2ull << lane) - 1ull);
  const int start = 63 - __clzll(stop & upto);  // my run's leader (or me)
  const unsigned long long above = stop & ~upto;
  const int end = above ? (__ffsll((long long)above) - 1) : 64;
  uint32_t old = 0;
  if (leader) old = atomicAdd(&base[key], (uint32_t)(end - lane));
  const uint32_t b = __shfl(old, start, 64);
  return b + (uint32_t)(lane - start);
}

// One axis of drift_and_predict (driftkick.hpp:5-30) in FP64 without
// contraction: the wrapped position, delta_sumr and the predicted velocity.
// Every kernel that drifts uses this, so they agree bit for bit.
struct DriftAxis {
  double p, dl, vb;
};
// the same with the replica's dt, 1 / box and box edge of the axis as values
__device__ __forceinline__ DriftAxis drift_axis_v(double dt, double inv_box, double box, double x,
                                                  double v, double ac, double dl) {
#pragma clang fp contract(off)
  const double dr = dt * (v + 0.5 * dt * ac);
  const double tr = x + dr;
  DriftAxis o;
  o.dl = dl + dr;
  o.p = tr - floor(tr * inv_box) * box;
  o.vb = v + 0.5 * dt * ac;
  return o;
}
__device__ __forceinline__ DriftAxis drift_axis(const Replica& R, int a, double x, double v,
                                                double ac, double dl) {
  return drift_axis_v(R.dt, R.inv_box[a], R.box[a], x, v, ac, dl);
}

}  // namespace mdpd
