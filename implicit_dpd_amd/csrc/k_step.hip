// Integrator, cell-list rebuild, bonded forces, reactions and reductions.
//
// FP64 kernels that must agree bit for bit with the oracle's restatement of
// the reference (drift, kick, bonded, the FP64 pair re-decision) are
// compiled without FMA contraction (the pragma at the top of each body).
#include <hip/hip_runtime.h>
#include <math.h>

#include "common.hpp"
#include "saru.hpp"
#include "state.hpp"
#include "bonded.hpp"
#include "step_dev.hpp"

namespace mdpd {

// stream tags of the keyed chemistry draws
constexpr uint32_t TAG_CONE = 0x9E3779B9u;
constexpr uint32_t TAG_TANG = 0x5bd1e995u;
constexpr uint32_t TAG_VEL = 0x68e31da4u;

static inline dim3 grid1(uint32_t n, uint32_t b) { return dim3(div_up(n ? n : 1, b)); }

// ---------------------------------------------------------------- fill
__global__ void k_fill_u32(uint32_t* __restrict__ p, uint32_t n, uint32_t v) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    p[i] = v;
}
void launch_fill_u32(uint32_t* p, uint32_t n, uint32_t v, hipStream_t st) {
  if (!n) return;
  const uint32_t b = 256;
  uint32_t g = div_up(n, b);
  if (g > 4096) g = 4096;
  hipLaunchKernelGGL(k_fill_u32, dim3(g), dim3(b), 0, st, p, n, v);
}

__global__ void k_fill_d(double* __restrict__ p, size_t n, double v) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    p[i] = v;
}

// ---------------------------------------------------------------- cells

// drift_and_predict, driftkick.hpp:5-30.  vel is left as the last kicked
// velocity; the predicted velocity (vel_buf == vel of the reference after
// the drift) goes to vbuf.  With `keys` the cell key and the cell
// histogram of the rebuild are produced in the same pass.

// store = 0 (the step): only the cell key and histogram of the drifted
// position; the permutation that follows drifts each bead again as it moves
// it and writes the new state once.
__global__ void __launch_bounds__(256) k_drift(DevState d, int keys, int predict_only, int tick,
                                               int store) {
  if (halted(d)) return;
  // tick: the step's clock advances here (no kernel of the step before
  // this one reads it)
  if (tick && blockIdx.x == 0 && threadIdx.x == 0) *d.d_time += 1u;
  const uint32_t i0 = blockIdx.x * blockDim.x + threadIdx.x;
  const bool active = i0 < dev_n(d);
  const uint32_t i = active ? i0 : dev_n(d) - 1;
  uint32_t k = 0;
  if (active) {
    const int r = d.n_rep > 1 ? find_rep(d.rep_off, d.n_rep, i) : 0;
    const Replica& R = d.reps[r];
    const uint32_t cap = d.cap;
    double p[3];
    bool fin = true;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      const double v = d.vel[a * cap + i];
      const double ac = (double)d.acc[a * cap + i];
      const double x = d.pos[a * cap + i];
      if (!predict_only) {
        const DriftAxis o = drift_axis(R, a, x, v, ac, store ? d.delta[a * cap + i] : 0.0);
        p[a] = o.p;
        if (store) {
          d.delta[a * cap + i] = o.dl;
          d.pos[a * cap + i] = o.p;
          d.vbuf[a * cap + i] = o.vb;
        }
      } else {
        p[a] = x;
        d.vbuf[a * cap + i] = drift_axis(R, a, x, v, ac, 0.0).vb;
      }
      fin = fin && isfinite(p[a]) && isfinite(v);
    }
    if (!fin) atomicOr(d.flags, FLAG_NONFINITE);
    if (keys) {
      k = cell_of(R, p);
      d.key[i] = k;
    }
  }
  if (keys) (void)agg_inc(d.cell_cnt, k, active);
}
void launch_drift(const DevState& d, bool keys, bool predict_only, hipStream_t st, bool tick) {
  if (!d.nl) return;
  hipLaunchKernelGGL(k_drift, grid1(d.nl, 256), dim3(256), 0, st, d, keys ? 1 : 0,
                     predict_only ? 1 : 0, tick ? 1 : 0, 1);
}
void launch_step_keys(const DevState& d, hipStream_t st) {
  if (!d.nl) return;
  hipLaunchKernelGGL(k_drift, grid1(d.nl, 256), dim3(256), 0, st, d, 1, 0, 0, 0);
}

__global__ void __launch_bounds__(256) k_keys(DevState d, int gate) {
  if (gate && *d.rebuild == 0u) return;
  const uint32_t i0 = blockIdx.x * blockDim.x + threadIdx.x;
  const bool active = i0 < dev_n(d);
  uint32_t k = 0;
  if (active) {
    const int r = d.n_rep > 1 ? find_rep(d.rep_off, d.n_rep, i0) : 0;
    const Replica& R = d.reps[r];
    double p[3];
    for (int a = 0; a < 3; ++a) p[a] = d.pos[a * d.cap + i0];
    k = cell_of(R, p);
    d.key[i0] = k;
  }
  (void)agg_inc(d.cell_cnt, k, active);
}
void launch_keys(const DevState& d, hipStream_t st) {
  if (!d.nl) return;
  hipLaunchKernelGGL(k_keys, grid1(d.nl, 256), dim3(256), 0, st, d, 0);
}

// exclusive scan of the cell histogram (1024 cells per block, three passes)
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t u = __shfl_up(v, o, 64);
    if (lane >= o) v += u;
  }
  return v;
}

// block of 256 threads: exclusive scan of per-thread totals, returns block total
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t& excl) {
  __shared__ uint32_t wsum[4];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const uint32_t inc = wave_incl_scan(v);
  if (lane == 63) wsum[w] = inc;
  __syncthreads();
  uint32_t off = 0, tot = 0;
  for (int k = 0; k < 4; ++k) {
    if (k < w) off += wsum[k];
    tot += wsum[k];
  }
  excl = off + inc - v;
  __syncthreads();
  return tot;
}

// The histogram is cleared as it is read, ready for the next rebuild.
__global__ void __launch_bounds__(256) k_scan1(uint32_t* __restrict__ in,
                                               uint32_t* __restrict__ out,
                                               uint32_t* __restrict__ bsum, uint32_t m,
                                               const uint32_t* __restrict__ gate) {
  if (gate && *gate == 0u) return;
  const uint32_t base = blockIdx.x * 1024u + threadIdx.x * 4u;
  uint32_t v[4], t = 0;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    v[k] = (base + k < m) ? in[base + k] : 0u;
    if (base + k < m) in[base + k] = 0u;
    t += v[k];
  }
  uint32_t excl;
  const uint32_t tot = block_excl_scan(t, excl);
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    if (base + k < m) out[base + k] = excl;
    excl += v[k];
  }
  if (threadIdx.x == 0) bsum[blockIdx.x] = tot;
}

__global__ void __launch_bounds__(256) k_scan2(uint32_t* __restrict__ bsum, uint32_t nb,
                                               const uint32_t* __restrict__ gate) {
  if (gate && *gate == 0u) return;
  uint32_t carry = 0;
  for (uint32_t base = 0; base < nb; base += 1024u) {
    const uint32_t b0 = base + threadIdx.x * 4u;
    uint32_t v[4], t = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      v[k] = (b0 + k < nb) ? bsum[b0 + k] : 0u;
      t += v[k];
    }
    uint32_t excl;
    const uint32_t tot = block_excl_scan(t, excl);
    excl += carry;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      if (b0 + k < nb) bsum[b0 + k] = excl;
      excl += v[k];
    }
    carry += tot;
  }
}

// carry: bsum holds block totals and every block sums those before it
// (few blocks), else bsum already holds the exclusive block offsets
__global__ void __launch_bounds__(256) k_scan3(uint32_t* __restrict__ out,
                                               uint32_t* __restrict__ cursor,
                                               const uint32_t* __restrict__ bsum, uint32_t m,
                                               uint32_t total,
                                               const uint32_t* __restrict__ total_dev, int carry,
                                               const uint32_t* __restrict__ gate) {
  if (gate && *gate == 0u) return;
  const uint32_t base = blockIdx.x * 1024u;
  uint32_t add;
  if (carry) {
    __shared__ uint32_t part[4];
    uint32_t v = 0;
    for (uint32_t b = threadIdx.x; b < blockIdx.x; b += 256u) v += bsum[b];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = v;
    __syncthreads();
    add = part[0] + part[1] + part[2] + part[3];
  } else {
    add = bsum[blockIdx.x];
  }
  for (uint32_t k = threadIdx.x; k < 1024u; k += 256u) {
    const uint32_t idx = base + k;
    if (idx < m) {
      const uint32_t v = out[idx] + add;
      out[idx] = v;
      cursor[idx] = v;
    }
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) out[m] = total_dev ? *total_dev : total;
}

static void scan_cells(const DevState& d, const uint32_t* gate, hipStream_t st) {
  const uint32_t m = d.ncell;
  const uint32_t nb = div_up(m, 1024u);
  hipLaunchKernelGGL(k_scan1, dim3(nb), dim3(256), 0, st, d.cell_cnt, d.cell_start, d.bsum, m,
                     gate);
  const int carry = nb <= 4096u ? 1 : 0;  // up to 4M cells: no separate block-sum pass
  if (!carry) hipLaunchKernelGGL(k_scan2, dim3(1), dim3(256), 0, st, d.bsum, nb, gate);
  // in grow mode the bead count is the device's (chemistry may have added beads)
  const uint32_t* total_dev = d.run_ok ? d.rep_off + d.n_rep : nullptr;
  hipLaunchKernelGGL(k_scan3, dim3(nb), dim3(256), 0, st, d.cell_start, d.cursor, d.bsum, m, d.n,
                     total_dev, carry, gate);
}
// (grow mode: a run stopped by the chemistry skips its remaining steps)
void launch_scan_cells(const DevState& d, hipStream_t st) { scan_cells(d, d.run_ok, st); }

__global__ void __launch_bounds__(256) k_scatter(DevState d, int gate) {
  if (gate && *d.rebuild == 0u) return;
  if (halted(d)) return;
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  const bool active = i < dev_n(d);
  const uint32_t k = active ? d.key[i] : 0u;
  const uint32_t slot = agg_inc(d.cursor, k, active);
  if (active) d.perm[slot] = i;
}
void launch_scatter(const DevState& d, hipStream_t st) {
  if (!d.nl) return;
  hipLaunchKernelGGL(k_scatter, grid1(d.nl, 256), dim3(256), 0, st, d, 0);
}

// ascending bead slots inside each cell, so the sorted order (and every FP32
// sum over it) does not depend on the order of the scatter atomics
__global__ void __launch_bounds__(256) k_cell_sort(DevState d, int gate) {
  if (gate && *d.rebuild == 0u) return;
  const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= d.ncell) return;
  const uint32_t b = d.cell_start[c], e = d.cell_start[c + 1];
  for (uint32_t k = b + 1; k < e; ++k) {
    const uint32_t v = d.perm[k];
    uint32_t q = k;
    while (q > b) {
      const uint32_t u = d.perm[q - 1];
      if (u <= v) break;
      d.perm[q] = u;
      --q;
    }
    d.perm[q] = v;
  }
}
void launch_cell_sort(const DevState& d, hipStream_t st) {
  hipLaunchKernelGGL(k_cell_sort, grid1(d.ncell, 256), dim3(256), 0, st, d, 0);
}

// Moves every bead to its cell-order slot s (from old slot perm[s]) in the
// other master buffer set, and writes the FP32 copies the pair passes read:
// position inside the bead's cell (FP64 difference, then rounded) with
// id << 2 | type in .w, and the velocity the force uses.
// This form serves the explicit rebuild (full: also vel, acc and press, and
// the inverse map for the topology) and skin mode's rebuild steps; the
// every-step path is k_step_permute, which also drifts the bead.
__global__ void __launch_bounds__(256) k_permute(DevState d, int predicted, int full, int skin) {
  if (skin && *d.rebuild == 0u) return;
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= dev_n(d)) return;
  const uint32_t i = d.perm[s];
  if (full) d.inv_perm[i] = s;  // for the explicit rebuild's topology remap
  const int r = d.n_rep > 1 ? find_rep(d.rep_off, d.n_rep, s) : 0;
  const Replica& R = d.reps[r];
  const uint32_t key = d.key[i];
  const uint32_t c = key - R.cell_off;
  const uint32_t nx = (uint32_t)R.nc[0], ny = (uint32_t)R.nc[1];
  const uint32_t cc[3] = {c % nx, (c / nx) % ny, c / (nx * ny)};
  const uint32_t cap = d.cap;
  float l[3], v[3];
  double xs[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const double x = d.pos[a * cap + i];
    const double dl = d.delta[a * cap + i];
    const double vb = d.vbuf[a * cap + i];
    xs[a] = x;
    d.alt.pos[a * cap + s] = x;
    d.alt.vbuf[a * cap + s] = vb;
    d.alt.delta[a * cap + s] = dl;
    // skin mode: the displacement test of the coming steps starts here
    if (skin) d.dref[a * cap + s] = dl;
    l[a] = (float)(x - (double)cc[a] * R.h[a]);
    if (a == 0) l[0] = __uint_as_float(fix_x(x - (double)cc[0] * R.h[0], cc[0], R.inv_h[0]));
    if (full) {
      d.alt.vel[a * cap + s] = d.vel[a * cap + i];
      d.alt.acc[a * cap + s] = d.acc[a * cap + i];
      d.alt.press[a * cap + s] = d.press[a * cap + i];
    }
    v[a] = (float)(predicted ? vb : d.vel[a * cap + i]);
  }
  const uint32_t id = d.id[i], pr = d.prop[i];
  d.alt.id[s] = id;
  d.alt.prop[s] = pr;
  d.alt.amp_id[s] = d.amp_id[i];
  d.alt.unit[s] = d.unit[i];
  d.alt.orig[s] = d.orig[i];
  const uint32_t ch = d.chain[i];
  d.alt.chain[s] = ch;
  if (ch != 0xffffffffu) d.chainpos[ch] = make_double4(xs[0], xs[1], xs[2], 0.0);
  const float4 ps = make_float4(l[0], l[1], l[2], __uint_as_float((id << 2) | (pr & 3u)));
  d.spos[s] = ps;
  d.sb[s].pos = ps;
  d.sb[s].vel = make_float4(v[0], v[1], v[2], 0.f);
  d.scell[s] = key;
}
void launch_permute(const DevState& d, bool predicted, bool full, hipStream_t st) {
  if (!d.nl) return;
  hipLaunchKernelGGL(k_permute, grid1(d.nl, 256), dim3(256), 0, st, d, predicted ? 1 : 0,
                     full ? 1 : 0, 0);
}

// The step's permutation (k_permute with the clock advance, the drift and
// the predicted velocity): every input of the bead is read before anything
// is written, so a wave has all of its old-set gathers in flight at once
// (stores into the new set could otherwise alias later loads for the
// compiler, which then serialises the axes), and with one system the
// replica constants are wave-uniform.  Same arithmetic as k_permute.
template <bool ST, bool MULTI>
__global__ void __launch_bounds__(256) k_step_permute(DevState d) {
  if (halted(d)) return;
  if (blockIdx.x == 0 && threadIdx.x == 0) *d.d_time += 1u;
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= dev_n(d)) return;
  const uint32_t i = d.perm[s];
  const int r = MULTI ? find_rep(d.rep_off, d.n_rep, s) : 0;
  const Replica& R = d.reps[r];
  const uint32_t cap = d.cap;
  // the old set is read here for the last time in this step
  const uint32_t key = d.key[i];
  double x[3], dl[3], vo[3];
  float ac[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    x[a] = ld_hint<ST>(&d.pos[a * cap + i]);
    dl[a] = ld_hint<ST>(&d.delta[a * cap + i]);
    vo[a] = ld_hint<ST>(&d.vel[a * cap + i]);
    ac[a] = ld_hint<ST>(&d.acc[a * cap + i]);
  }
  const uint32_t id = ld_hint<ST>(&d.id[i]), pr = d.prop[i], am = d.amp_id[i], un = d.unit[i];
  const uint32_t og = ld_hint<ST>(&d.orig[i]), ch = ld_hint<ST>(&d.chain[i]);
  // the scatter left each cell's beads in atomic order; the bead's slot is
  // its cell's start plus its rank by previous slot among the cell's beads
  // (the order k_cell_sort gives, without that pass)
  const uint32_t cb = d.cell_start[key], ce = d.cell_start[key + 1];
  uint32_t t = cb;
  for (uint32_t k = cb; k < ce; ++k) t += d.perm[k] < i ? 1u : 0u;
  // every value in registers before the first store: the compiler would
  // otherwise sink the loads to their uses, behind the earlier axes' stores
  asm volatile("" ::"v"(x[0]), "v"(x[1]), "v"(x[2]), "v"(dl[0]), "v"(dl[1]), "v"(dl[2]),
               "v"(vo[0]), "v"(vo[1]), "v"(vo[2]), "v"(ac[0]), "v"(ac[1]), "v"(ac[2]));
  asm volatile("" ::"v"(id), "v"(pr), "v"(am), "v"(un), "v"(og), "v"(ch), "v"(key));
  // the replica constants too, before any store
  const double dt = R.dt;
  double ib[3], bx[3], h[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    ib[a] = R.inv_box[a];
    bx[a] = R.box[a];
    h[a] = R.h[a];
  }
  const double ihx = R.inv_h[0];
  const uint32_t c = key - R.cell_off;
  const uint32_t nx = (uint32_t)R.nc[0], ny = (uint32_t)R.nc[1];
  const uint32_t cc[3] = {c % nx, (c / nx) % ny, c / (nx * ny)};
  float l[3], v[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const DriftAxis o = drift_axis_v(dt, ib[a], bx[a], x[a], vo[a], (double)ac[a], dl[a]);
    x[a] = o.p;
    // positions and delta_sumr are next read by the following step's
    // permutation: streaming stores keep them from displacing this step's
    // pair-pass data in L2 (vel_buf is read by this step's kick)
    st_hint<ST>(&d.alt.pos[a * cap + t], o.p);
    d.alt.vbuf[a * cap + t] = o.vb;
    st_hint<ST>(&d.alt.delta[a * cap + t], o.dl);
    l[a] = a == 0 ? __uint_as_float(fix_x(o.p - (double)cc[0] * h[0], cc[0], ihx))
                  : (float)(o.p - (double)cc[a] * h[a]);
    v[a] = (float)o.vb;
  }
  st_hint<ST>(&d.alt.id[t], id);
  d.alt.prop[t] = pr;
  d.alt.amp_id[t] = am;
  d.alt.unit[t] = un;
  st_hint<ST>(&d.alt.orig[t], og);
  st_hint<ST>(&d.alt.chain[t], ch);
  if (ch != 0xffffffffu) d.chainpos[ch] = make_double4(x[0], x[1], x[2], 0.0);
  const float4 ps = make_float4(l[0], l[1], l[2], __uint_as_float((id << 2) | (pr & 3u)));
  d.spos[t] = ps;
  d.sb[t].pos = ps;
  d.sb[t].vel = make_float4(v[0], v[1], v[2], 0.f);
  d.scell[t] = key;
}

void launch_step_permute(const DevState& d, hipStream_t st) {
  if (!d.nl) return;
  const bool big = d.nl >= kStreamMin, multi = d.n_rep > 1;
  const dim3 g = grid1(d.nl, 256), b(256);
  if (big) {
    if (multi) hipLaunchKernelGGL((k_step_permute<true, true>), g, b, 0, st, d);
    else hipLaunchKernelGGL((k_step_permute<true, false>), g, b, 0, st, d);
  } else {
    if (multi) hipLaunchKernelGGL((k_step_permute<false, true>), g, b, 0, st, d);
    else hipLaunchKernelGGL((k_step_permute<false, false>), g, b, 0, st, d);
  }
}

// chain bead slots follow the beads to their new slots
__global__ void __launch_bounds__(256) k_topo_remap(DevState d) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= 4u * d.n_amp) return;
  d.topo[k] = d.inv_perm[d.topo[k]];
}
void launch_topo_remap(const DevState& d, hipStream_t st) {
  if (!d.n_amp) return;
  hipLaunchKernelGGL(k_topo_remap, grid1(4 * d.n_amp, 256), dim3(256), 0, st, d);
}

// the topology from the chain map (every chain bead knows its chain slot)
__global__ void __launch_bounds__(256) k_topo_from_chain(DevState d) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= dev_n(d)) return;
  const uint32_t ch = d.chain[s];
  if (ch != 0xffffffffu) d.topo[ch] = s;
}
void launch_topo_from_chain(const DevState& d, hipStream_t st) {
  if (!d.nl || !d.n_amp) return;
  hipLaunchKernelGGL(k_topo_from_chain, grid1(d.nl, 256), dim3(256), 0, st, d);
}

// ---------------------------------------------------------------- kick
// driftkick.hpp:32-40: vel = vel_buf + dt * 0.5 * acc
__global__ void __launch_bounds__(256) k_kick(DevState d) {
#pragma clang fp contract(off)
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= dev_n(d)) return;
  const int r = d.n_rep > 1 ? find_rep(d.rep_off, d.n_rep, i) : 0;
  const double dt = d.reps[r].dt;
  const uint32_t cap = d.cap;
#pragma unroll
  for (int a = 0; a < 3; ++a)
    d.vel[a * cap + i] = d.vbuf[a * cap + i] + dt * 0.5 * (double)d.acc[a * cap + i];
}
void launch_kick(const DevState& d, hipStream_t st) {
  if (!d.nl) return;
  hipLaunchKernelGGL(k_kick, grid1(d.nl, 256), dim3(256), 0, st, d);
}

// ---------------------------------------------------------------- reductions
// Neumaier-compensated FP64 sums; fixed grid and fixed trees, so the result
// is the same on every run.
struct KSum {
  double s, c;
  __device__ __forceinline__ void add(double x) {
#pragma clang fp contract(off)
    const double t = s + x;
    if (fabs(s) >= fabs(x)) c += (s - t) + x;
    else c += (x - t) + s;
    s = t;
  }
  __device__ __forceinline__ void merge(const KSum& o) {
    add(o.s);
    c += o.c;
  }
};

constexpr int kRedBlocks = 64;   // partial blocks per replica
constexpr int kRedThreads = 256;

template <int Q>
__device__ __forceinline__ void block_reduce_write(KSum (&acc)[Q], double* __restrict__ out2) {
  // out2[q*2 + {0,1}] = (s, c) of the block
  __shared__ double sh_s[Q][kRedThreads / 64];
  __shared__ double sh_c[Q][kRedThreads / 64];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    KSum v = acc[q];
    for (int o = 32; o > 0; o >>= 1) {
      KSum u;
      u.s = __shfl_down(v.s, o, 64);
      u.c = __shfl_down(v.c, o, 64);
      if (lane < o) v.merge(u);
    }
    if (lane == 0) { sh_s[q][w] = v.s; sh_c[q][w] = v.c; }
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      KSum v{sh_s[q][0], sh_c[q][0]};
      for (int k = 1; k < kRedThreads / 64; ++k) v.merge(KSum{sh_s[q][k], sh_c[q][k]});
      out2[2 * q] = v.s;
      out2[2 * q + 1] = v.c;
    }
  }
}

// final: one block per replica sums its kRedBlocks partials in order
template <int Q>
__global__ void k_red_final(const double* __restrict__ part, double* __restrict__ out,
                            const uint32_t* __restrict__ gate) {
  if (gate && *gate == 0u) return;
  const int r = blockIdx.x;
  if (threadIdx.x < Q) {
    const int q = threadIdx.x;
    KSum v{0.0, 0.0};
    for (int b = 0; b < kRedBlocks; ++b) {
      const double* p = part + ((size_t)r * kRedBlocks + b) * 2 * Q;
      v.merge(KSum{p[2 * q], p[2 * q + 1]});
    }
    out[(size_t)r * Q + q] = v.s + v.c;
  }
}

// sum of velocities per replica (for remove_cmdrift_global)
__global__ void __launch_bounds__(kRedThreads) k_cm_sum(DevState d) {
  if (halted(d)) return;
  const int r = blockIdx.y;
  const uint32_t b0 = d.rep_off[r], n_r = d.rep_off[r + 1] - b0;
  KSum acc[3] = {{0, 0}, {0, 0}, {0, 0}};
  for (uint32_t k = blockIdx.x * kRedThreads + threadIdx.x; k < n_r; k += kRedBlocks * kRedThreads) {
    const uint32_t i = b0 + k;
#pragma unroll
    for (int a = 0; a < 3; ++a) acc[a].add(d.vel[a * d.cap + i]);
  }
  block_reduce_write<3>(acc, d.red + ((size_t)r * kRedBlocks + blockIdx.x) * 6);
}

__global__ void __launch_bounds__(256) k_cm_apply(DevState d) {
#pragma clang fp contract(off)
  if (halted(d)) return;
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= dev_n(d)) return;
  const int r = d.n_rep > 1 ? find_rep(d.rep_off, d.n_rep, i) : 0;
  const double n_r = (double)(d.rep_off[r + 1] - d.rep_off[r]);
#pragma unroll
  for (int a = 0; a < 3; ++a) d.vel[a * d.cap + i] -= d.red_out[(size_t)r * 3 + a] / n_r;
}

void launch_cm_drift(const DevState& d, hipStream_t st) {
  if (!d.nl) return;
  hipLaunchKernelGGL(k_cm_sum, dim3(kRedBlocks, d.n_rep), dim3(kRedThreads), 0, st, d);
  hipLaunchKernelGGL(k_red_final<3>, dim3(d.n_rep), dim3(64), 0, st, d.red, d.red_out,
                     (const uint32_t*)d.run_ok);
  hipLaunchKernelGGL(k_cm_apply, grid1(d.nl, 256), dim3(256), 0, st, d);
}

// Observer sums per replica (observer.hpp:82-178); amp_num[r] is the
// replica's amphiphile count.  Quantities:
//  0-2 sum v_a^2; 3-5 sum (0.5 press_a + v_a^2); 6-8 sum press_a;
//  9 sum |delta|^2; 10 sum |delta|^2 over local index >= amp_num (the
//  reference's solvent loop, observer.hpp:171); 11 sum over amphiphiles of
//  |mean delta of beads 4a..4a+3|^2 (observer.hpp:160-167); 12 number of
//  pair-list entries (sorted slots of the replica).
// delta_sumr in load order, dl[a][orig[i]] = delta[a][i]: the per-amphiphile
// term then reads each amphiphile's four beads as one contiguous run
__global__ void __launch_bounds__(256) k_delta_load_order(DevState d, double* __restrict__ dl) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= dev_n(d)) return;
  const uint32_t o = d.orig[i];
#pragma unroll
  for (int a = 0; a < 3; ++a) dl[a * (size_t)d.cap + o] = d.delta[a * d.cap + i];
}

__global__ void __launch_bounds__(kRedThreads) k_observe(DevState d,
                                                         const uint32_t* __restrict__ amp_num,
                                                         const double* __restrict__ dl,
                                                         double* __restrict__ part) {
#pragma clang fp contract(off)
  const int r = blockIdx.y;
  const uint32_t b0 = d.rep_off[r], n_r = d.rep_off[r + 1] - b0;
  const uint32_t na = amp_num[r];
  const uint32_t cap = d.cap;
  KSum acc[13];
#pragma unroll
  for (int q = 0; q < 13; ++q) acc[q] = KSum{0.0, 0.0};
  for (uint32_t k = blockIdx.x * kRedThreads + threadIdx.x; k < n_r; k += kRedBlocks * kRedThreads) {
    const uint32_t i = b0 + k;
    double dd = 0.0;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      const double v = d.vel[a * cap + i];
      const double pr = (double)d.press[a * cap + i];
      const double dl = d.delta[a * cap + i];
      acc[a].add(v * v);
      acc[3 + a].add(pr * 0.5 + v * v);
      acc[6 + a].add(pr);
      dd += dl * dl;
    }
    acc[9].add(dd);
    acc[12].add((double)min(d.ncount[i], (uint32_t)d.W));
    // the reference indexes beads by their load order: solvent loop from
    // index amp_num (observer.hpp:171), amphiphile a = beads 4a..4a+3 (:163-164)
    if (d.orig[i] - b0 >= na) acc[10].add(dd);
    if ((k & 3u) == 0 && (k >> 2) < na && k + 3 < n_r) {
      // load-order beads b0 + k .. b0 + k + 3
      const double* q = dl + b0 + k;
      double m[3];
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        const double* qa = q + a * (size_t)cap;
        m[a] = qa[0] + qa[1];
        m[a] += qa[2];
        m[a] += qa[3];
        m[a] /= 4.0;
      }
      acc[11].add(m[0] * m[0] + m[1] * m[1] + m[2] * m[2]);
    }
  }
  block_reduce_write<13>(acc, part + ((size_t)r * kRedBlocks + blockIdx.x) * 2 * 13);
}

void launch_observe(const DevState& d, const uint32_t* amp_num, double* dl_load, double* out,
                    hipStream_t st) {
  hipLaunchKernelGGL(k_delta_load_order, grid1(d.nl, 256), dim3(256), 0, st, d, dl_load);
  hipLaunchKernelGGL(k_observe, dim3(kRedBlocks, d.n_rep), dim3(kRedThreads), 0, st, d, amp_num,
                     (const double*)dl_load, d.red);
  hipLaunchKernelGGL(k_red_final<13>, dim3(d.n_rep), dim3(64), 0, st, d.red, out,
                     (const uint32_t*)nullptr);
}

// ---------------------------------------------------------------- bonded
// One thread per bead: the bead evaluates its own chain (four beads, a few
// hundred FP64 flops) and keeps its own row, so every write is to the bead's
// own slot; then kick (driftkick.hpp:32-40) from the same registers.  The
// chain virial is counted once, by the chain's first bead.
__global__ void __launch_bounds__(256) k_bonded_kick(DevState d, int virial, int kick,
                                                     double* __restrict__ vir_bead) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= dev_n(d)) return;
  const int r = d.n_rep > 1 ? find_rep(d.rep_off, d.n_rep, s) : 0;
  float af[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) af[c] = d.acc[c * d.cap + s];
  bonded_kick_bead(d, d.reps[r], s, af, virial != 0, kick != 0, vir_bead);
}

// chain-major FP64 positions of every chain bead from the master storage
// (the step's permutation writes them as it moves the beads)
__global__ void __launch_bounds__(256) k_chainpos_fill(DevState d) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= dev_n(d)) return;
  const uint32_t ch = d.chain[s];
  if (ch != 0xffffffffu)
    d.chainpos[ch] = make_double4(d.pos[s], d.pos[d.cap + s], d.pos[2 * (size_t)d.cap + s], 0.0);
}
void launch_chainpos_fill(const DevState& d, hipStream_t st) {
  if (!d.nl || !d.n_amp) return;
  hipLaunchKernelGGL(k_chainpos_fill, grid1(d.nl, 256), dim3(256), 0, st, d);
}

// per-replica sum of the per-bead chain virials
__global__ void __launch_bounds__(kRedThreads) k_bead_vir(DevState d, const double* __restrict__ vb,
                                                          double* __restrict__ part) {
  if (halted(d)) return;
  const int r = blockIdx.y;
  const uint32_t b0 = d.rep_off[r], nr = d.rep_off[r + 1] - b0;
  KSum acc[3] = {{0, 0}, {0, 0}, {0, 0}};
  for (uint32_t k = blockIdx.x * kRedThreads + threadIdx.x; k < nr; k += kRedBlocks * kRedThreads)
#pragma unroll
    for (int c = 0; c < 3; ++c) acc[c].add(vb[3 * (size_t)(b0 + k) + c]);
  block_reduce_write<3>(acc, part + ((size_t)r * kRedBlocks + blockIdx.x) * 6);
}

double* bead_vir_buffer(const DevState& d) {
  return d.red + (size_t)d.n_rep * kRedBlocks * 2 * 16;
}

void launch_bead_vir_reduce(const DevState& d, double* vir_out, hipStream_t st) {
  if (d.n_amp) {
    hipLaunchKernelGGL(k_bead_vir, dim3(kRedBlocks, d.n_rep), dim3(kRedThreads), 0, st, d,
                       (const double*)bead_vir_buffer(d), d.red);
    hipLaunchKernelGGL(k_red_final<3>, dim3(d.n_rep), dim3(64), 0, st, d.red, vir_out,
                       (const uint32_t*)d.run_ok);
  } else {
    hipLaunchKernelGGL(k_fill_d, dim3(1), dim3(256), 0, st, vir_out, (size_t)3 * d.n_rep, 0.0);
  }
}

void launch_bonded_kick(const DevState& d, bool virial, bool kick, double* vir_out,
                        hipStream_t st) {
  if (!d.nl) return;
  if (d.n_amp || kick)
    hipLaunchKernelGGL(k_bonded_kick, grid1(d.nl, 256), dim3(256), 0, st, d, virial ? 1 : 0,
                       kick ? 1 : 0, bead_vir_buffer(d));
  if (virial) launch_bead_vir_reduce(d, vir_out, st);
}

// ---------------------------------------------------------------- reactions
// head conversion: unit-0 bead of type conv_from -> conv_to when
// Saru(amp_id, time, chem_seed).uni_open() <= conv_p
__global__ void __launch_bounds__(256) k_head_conv(DevState d) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= dev_n(d)) return;
  if (d.unit[i] != 0u) return;
  const int r = d.n_rep > 1 ? find_rep(d.rep_off, d.n_rep, i) : 0;
  const Replica& R = d.reps[r];
  if (R.conv_p <= 0.0 || d.prop[i] != (uint32_t)R.conv_from) return;
  Saru s = saru_seed(d.amp_id[i], *d.d_time, R.chem_seed);
  if (saru_uni_f64(s) <= R.conv_p) {
    d.prop[i] = (uint32_t)R.conv_to;
    // the type also rides in the FP32 copy (read by the pair passes before
    // the next rebuild in skin mode)
    const float w = __uint_as_float((d.id[i] << 2) | ((uint32_t)R.conv_to & 3u));
    reinterpret_cast<float*>(&d.spos[i])[3] = w;
    reinterpret_cast<float*>(&d.sb[i].pos)[3] = w;
  }
}
void launch_head_conversion(const DevState& d, hipStream_t st) {
  if (!d.nl) return;
  hipLaunchKernelGGL(k_head_conv, grid1(d.nl, 256), dim3(256), 0, st, d);
}


// ---------------------------------------------------------------- skin mode
// drift_and_predict (driftkick.hpp:5-30, the same FP64 expressions as
// k_drift) in place, plus what the pair passes read between rebuilds: the
// FP32 position relative to the cell of the last build (minimum image, so
// a bead that crossed the box edge stays next to its cell), the predicted
// velocity and the chain-major copy of chain beads.  A bead that has moved
// more than skin / 2 since the last build raises the rebuild word.
__global__ void __launch_bounds__(256) k_drift_skin(DevState d) {
#pragma clang fp contract(off)
  if (blockIdx.x == 0 && threadIdx.x == 0) *d.d_time += 1u;  // the step's clock
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= dev_n(d)) return;
  const int r = d.n_rep > 1 ? find_rep(d.rep_off, d.n_rep, i) : 0;
  const Replica& R = d.reps[r];
  const uint32_t cap = d.cap;
  const uint32_t c = d.scell[i] - R.cell_off;
  const uint32_t nx = (uint32_t)R.nc[0], ny = (uint32_t)R.nc[1];
  const uint32_t cc[3] = {c % nx, (c / nx) % ny, c / (nx * ny)};
  double p[3], disp = 0.0;
  float l[3], v[3];
  bool fin = true;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const double vv = d.vel[a * cap + i];
    const DriftAxis o = drift_axis(R, a, d.pos[a * cap + i], vv, (double)d.acc[a * cap + i],
                                   d.delta[a * cap + i]);
    const double dl = o.dl, vb = o.vb;
    p[a] = o.p;
    d.delta[a * cap + i] = dl;
    d.pos[a * cap + i] = p[a];
    d.vbuf[a * cap + i] = vb;
    fin = fin && isfinite(p[a]) && isfinite(vv);
    const double e = dl - d.dref[a * cap + i];
    disp += e * e;
    double x = p[a] - (double)cc[a] * R.h[a];
    if (x > 0.5 * R.box[a]) x -= R.box[a];
    else if (x < -0.5 * R.box[a]) x += R.box[a];
    l[a] = a == 0 ? __uint_as_float(fix_x(x, cc[0], R.inv_h[0])) : (float)x;
    v[a] = (float)vb;
  }
  if (!fin) atomicOr(d.flags, FLAG_NONFINITE);
  if (!(disp <= d.disp2)) *d.rebuild = 1u;  // every writer stores the same value
  // .w (id, type) stays
  float* sp = reinterpret_cast<float*>(&d.spos[i]);
  float* sq = reinterpret_cast<float*>(&d.sb[i].pos);
  sp[0] = sq[0] = l[0];
  sp[1] = sq[1] = l[1];
  sp[2] = sq[2] = l[2];
  d.sb[i].vel = make_float4(v[0], v[1], v[2], 0.f);
  const uint32_t ch = d.chain[i];
  if (ch != 0xffffffffu) d.chainpos[ch] = make_double4(p[0], p[1], p[2], 0.0);
}
void launch_drift_skin(const DevState& d, hipStream_t st) {
  if (!d.nl) return;
  hipLaunchKernelGGL(k_drift_skin, grid1(d.nl, 256), dim3(256), 0, st, d);
}

void launch_keys_gated(const DevState& d, hipStream_t st) {
  if (!d.nl) return;
  hipLaunchKernelGGL(k_keys, grid1(d.nl, 256), dim3(256), 0, st, d, 1);
}
void launch_scan_cells_gated(const DevState& d, hipStream_t st) { scan_cells(d, d.rebuild, st); }
void launch_scatter_gated(const DevState& d, hipStream_t st) {
  if (!d.nl) return;
  hipLaunchKernelGGL(k_scatter, grid1(d.nl, 256), dim3(256), 0, st, d, 1);
}
void launch_cell_sort_gated(const DevState& d, hipStream_t st) {
  hipLaunchKernelGGL(k_cell_sort, grid1(d.ncell, 256), dim3(256), 0, st, d, 1);
}
void launch_permute_gated(const DevState& d, hipStream_t st) {
  if (!d.nl) return;
  hipLaunchKernelGGL(k_permute, grid1(d.nl, 256), dim3(256), 0, st, d, 1, 0, 1);
}

// The rebuild's permutation wrote the other buffer set; copy it back so the
// master storage keeps one address whether or not this step rebuilt (the
// captured step cannot know).  vel, acc and press are rewritten by the
// force pass of the step and need no copy.
__global__ void __launch_bounds__(256) k_copy_back(DevState d) {
  if (*d.rebuild == 0u) return;
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= dev_n(d)) return;
  const uint32_t cap = d.cap;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    d.pos[a * cap + s] = d.alt.pos[a * cap + s];
    d.vbuf[a * cap + s] = d.alt.vbuf[a * cap + s];
    d.delta[a * cap + s] = d.alt.delta[a * cap + s];
  }
  d.id[s] = d.alt.id[s];
  d.prop[s] = d.alt.prop[s];
  d.amp_id[s] = d.alt.amp_id[s];
  d.unit[s] = d.alt.unit[s];
  d.orig[s] = d.alt.orig[s];
  d.chain[s] = d.alt.chain[s];
}
void launch_copy_back_gated(const DevState& d, hipStream_t st) {
  if (!d.nl) return;
  hipLaunchKernelGGL(k_copy_back, grid1(d.nl, 256), dim3(256), 0, st, d);
}

// bead counts along one axis per replica
__global__ void __launch_bounds__(256) k_profile(DevState d, int axis, int nbins, uint32_t* hist) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= dev_n(d)) return;
  const int r = d.n_rep > 1 ? find_rep(d.rep_off, d.n_rep, i) : 0;
  const Replica& R = d.reps[r];
  int b = (int)floor(d.pos[axis * d.cap + i] * R.inv_box[axis] * nbins);
  b = b < 0 ? 0 : (b >= nbins ? nbins - 1 : b);
  atomicAdd(&hist[(size_t)r * nbins + b], 1u);
}
void launch_density_profile(const DevState& d, int axis, int nbins, uint32_t* hist, hipStream_t st) {
  if (!d.nl) return;
  hipLaunchKernelGGL(k_profile, grid1(d.nl, 256), dim3(256), 0, st, d, axis, nbins, hist);
}

// ---------------------------------------------------------------- neighbour CM
// CalcDensity's centre-of-mass sums (nbdf_calculator.hpp:22-58) for a list
// of beads, FP64 from the master positions: hydrophilic CM of the bead and
// its Hyphil neighbours with r^2 < 1, hydrophobic CM of the Hyphob
// neighbours with r^2 < 1.44, j at its image position r_i - d.
__device__ __forceinline__ double fp64_disp(const DevState& d, const Replica& R, uint32_t i,
                                            uint32_t j, double dv[3]) {
#pragma clang fp contract(off)
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    double x = d.pos[a * d.cap + i] - d.pos[a * d.cap + j];
    if (x > 0.5 * R.box[a]) x -= R.box[a];
    else if (x < -0.5 * R.box[a]) x += R.box[a];
    dv[a] = x;
  }
  return dv[0] * dv[0] + dv[1] * dv[1] + dv[2] * dv[2];
}

// out[0..2] hydrophilic CM, out[3..5] hydrophobic CM, out[6..7] counts
__device__ void neighbor_cm_at(const DevState& d, uint32_t i, double out[8]) {
#pragma clang fp contract(off)
  const int r = d.n_rep > 1 ? find_rep(d.rep_off, d.n_rep, i) : 0;
  const Replica& R = d.reps[r];
  double ri[3], hl[3], hb[3] = {0.0, 0.0, 0.0};
  for (int a = 0; a < 3; ++a) { ri[a] = d.pos[a * d.cap + i]; hl[a] = ri[a]; }
  uint32_t cl = 1, cb = 0;
  int rad[3];
  bool brute = false;
  for (int a = 0; a < 3; ++a) {
    // beads may sit up to cell_slack / 2 outside their cell (skin mode)
    rad[a] = (int)ceil((1.2 + d.cell_slack) * R.inv_h[a] - 1e-12);
    if (2 * rad[a] + 1 > R.nc[a]) brute = true;
  }
  auto visit = [&](uint32_t j) {
    if (j == i) return;
    double dv[3];
    const double r2 = fp64_disp(d, R, i, j, dv);
    if (r2 == 0.0) return;
    const uint32_t pj = d.prop[j];
    if (r2 < 1.44 && pj == 1u) { for (int a = 0; a < 3; ++a) hb[a] += ri[a] - dv[a]; cb++; }
    if (r2 < 1.0 && pj == 0u) { for (int a = 0; a < 3; ++a) hl[a] += ri[a] - dv[a]; cl++; }
  };
  if (brute) {
    const uint32_t b0 = d.rep_off[r], b1 = d.rep_off[r + 1];
    for (uint32_t j = b0; j < b1; ++j) visit(j);
  } else {
    const uint32_t c = d.scell[i] - R.cell_off;
    const int nx = R.nc[0], ny = R.nc[1], nz = R.nc[2];
    const int cx = (int)(c % (uint32_t)nx), cy = (int)((c / (uint32_t)nx) % (uint32_t)ny),
              cz = (int)(c / (uint32_t)(nx * ny));
    for (int oz = -rad[2]; oz <= rad[2]; ++oz)
      for (int oy = -rad[1]; oy <= rad[1]; ++oy)
        for (int ox = -rad[0]; ox <= rad[0]; ++ox) {
          const int qx = (cx + ox + nx) % nx, qy = (cy + oy + ny) % ny, qz = (cz + oz + nz) % nz;
          const uint32_t q = R.cell_off + (uint32_t)((qz * ny + qy) * nx + qx);
          for (uint32_t t = d.cell_start[q]; t < d.cell_start[q + 1]; ++t) visit(t);
        }
  }
  for (int a = 0; a < 3; ++a) {
    out[a] = hl[a] / (double)cl;
    out[3 + a] = hb[a] / (double)cb;
  }
  out[6] = (double)cl;
  out[7] = (double)cb;
}

__global__ void k_neighbor_cm(DevState d, const uint32_t* __restrict__ beads, uint32_t nb,
                              double* __restrict__ out) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nb) return;
  double o[8];
  neighbor_cm_at(d, beads[k], o);
  for (int q = 0; q < 8; ++q) out[8 * (size_t)k + q] = o[q];
}
void launch_neighbor_cm(const DevState& d, const uint32_t* beads, uint32_t nb, double* out,
                        hipStream_t st) {
  if (!nb) return;
  hipLaunchKernelGGL(k_neighbor_cm, grid1(nb, 64), dim3(64), 0, st, d, beads, nb, out);
}

// ---------------------------------------------------------------- Saru check
__global__ void k_saru_test(const uint32_t* s1, const uint32_t* s2, const uint32_t* s3, uint32_t n,
                            uint32_t* u32out, float* nrml_out) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  Saru a = saru_seed(s1[k], s2[k], s3[k]);
  for (int q = 0; q < 8; ++q) u32out[8 * (size_t)k + q] = saru_u32(a);
  Saru b = saru_seed(s1[k], s2[k], s3[k]);
  nrml_out[k] = saru_nrml_f32(b);
}
void launch_saru_test(const uint32_t* s1, const uint32_t* s2, const uint32_t* s3, uint32_t n,
                      uint32_t* u32out, float* nrml_out, hipStream_t st) {
  if (!n) return;
  hipLaunchKernelGGL(k_saru_test, grid1(n, 256), dim3(256), 0, st, s1, s2, s3, n, u32out, nrml_out);
}

// ---------------------------------------------------------------- chemistry
// ChemManager::RandomChemEvent, serial version (chemmanager.hpp:244-281),
// with DetermineTargetIdWithCorePos (:116-164), GetCorePos (:84-114),
// GetMembraneTangRandomVec (:166-173) and CreateAmpPart (:35-67), entirely
// on the device so that it can sit inside a captured step:
//   k_chem_find    storage slot of every core head (bead ids in cd->core_id)
//   k_chem_core    per core: neighbour CMs, membrane normal h2t, the head
//                  moved along h2t by the projection of its hydrophilic CM
//   k_chem_flags   one thread per amphiphile, every core in turn: is its
//                  head in the core's truncated cone with a draw <=
//                  p_thresld; the flags as one 64-bit mask per (core, wave),
//                  their count per (core, block of 256)
//   k_chem_scan    one block: exclusive offsets of those counts in (core,
//                  amphiphile) order, the capacity check, the stop test
//   k_chem_scatter the targets' amphiphile indices at those offsets, from
//                  the masks
//   k_chem_insert  four new beads per target, appended at n0 + 4t + u
//   k_chem_commit  the new bead and chain counts (rep_off / amp_off on the
//                  device), the stop of the run, the core slots reset
// Head and unit-2 positions come from the chain-major FP64 copy (chainpos,
// written by every permutation), so no topology pass is needed.  Draws are
// keyed Saru streams: the cone test (amphiphile j, time, chem_seed +
// TAG_CONE (core + 1)), the tangent angle (new amphiphile id, time,
// chem_seed ^ TAG_TANG), the velocities (new bead id, time, chem_seed ^
// TAG_VEL).  `time` is the device clock.  In the reference-MT mode (ChemMt)
// the draws are the outputs of one MT19937 stream instead, in the
// reference's order: k_chem_flags first marks and counts the heads inside
// each cone, k_chem_scan turns the counts into each head's rank, k_chem_mt
// draws one genrand_real1 per in-cone head in rank order, k_chem_decide
// turns the cone masks into target masks by those decisions, and after the
// target scan k_chem_mt draws the targets' 25 outputs each for
// k_chem_insert.
constexpr int kChemBlock = 256;

__global__ void __launch_bounds__(256) k_chem_find(DevState d, const ChemDev* __restrict__ cd,
                                                   ChemState* __restrict__ cs) {
  if (halted(d)) return;
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= dev_n(d)) return;
  const uint32_t id = d.id[s], nc = cd->n_core;
  for (uint32_t c = 0; c < nc; ++c)
    if (cd->core_id[c] == id) cs->core_slot[c] = s;
}

// one block of kCoreBlock threads per core: the neighbour sums of the core
// head spread over the threads (P threads per cell of its window, each
// taking every P-th bead of the cell, or beads on the all-pairs path), then
// a fixed-order tree reduction, so the result does not depend on scheduling
constexpr int kCoreBlock = 512;
__global__ void __launch_bounds__(kCoreBlock) k_chem_core(DevState d, const ChemDev* __restrict__ cd,
                                                          ChemState* __restrict__ cs) {
#pragma clang fp contract(off)
  if (halted(d)) return;
  const uint32_t c = blockIdx.x;
  if (c >= cd->n_core) return;
  __shared__ double red[8][kCoreBlock];
  const Replica& R = d.reps[0];
  const uint32_t slot = cs->core_slot[c];
  const int tid = threadIdx.x;
  double sl[3] = {0.0, 0.0, 0.0}, sb[3] = {0.0, 0.0, 0.0}, nl = 0.0, nb = 0.0;
  if (slot != 0xffffffffu) {
    double ri[3];
    for (int a = 0; a < 3; ++a) ri[a] = d.pos[a * d.cap + slot];
    auto visit = [&](uint32_t j) {
      if (j == slot) return;
      double dv[3];
      const double r2 = fp64_disp(d, R, slot, j, dv);
      if (r2 == 0.0) return;
      const uint32_t pj = d.prop[j];
      if (r2 < 1.44 && pj == 1u) { for (int a = 0; a < 3; ++a) sb[a] += ri[a] - dv[a]; nb += 1.0; }
      if (r2 < 1.0 && pj == 0u) { for (int a = 0; a < 3; ++a) sl[a] += ri[a] - dv[a]; nl += 1.0; }
    };
    int rad[3];
    bool brute = false;
    for (int a = 0; a < 3; ++a) {
      rad[a] = (int)ceil((1.2 + d.cell_slack) * R.inv_h[a] - 1e-12);
      if (2 * rad[a] + 1 > R.nc[a]) brute = true;
    }
    if (brute) {
      const uint32_t b1 = d.rep_off[1];
      for (uint32_t j = (uint32_t)tid; j < b1; j += kCoreBlock) visit(j);
    } else {
      const uint32_t cc = d.scell[slot] - R.cell_off;
      const int nx = R.nc[0], ny = R.nc[1], nz = R.nc[2];
      const int cx = (int)(cc % (uint32_t)nx), cy = (int)((cc / (uint32_t)nx) % (uint32_t)ny),
                cz = (int)(cc / (uint32_t)(nx * ny));
      const int wx = 2 * rad[0] + 1, wy = 2 * rad[1] + 1, wz = 2 * rad[2] + 1;
      const int wn = wx * wy * wz;
      const int P = wn >= kCoreBlock ? 1 : kCoreBlock / wn;
      for (int w = tid; w < wn * P; w += kCoreBlock) {
        const int cell = w % wn, k0 = w / wn;
        const int ox = cell % wx - rad[0], oy = (cell / wx) % wy - rad[1], oz = cell / (wx * wy) - rad[2];
        const int qx = (cx + ox + nx) % nx, qy = (cy + oy + ny) % ny, qz = (cz + oz + nz) % nz;
        const uint32_t q = R.cell_off + (uint32_t)((qz * ny + qy) * nx + qx);
        for (uint32_t t = d.cell_start[q] + (uint32_t)k0; t < d.cell_start[q + 1]; t += (uint32_t)P)
          visit(t);
      }
    }
  }
  for (int a = 0; a < 3; ++a) {
    red[a][tid] = sl[a];
    red[3 + a][tid] = sb[a];
  }
  red[6][tid] = nl;
  red[7][tid] = nb;
  __syncthreads();
  for (int h = kCoreBlock / 2; h > 0; h >>= 1) {
    if (tid < h)
      for (int q = 0; q < 8; ++q) red[q][tid] += red[q][tid + h];
    __syncthreads();
  }
  if (tid != 0) return;
  // GetCorePos: (-1, -1, -1) when the core head is absent
  double h[3] = {-1.0, -1.0, -1.0}, t[3] = {-1.0, -1.0, -1.0}, m[3] = {-1.0, -1.0, -1.0};
  if (slot != 0xffffffffu) {
    // the hydrophilic sum starts with the head itself (count 1)
    const double cl = 1.0 + red[6][0], cb = red[7][0];
    for (int a = 0; a < 3; ++a) {
      h[a] = d.pos[a * d.cap + slot];
      m[a] = (h[a] + red[a][0]) / cl;
      t[a] = red[3 + a][0] / cb;
    }
  }
  double h2t[3] = {t[0] - h[0], t[1] - h[1], t[2] - h[2]};
  bd_min_image(R, h2t);
  double l = sqrt(h2t[0] * h2t[0] + h2t[1] * h2t[1] + h2t[2] * h2t[2]);
  for (int a = 0; a < 3; ++a) h2t[a] /= l;
  double c2cm[3] = {m[0] - h[0], m[1] - h[1], m[2] - h[2]};
  bd_min_image(R, c2cm);
  const double prj = c2cm[0] * h2t[0] + c2cm[1] * h2t[1] + c2cm[2] * h2t[2];
  for (int a = 0; a < 3; ++a) h[a] += prj * h2t[a];
  for (int a = 0; a < 3; ++a) h[a] -= floor(h[a] * R.inv_box[a]) * R.box[a];
  for (int a = 0; a < 3; ++a) {
    cs->core_geo[c][a] = h[a];
    cs->core_geo[c][3 + a] = h2t[a];
  }
}

// DetermineTargetIdWithCorePos loop body (chemmanager.hpp:140-161): is the
// head of chain j inside the truncated cone of core c?
__device__ __forceinline__ bool chem_in_cone(const DevState& d, const ChemDev* __restrict__ cd,
                                             const ChemState* __restrict__ cs, uint32_t c,
                                             const double4& hp) {
#pragma clang fp contract(off)
  const Replica& R = d.reps[0];
  const double* ch = cs->core_geo[c];
  const double* h2t = ch + 3;
  double c2p[3] = {hp.x - ch[0], hp.y - ch[1], hp.z - ch[2]};
  bd_min_image(R, c2p);
  const double depth = h2t[0] * c2p[0] + h2t[1] * c2p[1] + h2t[2] * c2p[2];
  double tg[3];
  for (int a = 0; a < 3; ++a) tg[a] = c2p[a] - depth * h2t[a];
  const double rad = sqrt(tg[0] * tg[0] + tg[1] * tg[1] + tg[2] * tg[2]);
  const double rad_thr = cd->influ_grd * depth + cd->influ_rad;
  return (rad <= rad_thr) && (depth >= -cd->influ_hei) && (depth <= cd->influ_dep);
}

// rank of this thread's flag among the set flags of the block, threads in
// order (every thread of the block calls it; wcnt: kChemBlock / 64 words of
// LDS not in use by the block)
__device__ __forceinline__ uint32_t chem_block_rank(bool f, uint32_t* wcnt) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const unsigned long long bal = __ballot(f);
  if (lane == 0) wcnt[w] = (uint32_t)__popcll(bal);
  __syncthreads();
  uint32_t r = (uint32_t)__popcll(bal & ((1ull << lane) - 1ull));
  for (int k = 0; k < w; ++k) r += wcnt[k];
  return r;
}

// one thread per amphiphile j (grid over the chain capacity), every core c
// in turn: kMode 0 the keyed target test, 1 the cone test alone
// (reference-MT mode).  A head outside the cone keeps rnd = 2.0, a target
// only if p_thresld >= 2.  The flags go to bit j % 64 of
// masks[c * words + j / 64], their count per block to counts[c * gridDim.x +
// blockIdx.x]; the head position is read once for all cores.
template <int kMode>
__global__ void __launch_bounds__(kChemBlock) k_chem_flags(DevState d, const ChemDev* __restrict__ cd,
                                                           const ChemState* __restrict__ cs,
                                                           uint32_t* __restrict__ counts,
                                                           unsigned long long* __restrict__ masks) {
  if (halted(d)) return;
  const uint32_t j = blockIdx.x * kChemBlock + threadIdx.x;
  const bool live = j < dev_namp(d);
  const double4 hp = live ? d.chainpos[4 * (size_t)j] : make_double4(0.0, 0.0, 0.0, 0.0);
  const uint32_t nc = cd->n_core, time = *d.d_time, seed = d.reps[0].chem_seed;
  const size_t words = (size_t)gridDim.x * (kChemBlock / 64);
  const double p = cd->p_thresld;
  for (uint32_t c = 0; c < nc; ++c) {
    const bool in = live && chem_in_cone(d, cd, cs, c, hp);
    bool f = in;
    if (kMode == 0) {
      double rnd = 2.0;
      if (in) {
        Saru q = saru_seed(j, time, seed + TAG_CONE * (c + 1u));
        rnd = saru_uni_f64(q);
      }
      f = live && rnd <= p;
    }
    const unsigned long long bal = __ballot(f);
    if ((threadIdx.x & 63u) == 0u) masks[(size_t)c * words + (j >> 6)] = bal;
    const int cnt = __syncthreads_count(f ? 1 : 0);
    if (threadIdx.x == 0) counts[(size_t)c * gridDim.x + blockIdx.x] = (uint32_t)cnt;
  }
}

// reference-MT mode, grid (chain blocks, cores): the cone masks become
// target masks in place, an in-cone head taking the decision of its rank
// (mt.cone_offs + its rank in the block) from mt.cone_bits; the targets'
// count per block to counts.  Each wave reads and rewrites only its own word.
__global__ void __launch_bounds__(kChemBlock) k_chem_decide(DevState d, const ChemDev* __restrict__ cd,
                                                            ChemMt mt, uint32_t* __restrict__ counts,
                                                            unsigned long long* __restrict__ masks) {
  if (halted(d)) return;
  __shared__ uint32_t wcnt[kChemBlock / 64];
  const uint32_t j = blockIdx.x * kChemBlock + threadIdx.x, c = blockIdx.y;
  const uint32_t lane = threadIdx.x & 63u;
  unsigned long long* mw = masks + (size_t)c * gridDim.x * (kChemBlock / 64) + (j >> 6);
  const bool in = ((*mw >> lane) & 1ull) != 0ull;
  const bool live = j < dev_namp(d);
  const uint32_t r = mt.cone_offs[(size_t)c * gridDim.x + blockIdx.x] + chem_block_rank(in, wcnt);
  const bool f = in ? ((mt.cone_bits[r >> 5] >> (r & 31u)) & 1u) != 0u
                    : (live && 2.0 <= cd->p_thresld);
  const unsigned long long bal = __ballot(f);
  if (lane == 0u) *mw = bal;
  const int cnt = __syncthreads_count(f ? 1 : 0);
  if (threadIdx.x == 0) counts[(size_t)c * gridDim.x + blockIdx.x] = (uint32_t)cnt;
}

// cone: the counts are in-cone heads (reference-MT mode), whose total is
// the number of draws of the cone tests
__global__ void __launch_bounds__(1024) k_chem_scan(DevState d, const ChemDev* __restrict__ cd,
                                                    ChemState* __restrict__ cs,
                                                    uint32_t* __restrict__ counts, uint32_t m,
                                                    int cone) {
  if (halted(d)) return;
  __shared__ uint32_t wsum[16];
  const uint32_t per = (m + 1023u) / 1024u;
  const uint32_t b0 = threadIdx.x * per, b1 = min(b0 + per, m);
  uint32_t tsum = 0;
  for (uint32_t k = b0; k < b1; ++k) tsum += counts[k];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const uint32_t inc = wave_incl_scan(tsum);
  if (lane == 63) wsum[w] = inc;
  __syncthreads();
  uint32_t off = 0, tot = 0;
  for (int k = 0; k < 16; ++k) {
    if (k < w) off += wsum[k];
    tot += wsum[k];
  }
  uint32_t run = off + inc - tsum;
  for (uint32_t k = b0; k < b1; ++k) {
    const uint32_t v = counts[k];
    counts[k] = run;
    run += v;
  }
  if (threadIdx.x == 0 && cone) cs->n_in = tot;
  if (threadIdx.x == 0 && !cone) {
    const uint32_t n0 = dev_n(d), na0 = dev_namp(d);
    // amp_cap = cap / 4 + 1 chains fit the chain arrays
    const bool fits = (uint64_t)n0 + 4ull * tot <= (uint64_t)d.cap &&
                      (uint64_t)na0 + tot <= (uint64_t)d.cap / 4u + 1u;
    cs->n0 = n0;
    cs->namp0 = na0;
    cs->nt = fits ? tot : 0u;
    cs->refused = fits ? 0u : 1u;
    cs->stop = (fits && na0 + tot >= cd->max_amp_num) ? 1u : 0u;
  }
}

// grid (chain blocks, cores): the targets' amphiphile indices at their
// offsets, from the target masks
__global__ void __launch_bounds__(kChemBlock) k_chem_scatter(DevState d, const ChemState* __restrict__ cs,
                                                             const unsigned long long* __restrict__ masks,
                                                             const uint32_t* __restrict__ offs,
                                                             uint32_t* __restrict__ targets) {
  if (halted(d) || cs->nt == 0u) return;
  __shared__ uint32_t wcnt[kChemBlock / 64];
  const uint32_t j = blockIdx.x * kChemBlock + threadIdx.x, c = blockIdx.y;
  const unsigned long long m = masks[(size_t)c * gridDim.x * (kChemBlock / 64) + (j >> 6)];
  const bool f = ((m >> (threadIdx.x & 63u)) & 1ull) != 0ull;
  const uint32_t r = chem_block_rank(f, wcnt);
  if (f) targets[offs[(size_t)c * gridDim.x + blockIdx.x] + r] = j;
}

// ---- MT19937 (mt19937ar: init_genrand / genrand_int32 of M. Matsumoto and
// T. Nishimura, ACM TOMACS 8 (1998) 3-30, the 2002 initialisation), the
// generator behind the reference's PS::MT.  One block keeps the 624-word
// state in LDS and serves a stream of outputs in order.
constexpr int kMtBlock = 640;  // ten waves: one thread per state word
__device__ __forceinline__ uint32_t mt_temper(uint32_t y) {
  y ^= y >> 11;
  y ^= (y << 7) & 0x9d2c5680u;
  y ^= (y << 15) & 0xefc60000u;
  y ^= y >> 18;
  return y;
}
// the twist of genrand_int32 (word k from words k, k + 1 and k + 397 mod 624
// as they stand when the serial loop reaches k) in three bands of parallel
// words: [0, 227) reads only old words; [227, 454) reads k - 227 from the
// first band; [454, 624) reads k - 227 from the second and, for k = 623,
// word 0 from the first.  Every thread of the block calls it.
__device__ void mt_twist(uint32_t* s) {
  const int t = threadIdx.x;
  for (int band = 0; band < 3; ++band) {
    const int k0 = band * 227, k1 = band == 2 ? (int)kMtN : k0 + 227;
    const int k = k0 + t;
    uint32_t v = 0;
    if (k < k1) {
      const uint32_t y = (s[k] & 0x80000000u) | (s[(k + 1) % (int)kMtN] & 0x7fffffffu);
      v = s[(k + 397) % (int)kMtN] ^ (y >> 1) ^ ((y & 1u) ? 0x9908b0dfu : 0u);
    }
    __syncthreads();
    if (k < k1) s[k] = v;
    __syncthreads();
  }
}
// genrand_real1: [0, 1]; genrand_real3: (0, 1)
__device__ __forceinline__ double mt_real1(uint32_t y) { return (double)y * (1.0 / 4294967295.0); }
__device__ __forceinline__ double mt_real3(uint32_t y) {
  return ((double)y + 0.5) * (1.0 / 4294967296.0);
}

// one block: the next outputs of the stream.  targets = 0: the n_in cone
// draws, bit r of mt.cone_bits = (genrand_real1 of draw r <= p_thresld)
// (the bits start cleared); targets = 1: the targets' 25 nt raw outputs
// into mt.tdraws (none when the event was refused)
__global__ void __launch_bounds__(kMtBlock) k_chem_mt(DevState d, const ChemDev* __restrict__ cd,
                                                     const ChemState* __restrict__ cs, ChemMt mt,
                                                     int targets) {
#pragma clang fp contract(off)
  if (halted(d)) return;
  const uint32_t n = targets ? (cs->refused ? 0u : kMtDrawsPerTarget * cs->nt) : cs->n_in;
  if (n == 0u) return;
  __shared__ uint32_t s[kMtN];
  MtState* g = mt.g;
  const int t = threadIdx.x;
  if (t < (int)kMtN) s[t] = g->mt[t];
  uint32_t mti = g->mti;
  __syncthreads();
  const double p = cd->p_thresld;
  for (uint32_t done = 0; done < n;) {
    if (mti >= kMtN) {
      mt_twist(s);
      mti = 0;
    }
    const uint32_t cnt = min(kMtN - mti, n - done);
    const bool act = (uint32_t)t < cnt;
    const uint32_t y = act ? mt_temper(s[mti + t]) : 0u;
    if (targets) {
      if (act) mt.tdraws[done + t] = y;
    } else {
      // 64 draws per wave: draws r0 .. r0 + 63 straddle at most three words
      const unsigned long long bal = __ballot(act && mt_real1(y) <= p);
      if ((t & 63) == 0 && bal) {
        const uint32_t r0 = done + (uint32_t)t, w0 = r0 >> 5, sh = r0 & 31u;
        const unsigned long long lo = bal << sh;
        const uint32_t hi = sh ? (uint32_t)(bal >> (64u - sh)) : 0u;
        if ((uint32_t)lo) atomicOr(&mt.cone_bits[w0], (uint32_t)lo);
        if ((uint32_t)(lo >> 32)) atomicOr(&mt.cone_bits[w0 + 1], (uint32_t)(lo >> 32));
        if (hi) atomicOr(&mt.cone_bits[w0 + 2], hi);
      }
    }
    done += cnt;
    mti += cnt;
  }
  __syncthreads();
  if (t < (int)kMtN) g->mt[t] = s[t];
  if (t == 0) {
    g->mti = mti;
    g->draws += n;
  }
}

// one thread per new bead (grid-stride): target t makes amphiphile namp0 + t
// of beads n0 + 4t + u.  With `keys` (this step's force pass binned every
// other bead for the next step) the new bead's cell key and histogram entry
// of its next drift are added too; a run that stops here clears the
// histogram instead (the next run bins every bead itself).
__global__ void __launch_bounds__(256) k_chem_insert(DevState d, const ChemDev* __restrict__ cd,
                                                     const ChemState* __restrict__ cs,
                                                     const uint32_t* __restrict__ targets,
                                                     const uint32_t* __restrict__ tdraws,
                                                     int keys) {
#pragma clang fp contract(off)
  if (halted(d)) return;
  const uint32_t nt = cs->nt, n0 = cs->n0, na0 = cs->namp0;
  const bool stop_run = cd->in_run && (cs->stop || cs->refused);
  const uint32_t gsz = gridDim.x * blockDim.x;
  const uint32_t g0 = blockIdx.x * blockDim.x + threadIdx.x;
  if (stop_run && keys)
    for (uint32_t k = g0; k < d.ncell; k += gsz) d.cell_cnt[k] = 0u;
  const Replica& R = d.reps[0];
  const uint32_t cap = d.cap;
  const uint32_t time = *d.d_time;
  for (uint32_t g = g0; g < 4u * nt; g += gsz) {
    const uint32_t t = g >> 2, u = g & 3u;
    const uint32_t j = targets[t];
    const double4 hp = d.chainpos[4 * (size_t)j], tp = d.chainpos[4 * (size_t)j + 2];
    const double hd[3] = {hp.x, hp.y, hp.z};
    double h2e[3] = {tp.x - hp.x, tp.y - hp.y, tp.z - hp.z};
    bd_min_image(R, h2e);
    double l = sqrt(h2e[0] * h2e[0] + h2e[1] * h2e[1] + h2e[2] * h2e[2]);
    for (int a = 0; a < 3; ++a) h2e[a] /= l;
    const uint32_t new_amp = na0 + t;
    // reference-MT mode: target t's outputs are tdraws[25 t ...]: the angle,
    // then per bead u and axis a the log's and the cosine's draws
    const uint32_t* td = tdraws ? tdraws + (size_t)kMtDrawsPerTarget * t : nullptr;
    double theta;
    if (td) {
      theta = 2.0 * M_PI * mt_real1(td[0]);
    } else {
      Saru st = saru_seed(new_amp, time, R.chem_seed ^ TAG_TANG);
      theta = 2.0 * M_PI * saru_uni_f64(st);
    }
    double base[3] = {cos(theta), sin(theta), 0.0};
    const double bd = base[0] * h2e[0] + base[1] * h2e[1] + base[2] * h2e[2];
    double tv[3];
    for (int a = 0; a < 3; ++a) tv[a] = base[a] - h2e[a] * bd;
    l = sqrt(tv[0] * tv[0] + tv[1] * tv[1] + tv[2] * tv[2]);
    for (int a = 0; a < 3; ++a) tv[a] /= l;
    const uint32_t i = n0 + g;
    double np[3], vv[3];
    for (int a = 0; a < 3; ++a) {
      tv[a] *= cd->eps;
      h2e[a] *= cd->bond_leng;
      double x = (tv[a] + hd[a]) + h2e[a] * (double)u;
      x -= floor(x * R.inv_box[a]) * R.box[a];
      d.pos[a * cap + i] = x;
      np[a] = x;
    }
    d.chainpos[4 * (size_t)new_amp + u] = make_double4(np[0], np[1], np[2], 0.0);
    Saru sv = saru_seed(i, time, R.chem_seed ^ TAG_VEL);
    for (int a = 0; a < 3; ++a) {
      double u1, u2;
      if (td) {
        u1 = mt_real3(td[1 + 6 * u + 2 * a]);
        u2 = mt_real3(td[2 + 6 * u + 2 * a]);
      } else {
        u1 = saru_uni_f64(sv);
        u2 = saru_uni_f64(sv);
      }
      vv[a] = sqrt(-2.0 * log(u1)) * cos(2.0 * M_PI * u2);
      d.vel[a * cap + i] = vv[a];
      d.vbuf[a * cap + i] = 0.0;
      d.acc[a * cap + i] = 0.0f;
      d.press[a * cap + i] = 0.0f;
      d.delta[a * cap + i] = 0.0;
    }
    d.id[i] = i;
    d.orig[i] = i;
    d.prop[i] = (u == 0) ? 0u : 1u;
    d.amp_id[i] = new_amp;
    d.unit[i] = u;
    d.topo[4 * (size_t)new_amp + u] = i;
    d.chain[i] = (new_amp << 2) | u;
    if (u == 0) {
      d.amp_rep[new_amp] = 0u;
      d.cf_bend[new_amp] = cd->cf_b;  // AppendCfBend, bdf_calculator.hpp:21-26
    }
    if (keys && !stop_run) {
      // the next step's drift of the new bead (acc = 0), as k_drift bins it
      double p[3];
      for (int a = 0; a < 3; ++a) p[a] = drift_axis(R, a, np[a], vv[a], 0.0, 0.0).p;
      atomicAdd(&d.cell_cnt[cell_of(R, p)], 1u);
      d.key[i] = cell_of(R, p);
    }
  }
}

__global__ void k_chem_commit(DevState d, const ChemDev* __restrict__ cd,
                              ChemState* __restrict__ cs, uint32_t* __restrict__ amp_num) {
  if (halted(d)) return;
  const uint32_t nt = cs->nt;
  if (threadIdx.x == 0) {
    if (nt) {
      d.rep_off[1] = cs->n0 + 4u * nt;
      d.amp_off[1] = cs->namp0 + nt;
      amp_num[0] += nt;
    }
    cs->events += 1u;
    cs->total += nt;
    if (cs->refused) atomicOr(d.flags, FLAG_CHEM_CAPACITY);
    // main.cpp:188: the run ends after the step whose event reached
    // max_amp_num
    if (cd->in_run && d.run_ok && (cs->stop || cs->refused)) *d.run_ok = 0u;
  }
  for (uint32_t c = threadIdx.x; c < cd->n_core; c += blockDim.x) cs->core_slot[c] = 0xffffffffu;
}

void launch_chem_event(const DevState& d, const ChemDev* cd, ChemState* cs, uint32_t n_core,
                       uint32_t* counts, unsigned long long* masks, uint32_t* targets,
                       uint32_t* amp_num, bool keys, const ChemMt& mt, hipStream_t st) {
  const uint32_t amp_cap = d.cap / 4u + 1u;
  const uint32_t nblk = div_up(amp_cap, (uint32_t)kChemBlock);
  const bool mtm = mt.g != nullptr;
  const ChemState* ccs = cs;
  if (n_core) {
    hipLaunchKernelGGL(k_chem_find, grid1(d.nl, 256), dim3(256), 0, st, d, cd, cs);
    hipLaunchKernelGGL(k_chem_core, dim3(n_core), dim3(kCoreBlock), 0, st, d, cd, cs);
    if (mtm) {
      // at most one draw per (core, chain): n_core amp_cap bits
      launch_fill_u32(mt.cone_bits, (uint32_t)(((uint64_t)n_core * amp_cap + 31u) / 32u) + 2u, 0u,
                      st);
      hipLaunchKernelGGL(k_chem_flags<1>, dim3(nblk), dim3(kChemBlock), 0, st, d, cd, ccs,
                         mt.cone_offs, masks);
      hipLaunchKernelGGL(k_chem_scan, dim3(1), dim3(1024), 0, st, d, cd, cs, mt.cone_offs,
                         n_core * nblk, 1);
      hipLaunchKernelGGL(k_chem_mt, dim3(1), dim3(kMtBlock), 0, st, d, cd, ccs, mt, 0);
      hipLaunchKernelGGL(k_chem_decide, dim3(nblk, n_core), dim3(kChemBlock), 0, st, d, cd, mt,
                         counts, masks);
    } else {
      hipLaunchKernelGGL(k_chem_flags<0>, dim3(nblk), dim3(kChemBlock), 0, st, d, cd, ccs, counts,
                         masks);
    }
  } else {
    hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(256), 0, st, counts, 1u, 0u);
  }
  const uint32_t m = n_core ? n_core * nblk : 1u;
  hipLaunchKernelGGL(k_chem_scan, dim3(1), dim3(1024), 0, st, d, cd, cs, counts, m, 0);
  if (n_core)
    hipLaunchKernelGGL(k_chem_scatter, dim3(nblk, n_core), dim3(kChemBlock), 0, st, d, ccs,
                       (const unsigned long long*)masks, (const uint32_t*)counts, targets);
  if (mtm) hipLaunchKernelGGL(k_chem_mt, dim3(1), dim3(kMtBlock), 0, st, d, cd, ccs, mt, 1);
  hipLaunchKernelGGL(k_chem_insert, dim3(256), dim3(256), 0, st, d, cd, ccs,
                     (const uint32_t*)targets, (const uint32_t*)(mtm ? mt.tdraws : nullptr),
                     keys ? 1 : 0);
  hipLaunchKernelGGL(k_chem_commit, dim3(1), dim3(256), 0, st, d, cd, cs, amp_num);
}
uint32_t chem_count_blocks(uint32_t cap) { return div_up(cap / 4u + 1u, (uint32_t)kChemBlock); }
size_t chem_mask_words(uint32_t cap) {
  return (size_t)kMaxCore * chem_count_blocks(cap) * (kChemBlock / 64);
}

__global__ void k_copy_d(double* __restrict__ dst, const double* __restrict__ src, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
}
void launch_copy_d(double* dst, const double* src, size_t n, hipStream_t st) {
  if (!n) return;
  hipLaunchKernelGGL(k_copy_d, dim3(1024), dim3(256), 0, st, dst, src, n);
}

// The 17 columns of FPDPD::writeAscii (ptcl_class.hpp:67-73) of the bead in
// storage slot i, at row orig[i] (its upload slot) of out.
__global__ void __launch_bounds__(256) k_xyz_gather(DevState d, const double* __restrict__ vel,
                                                    XyzRec* __restrict__ out) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= dev_n(d)) return;
  XyzRec r;
  const size_t cap = d.cap;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    r.pos[a] = d.pos[a * cap + i];
    r.vel[a] = vel[a * cap + i];
    r.vbuf[a] = d.vbuf[a * cap + i];
    r.acc[a] = d.acc[a * cap + i];
  }
  r.id = d.id[i];
  r.prop = d.prop[i];
  r.amp_id = d.amp_id[i];
  r.unit = d.unit[i];
  out[d.orig[i]] = r;
}
void launch_xyz_gather(const DevState& d, const double* vel, XyzRec* out, hipStream_t st) {
  if (!d.nl) return;
  hipLaunchKernelGGL(k_xyz_gather, grid1(d.nl, 256), dim3(256), 0, st, d, vel, out);
}

}  // namespace mdpd
