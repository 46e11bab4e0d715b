// Host side of the engine: context, device memory, step orchestration,
// hipGraph capture and the C ABI of include/mdpd_capi.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "ctx.hpp"

using namespace mdpd;

namespace mdpd {
thread_local std::string g_err;
const char* kPhaseNames[kPhases] = {"drift", "sort", "gather", "density", "force",
                                    "bonded_kick", "kick", "cm_drift", "reactions", "clock"};

int fail(mdpd_ctx* c, int code, const std::string& msg) {
  if (c) c->err = msg;
  else g_err = msg;
  return code;
}

void invalidate_graphs(mdpd_ctx* c) {
  for (auto& row : c->gexec)
    for (auto& g : row)
      if (g) {
        hipGraphExecDestroy(g);
        g = nullptr;
      }
}

int enter_grow(mdpd_ctx* c) {
  if (c->n_rep != 1 || !c->chem_dev)
    return fail(c, MDPD_ERR_UNSUPPORTED, "chemistry needs a single-system context");
  if (!c->grow) {
    c->grow = true;
    c->d.run_ok = c->run_ok_word;
    set_launch_bound(c);
    invalidate_graphs(c);
  }
  return MDPD_OK;
}

int pull_counts(mdpd_ctx* c) {
  uint32_t v[3];
  HIPCHK(c, hipMemcpyAsync(&v[0], c->d.rep_off + 1, 4, hipMemcpyDeviceToHost, c->st));
  HIPCHK(c, hipMemcpyAsync(&v[1], c->d.amp_off + 1, 4, hipMemcpyDeviceToHost, c->st));
  HIPCHK(c, hipMemcpyAsync(&v[2], c->d_amp_num, 4, hipMemcpyDeviceToHost, c->st));
  HIPCHK(c, hipStreamSynchronize(c->st));
  c->d.n = v[0];
  c->rep_off[1] = v[0];
  c->R[0].n_beads = v[0];
  c->d.n_amp = v[1];
  c->n_amp = v[1];
  c->amp_off[1] = v[1];
  c->amp_num_rep[0] = v[2];
  set_launch_bound(c);
  return MDPD_OK;
}

// the permutation wrote the other buffer set: make it the current one
void swap_master(mdpd_ctx* c) {
  Master cur = static_cast<Master&>(c->d);
  static_cast<Master&>(c->d) = c->d.alt;
  c->d.alt = cur;
  c->parity ^= 1;
}
}  // namespace mdpd

namespace {

template <class T>
int dalloc(mdpd_ctx* c, T** p, size_t count) {
  void* q = nullptr;
  const size_t bytes = std::max<size_t>(count, 1) * sizeof(T);
  HIPCHK(c, hipMalloc(&q, bytes));
  HIPCHK(c, hipMemsetAsync(q, 0, bytes, c->st));
  c->allocs.push_back(q);
  *p = (T*)q;
  return MDPD_OK;
}

int ensure_pinned(mdpd_ctx* c, size_t bytes) {
  if (bytes <= c->pinned_bytes) return MDPD_OK;
  if (c->pinned) hipHostFree(c->pinned);
  c->pinned = nullptr;
  c->pinned_bytes = 0;
  HIPCHK(c, hipHostMalloc(&c->pinned, bytes, hipHostMallocDefault));
  c->pinned_bytes = bytes;
  return MDPD_OK;
}

int check_params(const mdpd_params& p, std::string& why) {
  for (int a = 0; a < 3; ++a)
    if (!(p.box[a] >= 3.0) || !std::isfinite(p.box[a])) {
      why = "box edges must be finite and >= 3 cutoffs";
      return MDPD_ERR_ARG;
    }
  if (!(p.dt > 0.0)) { why = "dt must be > 0"; return MDPD_ERR_ARG; }
  if (p.n_species < 1 || p.n_species > 3) { why = "n_species must be 1..3"; return MDPD_ERR_ARG; }
  if (p.n_dens != 1 && p.n_dens != 3) { why = "n_dens must be 1 or 3"; return MDPD_ERR_ARG; }
  if (p.envelope != 0 && p.envelope != 1) { why = "envelope must be 0 or 1"; return MDPD_ERR_ARG; }
  if (!(p.rd > 0.0 && p.rd <= 1.0)) { why = "rd must be in (0, 1]"; return MDPD_ERR_ARG; }
  if (!(p.rdm > 0.0 && p.rdm <= 1.0)) { why = "rdm must be in (0, 1]"; return MDPD_ERR_ARG; }
  if (p.conv_p > 0.0 && (p.conv_from < 0 || p.conv_from > 2 || p.conv_to < 0 || p.conv_to > 2)) {
    why = "conv_from/conv_to must be bead types 0..2";
    return MDPD_ERR_ARG;
  }
  return MDPD_OK;
}

// cell grid of a replica: the most cells per edge that are at least
// min_cell wide (the cutoff, or 1 + skin in skin mode)
void set_cells(Replica& r, double min_cell) {
  for (int a = 0; a < 3; ++a) {
    r.nc[a] = (int)std::floor(r.box[a] / min_cell);
    r.h[a] = r.box[a] / r.nc[a];
    r.inv_h[a] = 1.0 / r.h[a];
    r.hf[a] = (float)r.h[a];
  }
  r.n_cells = (uint32_t)r.nc[0] * r.nc[1] * r.nc[2];
}

Replica make_replica(const mdpd_params& p, uint32_t cell_off, uint32_t bead_off, uint32_t nb,
                     double min_cell) {
  Replica r{};
  for (int a = 0; a < 3; ++a) {
    r.box[a] = p.box[a];
    r.inv_box[a] = 1.0 / p.box[a];
  }
  set_cells(r, min_cell);
  r.dt = p.dt;
  r.rd2 = p.rd * p.rd;
  r.conv_p = p.conv_p;
  r.cf_s = p.cf_s;
  r.ibond = (p.bond_leng != 0.0) ? 1.0 / p.bond_leng : 0.0;
  r.cell_off = cell_off;
  r.bead_off = bead_off;
  r.n_beads = nb;
  r.pair_seed = p.pair_seed;
  r.chem_seed = p.chem_seed;
  r.conv_from = p.conv_from;
  r.conv_to = p.conv_to;
  r.n_species = p.n_species;
  r.n_dens = p.n_dens;
  r.envelope = p.envelope;
  r.rd = (float)p.rd;
  r.inv_rd = (float)(1.0 / p.rd);
  r.crho = (float)p.crho;
  r.inv_rdm = (float)(1.0 / p.rdm);
  r.arc = (float)p.arc;
  r.rd2f = (float)(p.rd * p.rd);
  r.rdm_is_one = (p.rdm == 1.0) ? 1 : 0;
  r.has_c1 = 0;
  for (int i = 0; i < 9; ++i) r.has_c1 |= (p.c1[i / 3][i % 3] != 0.0) ? 1 : 0;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      r.a0[i][j] = (float)p.a0[i][j];
      r.c1[i][j] = (float)p.c1[i][j];
      r.sigma[i][j] = (float)p.sigma[i][j];
      r.gamma[i][j] = (float)p.gamma[i][j];
      for (int k = 0; k < 3; ++k) r.mb[i][j][k] = (float)p.mb[i][j][k];
    }
  return r;
}

// the chemistry slots' state: no core head found yet, no events; the run word set
int chem_init_slots(mdpd_ctx* c) {
  static ChemState fresh;
  static bool once = false;
  if (!once) {
    memset(&fresh, 0, sizeof(fresh));
    for (uint32_t k = 0; k < kMaxCore; ++k) fresh.core_slot[k] = 0xffffffffu;
    once = true;
  }
  for (int q = 0; q < 2; ++q)
    HIPCHK(c, hipMemcpyAsync(c->chem_state + q, &fresh, sizeof(ChemState), hipMemcpyHostToDevice, c->st));
  launch_fill_u32(c->run_ok_word, 1, 1u, c->st);
  HIPCHK(c, hipStreamSynchronize(c->st));
  return MDPD_OK;
}

int create_common(const mdpd_params* p, const uint32_t* nb, int n_rep, uint32_t capacity,
                  mdpd_ctx** out) {
  if (!p || !out || n_rep < 1 || n_rep > kMaxRep) return fail(nullptr, MDPD_ERR_ARG, "bad arguments");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
    return fail(nullptr, MDPD_ERR_NO_DEVICE, "no HIP device");
  std::string why;
  for (int r = 0; r < n_rep; ++r) {
    int rc = check_params(p[r], why);
    if (rc) return fail(nullptr, rc, "replica " + std::to_string(r) + ": " + why);
    if (p[r].n_dens != p[0].n_dens || p[r].envelope != p[0].envelope)
      return fail(nullptr, MDPD_ERR_ARG, "replicas must share n_dens and envelope");
  }
  mdpd_ctx* c = new mdpd_ctx();
  hipGetDevice(&c->dev);
  if (int rc = noalloc_scratch_init(c->dev)) {
    const std::string why2 = g_err;
    delete c;
    return fail(nullptr, rc, why2);
  }
  if (hipStreamCreateWithFlags(&c->st, hipStreamNonBlocking) != hipSuccess) {
    delete c;
    return fail(nullptr, MDPD_ERR_HIP, "hipStreamCreate failed");
  }
  c->n_rep = n_rep;
  c->P.assign(p, p + n_rep);
  c->nd = p[0].n_dens;
  c->env = p[0].envelope;
  c->ns = 1;
  c->cm_corr = p[0].cm_corr ? p[0].cm_corr : 100u;
  for (int r = 0; r < n_rep; ++r) {
    if (p[r].n_species > 1) c->ns = 3;
    if (p[r].conv_p > 0.0) c->any_conv = true;
  }
  c->rep_off.assign(n_rep + 1, 0);
  c->cell_off.assign(n_rep + 1, 0);
  uint32_t tot = 0, cells = 0;
  for (int r = 0; r < n_rep; ++r) {
    c->rep_off[r] = tot;
    c->cell_off[r] = cells;
    const uint32_t nbr = nb ? nb[r] : capacity;
    Replica rr = make_replica(p[r], cells, tot, nbr, 1.0);
    c->R.push_back(rr);
    tot += nbr;
    cells += rr.n_cells;
  }
  c->rep_off[n_rep] = tot;
  c->cell_off[n_rep] = cells;
  c->cap = nb ? tot : capacity;
  if (c->cap == 0 || c->cap > kIndexMask) {
    mdpd_destroy(c);
    return fail(nullptr, MDPD_ERR_ARG, "capacity must be in 1.." + std::to_string(kIndexMask));
  }
  c->ncell = cells;
  c->ncell_alloc = cells;
  c->amp_cap = c->cap / 4 + 1;
  c->amp_off.assign(n_rep + 1, 0);
  c->amp_num_rep.assign(n_rep, 0);
  c->sol_num_rep.assign(n_rep, 0);
  c->tune.block_scan = 128;
  c->tune.block_force = 128;
  c->tune.cell_div = 1;
  c->tune.nlist_width = 96;
  c->tune.reorder_every = 0;
  c->tune.use_graph = 1;
  c->tune.verlet_skin = 0.0f;
  c->tune.verlet_width = 256;

  DevState& d = c->d;
  const uint32_t cap = c->cap;
  const uint32_t cap64 = div_up(cap, 64) * 64;
  int rc = MDPD_OK;
#define ALLOC(ptr, cnt) \
  if ((rc = dalloc(c, &(ptr), (cnt))) != MDPD_OK) { g_err = c->err; mdpd_destroy(c); return rc; }
  ALLOC(d.pos, 3 * (size_t)cap);
  ALLOC(d.vel, 3 * (size_t)cap);
  ALLOC(d.vbuf, 3 * (size_t)cap);
  ALLOC(d.delta, 3 * (size_t)cap);
  ALLOC(d.acc, 3 * (size_t)cap);
  ALLOC(d.press, 3 * (size_t)cap);
  ALLOC(d.id, cap);
  ALLOC(d.prop, cap);
  ALLOC(d.amp_id, cap);
  ALLOC(d.unit, cap);
  ALLOC(d.orig, cap);
  ALLOC(d.alt.pos, 3 * (size_t)cap);
  ALLOC(d.alt.vel, 3 * (size_t)cap);
  ALLOC(d.alt.vbuf, 3 * (size_t)cap);
  ALLOC(d.alt.delta, 3 * (size_t)cap);
  ALLOC(d.alt.acc, 3 * (size_t)cap);
  ALLOC(d.alt.press, 3 * (size_t)cap);
  ALLOC(d.alt.id, cap);
  ALLOC(d.alt.prop, cap);
  ALLOC(d.alt.amp_id, cap);
  ALLOC(d.alt.unit, cap);
  ALLOC(d.alt.orig, cap);
  ALLOC(d.chain, cap);
  ALLOC(d.alt.chain, cap);
  ALLOC(d.inv_perm, cap);
  ALLOC(c->dl_load, 3 * (size_t)cap);
  ALLOC(d.key, cap);
  ALLOC(d.cell_cnt, (size_t)cells + 1);
  ALLOC(d.cell_start, (size_t)cells + 1);
  ALLOC(d.cursor, (size_t)cells + 1);
  ALLOC(d.bsum, (size_t)div_up(cells, 1024) + 1);
  ALLOC(d.perm, cap64);
  ALLOC(d.spos, cap64);
  ALLOC(d.sb, cap64);
  ALLOC(d.scell, cap64);
  ALLOC(d.sdens, (size_t)(c->nd == 1 ? 1 : 4) * cap64);  // one float, or float4 per bead
  ALLOC(d.nlist, (size_t)cap64 * c->tune.nlist_width);
  ALLOC(d.ncount, cap64);
  ALLOC(d.reps, n_rep);
  ALLOC(d.rep_off, n_rep + 1);
  ALLOC(d.rep_cell_off, n_rep + 1);
  ALLOC(d.topo, 4 * (size_t)c->amp_cap);
  ALLOC(d.chainpos, 4 * (size_t)c->amp_cap);
  ALLOC(d.cf_bend, c->amp_cap);
  ALLOC(d.amp_rep, c->amp_cap);
  ALLOC(d.amp_off, n_rep + 1);
  ALLOC(d.d_time, 1);
  ALLOC(d.flags, 1);
  ALLOC(d.rebuild, 2);
  ALLOC(d.red, (size_t)n_rep * 64 * 32 + 3 * (size_t)cap);
  ALLOC(d.red_out, (size_t)n_rep * 16);
  ALLOC(c->vir_bonded, (size_t)n_rep * 3);
  ALLOC(c->obs_out, (size_t)n_rep * 16);
  ALLOC(c->d_amp_num, n_rep);
  if (n_rep == 1) {
    ALLOC(c->chem_dev, 2);
    ALLOC(c->chem_state, 2);
    ALLOC(c->chem_counts, (size_t)kMaxCore * chem_count_blocks(cap));
    ALLOC(c->chem_targets, c->amp_cap);
    ALLOC(c->chem_masks, chem_mask_words(cap));
    ALLOC(c->run_ok_word, 1);
  }
#undef ALLOC
  d.cap = cap;
  d.n = 0;
  d.nl = 0;
  d.run_ok = nullptr;
  d.ncell = cells;
  d.W = c->tune.nlist_width;
  d.n_rep = n_rep;
  d.n_dens = c->nd;
  d.rd_le_1 = 1;
  for (int r = 0; r < n_rep; ++r)
    if (!(p[r].rd <= 1.0)) d.rd_le_1 = 0;
  d.n_amp = 0;
  if (hipMemcpyAsync(d.reps, c->R.data(), sizeof(Replica) * n_rep, hipMemcpyHostToDevice, c->st) !=
          hipSuccess ||
      hipMemcpyAsync(d.rep_off, c->rep_off.data(), sizeof(uint32_t) * (n_rep + 1),
                     hipMemcpyHostToDevice, c->st) != hipSuccess ||
      hipMemcpyAsync(d.rep_cell_off, c->cell_off.data(), sizeof(uint32_t) * (n_rep + 1),
                     hipMemcpyHostToDevice, c->st) != hipSuccess ||
      (n_rep == 1 && chem_init_slots(c) != MDPD_OK) ||
      hipStreamSynchronize(c->st) != hipSuccess) {
    mdpd_destroy(c);
    return fail(nullptr, MDPD_ERR_HIP, "initial upload failed");
  }
  *out = c;
  return MDPD_OK;
}

// one step of main.cpp:167-196 on the context stream.  first: the cell keys
// of the drifted positions come from a key pass (the first step of a run,
// and the step after a CM-drift removal, which changed the velocities);
// otherwise the previous step's force epilogue wrote them.
void enqueue_step(mdpd_ctx* c, bool virial, bool cm, bool first, bool chem) {
  const DevState& d = c->d;
  hipStream_t st = c->st;
  const bool prof = c->profiling;
  hipEvent_t* ev = nullptr;
  if (prof) {
    const size_t need = (size_t)(c->ev_steps + 1) * (kPhases + 1);
    while (c->evpool.size() < need) {
      hipEvent_t e;
      hipEventCreate(&e);
      c->evpool.push_back(e);
    }
    ev = &c->evpool[(size_t)c->ev_steps * (kPhases + 1)];
    c->ev_steps++;
  }
  auto mark = [&](int k) {
    if (prof) hipEventRecord(ev[k], st);
  };
  mark(0);
  if (c->tune.verlet_skin > 0.0f) {
    // skin mode: the drift raises the rebuild word; the cell list, the
    // storage permutation and the skin list run only when it is set
    launch_drift_skin(d, st);
    mark(1);
    launch_keys_gated(d, st);
    launch_scan_cells_gated(d, st);
    launch_scatter_gated(d, st);
    launch_cell_sort_gated(d, st);
    mark(2);
    launch_permute_gated(d, st);
    launch_copy_back_gated(d, st);
    launch_skin_build(d, st);
    mark(3);
    launch_density_skin(d, c->nd, c->tune.block_scan, st);
    mark(4);
    launch_force(d, c->ns, c->nd, c->env, virial, c->tune.block_force, st, true);
    mark(5);
    if (virial) launch_bead_vir_reduce(d, c->vir_bonded, st);
    mark(6);
    mark(7);
    if (cm) launch_cm_drift(d, st);
    mark(8);
    // (a separate pass here: it also rewrites the FP32 type the next
    // step's pair passes read without a rebuild)
    if (c->any_conv) launch_head_conversion(d, st);
    mark(9);
    mark(10);
    return;
  }
  // the cell histogram is zero before the force pass that fills it (or this
  // key pass): allocated zeroed and cleared by every scan
  if (first) launch_step_keys(d, st);
  mark(1);
  launch_scan_cells(d, st);
  launch_scatter(d, st);
  mark(2);
  // bead storage moves to the new cell order (each cell's beads by previous
  // slot, ranked inside the permutation); the slots below are the new ones
  launch_step_permute(d, st);
  swap_master(c);
  mark(3);
  launch_density(d, c->ns, c->nd, c->tune.block_scan, st);
  mark(4);
  // the force pass ends with ForceBonded::CalcListedForce and the kick of
  // each bead (bead-centric: every bead evaluates its own chain)
  // ... and, with any reactive replica, the head conversion of each bead
  // ... and the next step's cell keys, unless this is a run's last step (the
  // host may change the state before the next run) or a CM step
  launch_force(d, c->ns, c->nd, c->env, virial, c->tune.block_force, st, true, c->any_conv,
               !virial && !cm);
  mark(5);
  if (virial) launch_bead_vir_reduce(d, c->vir_bonded, st);
  mark(6);
  mark(7);
  // the CM drift removal only changes velocities, so the conversion
  // (main.cpp order: after it) may run inside the force pass
  if (cm) launch_cm_drift(d, st);
  mark(8);
  // RandomChemEvent (main.cpp:186-196); new beads join the next step (their
  // cell keys with the others' when this step's force pass binned them)
  if (chem)
    launch_chem_event(d, c->chem_dev, c->chem_state, c->chem_host[0].n_core, c->chem_counts,
                      c->chem_masks, c->chem_targets, c->d_amp_num, !virial && !cm, c->chem_mt, st);
  mark(9);
  mark(10);
}

}  // namespace

int mdpd::check_flags(mdpd_ctx* c) {
  uint32_t f = 0;
  HIPCHK(c, hipMemcpyAsync(&f, c->d.flags, sizeof(uint32_t), hipMemcpyDeviceToHost, c->st));
  HIPCHK(c, hipStreamSynchronize(c->st));
  if (f) {
    launch_fill_u32(c->d.flags, 1, 0u, c->st);
    hipStreamSynchronize(c->st);
    if (f & FLAG_NLIST_OVERFLOW)
      return fail(c, MDPD_ERR_NLIST_OVERFLOW,
                  "a bead has more neighbours than nlist_width = " + std::to_string(c->d.W));
    if (f & FLAG_NONFINITE) return fail(c, MDPD_ERR_NONFINITE, "non-finite position/velocity/force");
    if (f & FLAG_CHEM_CAPACITY)
      return fail(c, MDPD_ERR_CAPACITY, "inserted amphiphiles would exceed the capacity");
    return fail(c, MDPD_ERR_OUT_OF_BOX, "bead outside the box");
  }
  return MDPD_OK;
}

namespace {
// graph index of a step variant
int step_variant(bool virial, bool cm, bool first, bool chem) {
  return (virial ? 1 : 0) | (cm ? 2 : 0) | (first ? 4 : 0) | (chem ? 8 : 0);
}

// records the graph of step variant v for master-set parity p without
// running it; the host's parity is left as it was
int capture_variant(mdpd_ctx* c, int p, int v) {
  const int p0 = c->parity;
  if (c->parity != p) swap_master(c);
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamBeginCapture(c->st, hipStreamCaptureModeThreadLocal);
  if (e == hipSuccess) {
    enqueue_step(c, (v & 1) != 0, (v & 2) != 0, (v & 4) != 0, (v & 8) != 0);  // flips the parity
    e = hipStreamEndCapture(c->st, &g);
  }
  if (e == hipSuccess) e = hipGraphInstantiate(&c->gexec[p][v], g, nullptr, nullptr, 0);
  if (g) hipGraphDestroy(g);
  // the host's view of the master sets is the one it had, on every path
  if (c->parity != p0) swap_master(c);
  if (e != hipSuccess) {
    c->gexec[p][v] = nullptr;
    return fail(c, MDPD_ERR_HIP, std::string("step graph capture: ") + hipGetErrorString(e));
  }
  return MDPD_OK;
}
}  // namespace

// Captures, before anything is launched, the graph of every step variant
// (parity of the master set x virial / CM / key-pass / chemistry step) that
// the next nsteps steps of run_steps use, so that no capture falls between
// the launches of a run.  Skin mode and eager or profiled runs capture
// nothing.
int mdpd::prepare_graphs(mdpd_ctx* c, uint32_t nsteps) {
  if (!c->tune.use_graph || c->profiling || c->tune.verlet_skin > 0.0f || nsteps == 0)
    return MDPD_OK;
  // step k of the run: time c->time + k, parity c->parity ^ (k & 1), a key
  // pass when k = 0 or step k - 1 was a CM step, the virial when k is the
  // last.  Past the chemistry's start and two CM periods the kinds repeat
  // (period lcm(2, cm_corr)), so those steps and the last cover every
  // variant of the run.
  const uint32_t cmc = (uint32_t)c->cm_corr;
  uint32_t span = 2u * cmc + 2u;
  if (c->chem_on && c->beg_chem > c->time) span += c->beg_chem - c->time;
  auto visit = [&](uint32_t k) -> int {
    const uint32_t t = c->time + k;
    const int p = c->parity ^ (int)(k & 1u);
    const bool first = k == 0u || ((t - 1u) % cmc) == 0u;
    const int v = step_variant(k + 1u == nsteps, (t % cmc) == 0u, first,
                               c->chem_on && t >= c->beg_chem);
    return c->gexec[p][v] ? MDPD_OK : capture_variant(c, p, v);
  };
  for (uint32_t k = 0; k < nsteps && k <= span; ++k) {
    const int rc = visit(k);
    if (rc) return rc;
  }
  if (nsteps - 1u > span) return visit(nsteps - 1u);
  return MDPD_OK;
}

int mdpd::run_steps(mdpd_ctx* c, uint32_t nsteps) {
  const bool skin = c->tune.verlet_skin > 0.0f;
  if (c->chem_on && skin)
    return fail(c, MDPD_ERR_UNSUPPORTED,
                "chemistry inside mdpd_run needs the every-step cell list (verlet_skin = 0)");
  {
    const int rc = prepare_graphs(c, nsteps);
    if (rc) return rc;
  }
  const uint32_t time0 = c->time;
  // skin mode: the first step builds when the host has moved or reordered
  // beads since the last build
  if (skin && nsteps && !c->skin_valid) launch_fill_u32(c->d.rebuild, 1, 1u, c->st);
  // the device clock holds the previous step's time; each step's first
  // kernel advances it
  if (nsteps) launch_fill_u32(c->d.d_time, 1, c->time - 1u, c->st);
  bool first = true;
  for (uint32_t k = 0; k < nsteps; ++k) {
    // the last step of a run is its virial step
    const bool virial = (k + 1 == nsteps);
    const bool cm = (c->time % c->cm_corr) == 0;
    const bool chem = c->chem_on && c->time >= c->beg_chem;
    if (c->tune.use_graph && !c->profiling) {
      const int v = step_variant(virial, cm, first && !skin, chem);
      // (a skin-mode step keeps the storage where it is: no parity flip)
      const int p = c->parity;
      if (!c->gexec[p][v]) {
        // capturing records the step and flips the host's buffer parity, as
        // running it does
        hipGraph_t g = nullptr;
        HIPCHK(c, hipStreamBeginCapture(c->st, hipStreamCaptureModeThreadLocal));
        enqueue_step(c, virial, cm, first, chem);
        HIPCHK(c, hipStreamEndCapture(c->st, &g));
        HIPCHK(c, hipGraphInstantiate(&c->gexec[p][v], g, nullptr, nullptr, 0));
        hipGraphDestroy(g);
      } else if (!skin) {
        swap_master(c);
      }
      HIPCHK(c, hipGraphLaunch(c->gexec[p][v], c->st));
    } else {
      enqueue_step(c, virial, cm, first, chem);
    }
    first = cm;
    c->time++;
  }
  if (c->chem_on && nsteps) {
    // the chemistry changed the counts on the device; a stop skipped the
    // run's remaining steps
    uint32_t w[2] = {1u, 0u};
    HIPCHK(c, hipMemcpyAsync(&w[0], c->run_ok_word, 4, hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipMemcpyAsync(&w[1], c->d.d_time, 4, hipMemcpyDeviceToHost, c->st));
    int rc = pull_counts(c);
    if (rc) return rc;
    if (w[0] == 0u) {
      const uint32_t done = w[1] - (time0 - 1u);  // steps whose permutation ran
      if ((nsteps - done) & 1u) swap_master(c);  // undo the skipped steps' parity flips
      c->time = w[1];  // main.cpp:188 leaves the loop before time++
      c->chem_stopped = true;
      launch_fill_u32(c->run_ok_word, 1, 1u, c->st);
      HIPCHK(c, hipStreamSynchronize(c->st));
    }
  }
  c->drifted = false;
  c->have_cells = true;
  c->cells_fresh = true;
  c->have_density = true;
  // beads sit up to skin / 2 away from where the last build put them
  if (skin && nsteps) {
    c->d.cell_slack = c->tune.verlet_skin;
    c->skin_valid = true;
  }
  return MDPD_OK;
}


// ======================================================================
// C ABI
// ======================================================================
extern "C" {

const char* mdpd_global_error(void) { return g_err.c_str(); }
const char* mdpd_last_error(const mdpd_ctx* c) { return c ? c->err.c_str() : g_err.c_str(); }

int mdpd_device_name(char* buf, size_t len) {
  hipDeviceProp_t pr;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&pr, dev) != hipSuccess)
    return fail(nullptr, MDPD_ERR_NO_DEVICE, "no HIP device");
  snprintf(buf, len, "%s (%s, %d CUs)", pr.name, pr.gcnArchName, pr.multiProcessorCount);
  return MDPD_OK;
}

int mdpd_create(const mdpd_params* p, uint32_t capacity, mdpd_ctx** out) {
  return create_common(p, nullptr, 1, capacity, out);
}

int mdpd_create_ensemble(const mdpd_params* p, const uint32_t* n_beads, int32_t n_rep,
                         mdpd_ctx** out) {
  if (!n_beads) return fail(nullptr, MDPD_ERR_ARG, "n_beads is null");
  return create_common(p, n_beads, n_rep, 0, out);
}

int mdpd_destroy(mdpd_ctx* c) {
  if (!c) return MDPD_OK;
  xyz_writer_free(c->xyz);  // finishes a trajectory write in flight
  if (c->st) hipStreamSynchronize(c->st);
  invalidate_graphs(c);
  for (auto& e : c->evpool) hipEventDestroy(e);
  for (void* q : c->allocs) hipFree(q);
  if (c->pinned) hipHostFree(c->pinned);
  if (c->st) hipStreamDestroy(c->st);
  delete c;
  return MDPD_OK;
}

int mdpd_num_beads(const mdpd_ctx* c, uint32_t* n) {
  if (!c || !n) return MDPD_ERR_ARG;
  *n = c->d.n;
  return MDPD_OK;
}

int mdpd_upload_soa(mdpd_ctx* c, uint32_t n, const double* pos, const double* vel,
                    const double* acc, const uint32_t* id, const uint32_t* prop,
                    const uint32_t* amp_id, const uint32_t* unit) {
  if (c) c->skin_valid = false;
  if (!c || !pos || !vel || !id || !prop) return fail(c, MDPD_ERR_ARG, "null argument");
  if (n > c->cap) return fail(c, MDPD_ERR_CAPACITY, "more beads than the capacity");
  if (c->n_rep > 1 && n != c->cap)
    return fail(c, MDPD_ERR_ARG, "an ensemble upload must hold every replica's beads");
  // parameter.hpp:436-460 checks
  for (uint32_t i = 0; i < n; ++i) {
    const int r = (int)(std::upper_bound(c->rep_off.begin(), c->rep_off.begin() + c->n_rep, i) -
                        c->rep_off.begin()) - 1;
    for (int a = 0; a < 3; ++a) {
      const double x = pos[3 * (size_t)i + a];
      if (!std::isfinite(x) || !std::isfinite(vel[3 * (size_t)i + a]))
        return fail(c, MDPD_ERR_NONFINITE, "non-finite position or velocity at bead " + std::to_string(i));
      if (!(x <= c->R[r].box[a] && x >= 0.0))
        return fail(c, MDPD_ERR_OUT_OF_BOX, "There is a particle in outside range (bead " + std::to_string(i) + ")");
    }
    if (prop[i] > 2) return fail(c, MDPD_ERR_ARG, "bead type must be 0..2");
  }
  const size_t bytes = 3 * sizeof(double) * (size_t)n * 4 + 4 * sizeof(uint32_t) * (size_t)n +
                       3 * sizeof(float) * (size_t)n;
  int rc = ensure_pinned(c, bytes + 64);
  if (rc) return rc;
  double* hp = (double*)c->pinned;
  double* hv = hp + 3 * (size_t)n;
  double* hb = hv + 3 * (size_t)n;
  double* hd = hb + 3 * (size_t)n;
  float* ha = (float*)(hd + 3 * (size_t)n);
  uint32_t* hu = (uint32_t*)(ha + 3 * (size_t)n);
  for (uint32_t i = 0; i < n; ++i)
    for (int a = 0; a < 3; ++a) {
      hp[(size_t)a * n + i] = pos[3 * (size_t)i + a];
      hv[(size_t)a * n + i] = vel[3 * (size_t)i + a];
      hb[(size_t)a * n + i] = vel[3 * (size_t)i + a];
      hd[(size_t)a * n + i] = 0.0;
      ha[(size_t)a * n + i] = acc ? (float)acc[3 * (size_t)i + a] : 0.f;
    }
  for (uint32_t i = 0; i < n; ++i) {
    hu[i] = id[i];
    hu[n + i] = prop[i];
    hu[2 * (size_t)n + i] = amp_id ? amp_id[i] : MDPD_UNIT_NONE;
    hu[3 * (size_t)n + i] = unit ? unit[i] : MDPD_UNIT_NONE;
  }
  DevState& d = c->d;
  for (int a = 0; a < 3; ++a) {
    HIPCHK(c, hipMemcpyAsync(d.pos + (size_t)a * d.cap, hp + (size_t)a * n, 8 * (size_t)n, hipMemcpyHostToDevice, c->st));
    HIPCHK(c, hipMemcpyAsync(d.vel + (size_t)a * d.cap, hv + (size_t)a * n, 8 * (size_t)n, hipMemcpyHostToDevice, c->st));
    HIPCHK(c, hipMemcpyAsync(d.vbuf + (size_t)a * d.cap, hb + (size_t)a * n, 8 * (size_t)n, hipMemcpyHostToDevice, c->st));
    HIPCHK(c, hipMemcpyAsync(d.delta + (size_t)a * d.cap, hd + (size_t)a * n, 8 * (size_t)n, hipMemcpyHostToDevice, c->st));
    HIPCHK(c, hipMemcpyAsync(d.acc + (size_t)a * d.cap, ha + (size_t)a * n, 4 * (size_t)n, hipMemcpyHostToDevice, c->st));
  }
  HIPCHK(c, hipMemcpyAsync(d.id, hu, 4 * (size_t)n, hipMemcpyHostToDevice, c->st));
  HIPCHK(c, hipMemcpyAsync(d.prop, hu + n, 4 * (size_t)n, hipMemcpyHostToDevice, c->st));
  HIPCHK(c, hipMemcpyAsync(d.amp_id, hu + 2 * (size_t)n, 4 * (size_t)n, hipMemcpyHostToDevice, c->st));
  HIPCHK(c, hipMemcpyAsync(d.unit, hu + 3 * (size_t)n, 4 * (size_t)n, hipMemcpyHostToDevice, c->st));
  HIPCHK(c, hipStreamSynchronize(c->st));
  for (uint32_t i = 0; i < n; ++i) hu[i] = i;  // storage slot i holds upload slot i
  HIPCHK(c, hipMemcpyAsync(d.orig, hu, 4 * (size_t)n, hipMemcpyHostToDevice, c->st));
  launch_fill_u32(d.chain, n, 0xffffffffu, c->st);  // chains come with set_topology
  HIPCHK(c, hipStreamSynchronize(c->st));
  d.n_amp = 0;
  c->n_amp = 0;
  // solvent count per replica (unit == none)
  for (int r = 0; r < c->n_rep; ++r) {
    uint32_t s = 0;
    for (uint32_t i = c->rep_off[r]; i < std::min(n, c->rep_off[r + 1]); ++i)
      if (!unit || unit[i] == MDPD_UNIT_NONE) ++s;
    c->sol_num_rep[r] = s;
  }
  if (n != d.n) invalidate_graphs(c);
  d.n = n;
  set_launch_bound(c);
  if (c->n_rep == 1) {
    c->rep_off[1] = n;
    c->R[0].n_beads = n;
    // no chains until set_topology (the device's copies are the live counts
    // in grow mode)
    c->amp_off.assign(2, 0);
    c->amp_num_rep[0] = 0;
    HIPCHK(c, hipMemcpyAsync(d.rep_off, c->rep_off.data(), sizeof(uint32_t) * 2, hipMemcpyHostToDevice, c->st));
    HIPCHK(c, hipMemcpyAsync(d.amp_off, c->amp_off.data(), sizeof(uint32_t) * 2, hipMemcpyHostToDevice, c->st));
    HIPCHK(c, hipMemcpyAsync(c->d_amp_num, c->amp_num_rep.data(), sizeof(uint32_t), hipMemcpyHostToDevice, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
  }
  c->drifted = false;
  c->have_cells = c->cells_fresh = c->have_density = false;
  return MDPD_OK;
}

int mdpd_upload_fpdpd(mdpd_ctx* c, const void* aos, uint32_t n, size_t stride) {
  if (!c || !aos) return fail(c, MDPD_ERR_ARG, "null argument");
  if (stride < MDPD_FPDPD_STRIDE) return fail(c, MDPD_ERR_ARG, "stride < 232");
  std::vector<double> pos(3 * (size_t)n), vel(3 * (size_t)n), vb(3 * (size_t)n), acc(3 * (size_t)n),
      del(3 * (size_t)n);
  std::vector<uint32_t> id(n), prop(n), aid(n), unit(n);
  const char* b = (const char*)aos;
  for (uint32_t i = 0; i < n; ++i) {
    const char* r = b + (size_t)i * stride;
    memcpy(&id[i], r + 0, 4);
    memcpy(&prop[i], r + 4, 4);
    memcpy(&aid[i], r + 8, 4);
    memcpy(&unit[i], r + 12, 4);
    memcpy(&pos[3 * (size_t)i], r + 16, 24);
    memcpy(&del[3 * (size_t)i], r + 40, 24);
    memcpy(&vel[3 * (size_t)i], r + 64, 24);
    memcpy(&vb[3 * (size_t)i], r + 88, 24);
    memcpy(&acc[3 * (size_t)i], r + 112, 24);
  }
  int rc = mdpd_upload_soa(c, n, pos.data(), vel.data(), acc.data(), id.data(), prop.data(),
                           aid.data(), unit.data());
  if (rc) return rc;
  // vel_buf and delta_sumr of the records
  DevState& d = c->d;
  std::vector<double> t(n);
  for (int a = 0; a < 3; ++a) {
    for (uint32_t i = 0; i < n; ++i) t[i] = vb[3 * (size_t)i + a];
    HIPCHK(c, hipMemcpy(d.vbuf + (size_t)a * d.cap, t.data(), 8 * (size_t)n, hipMemcpyHostToDevice));
    for (uint32_t i = 0; i < n; ++i) t[i] = del[3 * (size_t)i + a];
    HIPCHK(c, hipMemcpy(d.delta + (size_t)a * d.cap, t.data(), 8 * (size_t)n, hipMemcpyHostToDevice));
  }
  return MDPD_OK;
}

int mdpd_download_soa(mdpd_ctx* c, uint32_t n, double* pos, double* vel, double* vel_buf,
                      double* acc, double* press, double* dens, double* delta, uint32_t* id,
                      uint32_t* prop, uint32_t* amp_id, uint32_t* unit) {
  if (!c) return MDPD_ERR_ARG;
  if (n != c->d.n) return fail(c, MDPD_ERR_ARG, "n must equal the bead count");
  HIPCHK(c, hipStreamSynchronize(c->st));
  const DevState& d = c->d;
  // storage is in cell order; orig[] puts every bead back at its upload slot
  std::vector<uint32_t> orig(n);
  HIPCHK(c, hipMemcpy(orig.data(), d.orig, 4 * (size_t)n, hipMemcpyDeviceToHost));
  for (uint32_t i = 0; i < n; ++i)
    if (orig[i] >= n) return fail(c, MDPD_ERR_STATE, "bead storage map is corrupt");
  std::vector<double> t(n);
  std::vector<float> f(n);
  std::vector<uint32_t> u(n);
  auto get3d = [&](const double* src, double* dst) -> int {
    if (!dst) return MDPD_OK;
    for (int a = 0; a < 3; ++a) {
      HIPCHK(c, hipMemcpy(t.data(), src + (size_t)a * d.cap, 8 * (size_t)n, hipMemcpyDeviceToHost));
      for (uint32_t i = 0; i < n; ++i) dst[3 * (size_t)orig[i] + a] = t[i];
    }
    return MDPD_OK;
  };
  auto get3f = [&](const float* src, double* dst) -> int {
    if (!dst) return MDPD_OK;
    for (int a = 0; a < 3; ++a) {
      HIPCHK(c, hipMemcpy(f.data(), src + (size_t)a * d.cap, 4 * (size_t)n, hipMemcpyDeviceToHost));
      for (uint32_t i = 0; i < n; ++i) dst[3 * (size_t)orig[i] + a] = f[i];
    }
    return MDPD_OK;
  };
  auto getu = [&](const uint32_t* src, uint32_t* dst) -> int {
    if (!dst) return MDPD_OK;
    HIPCHK(c, hipMemcpy(u.data(), src, 4 * (size_t)n, hipMemcpyDeviceToHost));
    for (uint32_t i = 0; i < n; ++i) dst[orig[i]] = u[i];
    return MDPD_OK;
  };
  int rc;
  if ((rc = get3d(d.pos, pos))) return rc;
  if ((rc = get3d(c->drifted ? d.vbuf : d.vel, vel))) return rc;
  if ((rc = get3d(d.vbuf, vel_buf))) return rc;
  if ((rc = get3d(d.delta, delta))) return rc;
  if ((rc = get3f(d.acc, acc))) return rc;
  if ((rc = get3f(d.press, press))) return rc;
  if ((rc = getu(d.id, id))) return rc;
  if ((rc = getu(d.prop, prop))) return rc;
  if ((rc = getu(d.amp_id, amp_id))) return rc;
  if ((rc = getu(d.unit, unit))) return rc;
  if (dens) {
    memset(dens, 0, sizeof(double) * 3 * (size_t)n);
    if (c->have_density) {
      // one float per bead, or (rho_0, rho_1, rho_2, 0) per bead
      const size_t per = c->nd == 1 ? 1 : 4;
      std::vector<float> g(per * (size_t)n);
      HIPCHK(c, hipMemcpy(g.data(), d.sdens, 4 * per * (size_t)n, hipMemcpyDeviceToHost));
      for (uint32_t i = 0; i < n; ++i)
        for (int k = 0; k < c->nd; ++k) dens[3 * (size_t)orig[i] + k] = g[per * i + k];
    }
  }
  return MDPD_OK;
}

int mdpd_download_fpdpd(mdpd_ctx* c, void* aos, uint32_t n, size_t stride) {
  if (!c || !aos) return fail(c, MDPD_ERR_ARG, "null argument");
  if (stride < MDPD_FPDPD_STRIDE) return fail(c, MDPD_ERR_ARG, "stride < 232");
  std::vector<double> pos(3 * (size_t)n), vel(3 * (size_t)n), vb(3 * (size_t)n), acc(3 * (size_t)n),
      pr(3 * (size_t)n), dn(3 * (size_t)n), del(3 * (size_t)n), cm(8 * (size_t)n, 0.0);
  std::vector<uint32_t> id(n), prop(n), aid(n), unit(n);
  int rc = mdpd_download_soa(c, n, pos.data(), vel.data(), vb.data(), acc.data(), pr.data(),
                             dn.data(), del.data(), id.data(), prop.data(), aid.data(), unit.data());
  if (rc) return rc;
  if (c->cells_fresh && n) {
    // nei_cm_pos (ptcl_class.hpp:53-57) from the positions of the last rebuild
    uint32_t* idx = nullptr;
    double* out = nullptr;
    HIPCHK(c, hipMalloc(&idx, 4 * (size_t)n));
    HIPCHK(c, hipMalloc(&out, 64 * (size_t)n));
    // row k of the output is the bead uploaded at slot k
    std::vector<uint32_t> orig(n), seq(n);
    HIPCHK(c, hipMemcpy(orig.data(), c->d.orig, 4 * (size_t)n, hipMemcpyDeviceToHost));
    for (uint32_t i = 0; i < n; ++i) seq[orig[i]] = i;
    HIPCHK(c, hipMemcpy(idx, seq.data(), 4 * (size_t)n, hipMemcpyHostToDevice));
    launch_neighbor_cm(c->d, idx, n, out, c->st);
    HIPCHK(c, hipStreamSynchronize(c->st));
    HIPCHK(c, hipMemcpy(cm.data(), out, 64 * (size_t)n, hipMemcpyDeviceToHost));
    hipFree(idx);
    hipFree(out);
  }
  char* b = (char*)aos;
  for (uint32_t i = 0; i < n; ++i) {
    char* r = b + (size_t)i * stride;
    memcpy(r + 0, &id[i], 4);
    memcpy(r + 4, &prop[i], 4);
    memcpy(r + 8, &aid[i], 4);
    memcpy(r + 12, &unit[i], 4);
    memcpy(r + 16, &pos[3 * (size_t)i], 24);
    memcpy(r + 40, &del[3 * (size_t)i], 24);
    memcpy(r + 64, &vel[3 * (size_t)i], 24);
    memcpy(r + 88, &vb[3 * (size_t)i], 24);
    memcpy(r + 112, &acc[3 * (size_t)i], 24);
    memcpy(r + 136, &pr[3 * (size_t)i], 24);
    memcpy(r + 160, &cm[8 * (size_t)i], 24);
    memcpy(r + 184, &cm[8 * (size_t)i + 3], 24);
    memcpy(r + 208, &dn[3 * (size_t)i], 24);
  }
  return MDPD_OK;
}

int mdpd_set_topology(mdpd_ctx* c, const uint32_t* topo, uint32_t n_amp, const uint32_t* core_amp_id,
                      uint32_t n_core) {
  if (c) c->skin_valid = false;
  if (!c || (n_amp && !topo)) return fail(c, MDPD_ERR_ARG, "null argument");
  if (n_amp > c->amp_cap) return fail(c, MDPD_ERR_CAPACITY, "too many amphiphiles");
  std::vector<double> cfb(n_amp);
  std::vector<uint32_t> arep(n_amp);
  std::fill(c->amp_num_rep.begin(), c->amp_num_rep.end(), 0u);
  int last = 0;
  for (uint32_t a = 0; a < n_amp; ++a) {
    for (int u = 0; u < 4; ++u)
      if (topo[4 * (size_t)a + u] >= c->d.n)
        return fail(c, MDPD_ERR_ARG, "topology index out of range (amphiphile " + std::to_string(a) + ")");
    const uint32_t b0 = topo[4 * (size_t)a];
    const int r = (int)(std::upper_bound(c->rep_off.begin(), c->rep_off.begin() + c->n_rep, b0) -
                        c->rep_off.begin()) - 1;
    if (r < last) return fail(c, MDPD_ERR_ARG, "chains must be grouped by replica");
    last = r;
    arep[a] = (uint32_t)r;
    c->amp_num_rep[r]++;
    cfb[a] = c->P[r].cf_b;  // bdf_calculator.hpp:77-86
    for (uint32_t k = 0; k < n_core; ++k)
      if (core_amp_id && a == core_amp_id[k]) cfb[a] = c->P[r].cf_b_rigid;
  }
  c->amp_off.assign(c->n_rep + 1, 0);
  for (int r = 0; r < c->n_rep; ++r) c->amp_off[r + 1] = c->amp_off[r] + c->amp_num_rep[r];
  DevState& d = c->d;
  if (n_amp) {
    // the topology names upload slots; storage is in cell order
    std::vector<uint32_t> orig(d.n), slot(d.n), tp(4 * (size_t)n_amp);
    HIPCHK(c, hipMemcpy(orig.data(), d.orig, 4 * (size_t)d.n, hipMemcpyDeviceToHost));
    for (uint32_t i = 0; i < d.n; ++i) slot[orig[i]] = i;
    for (size_t k = 0; k < tp.size(); ++k) tp[k] = slot[topo[k]];
    HIPCHK(c, hipMemcpy(d.topo, tp.data(), 16 * (size_t)n_amp, hipMemcpyHostToDevice));
    // each bead's chain and position in it (glob_topol_[amp * 4 + unit])
    std::vector<uint32_t> ch(d.n, 0xffffffffu);
    for (size_t k = 0; k < tp.size(); ++k) ch[tp[k]] = (uint32_t)((k / 4) << 2 | (k & 3));
    HIPCHK(c, hipMemcpy(d.chain, ch.data(), 4 * (size_t)d.n, hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(d.cf_bend, cfb.data(), 8 * (size_t)n_amp, hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(d.amp_rep, arep.data(), 4 * (size_t)n_amp, hipMemcpyHostToDevice));
  }
  HIPCHK(c, hipMemcpy(d.amp_off, c->amp_off.data(), 4 * (size_t)(c->n_rep + 1), hipMemcpyHostToDevice));
  HIPCHK(c, hipMemcpy(c->d_amp_num, c->amp_num_rep.data(), 4 * (size_t)c->n_rep, hipMemcpyHostToDevice));
  if (n_amp != d.n_amp) invalidate_graphs(c);
  d.n_amp = n_amp;
  c->n_amp = n_amp;
  return MDPD_OK;
}

int mdpd_set_clock(mdpd_ctx* c, uint32_t time, uint32_t pair_seed) {
  if (!c) return MDPD_ERR_ARG;
  c->time = time;
  if (c->n_rep == 1) {  // an ensemble keeps the seeds of its replicas
    c->P[0].pair_seed = pair_seed;
    c->R[0].pair_seed = pair_seed;
  }
  HIPCHK(c, hipMemcpyAsync(c->d.reps, c->R.data(), sizeof(Replica) * c->n_rep, hipMemcpyHostToDevice, c->st));
  HIPCHK(c, hipMemcpyAsync(c->d.d_time, &c->time, 4, hipMemcpyHostToDevice, c->st));
  HIPCHK(c, hipStreamSynchronize(c->st));
  return MDPD_OK;
}

int mdpd_get_clock(const mdpd_ctx* c, uint32_t* time) {
  if (!c || !time) return MDPD_ERR_ARG;
  *time = c->time;
  return MDPD_OK;
}

int mdpd_set_tuning(mdpd_ctx* c, const mdpd_tuning* t) {
  if (c) c->skin_valid = false;
  if (!c || !t) return MDPD_ERR_ARG;
  auto okb = [](int b) { return b == 64 || b == 128 || b == 256; };
  if (!okb(t->block_scan) || !okb(t->block_force)) return fail(c, MDPD_ERR_ARG, "block sizes must be 64, 128 or 256");
  if (t->cell_div != 1) return fail(c, MDPD_ERR_UNSUPPORTED, "cell_div must be 1");
  if (t->nlist_width < 8 || t->nlist_width > 1024 || (t->nlist_width & 3))
    return fail(c, MDPD_ERR_ARG, "nlist_width must be a multiple of 4 in 8..1024");
  const float sk = t->verlet_skin;
  const bool skin = sk > 0.0f;
  if (!(sk == 0.0f || (sk >= 0.05f && sk <= 1.0f)))
    return fail(c, MDPD_ERR_ARG, "verlet_skin must be 0 or in 0.05..1");
  if (skin && (t->verlet_width < 8 || t->verlet_width > 2048 || (t->verlet_width & 3)))
    return fail(c, MDPD_ERR_ARG, "verlet_width must be a multiple of 4 in 8..2048");
  if (skin)
    for (int r = 0; r < c->n_rep; ++r)
      for (int a = 0; a < 3; ++a)
        if (std::floor(c->R[r].box[a] / (1.0 + sk)) < 3.0)
          return fail(c, MDPD_ERR_ARG, "replica " + std::to_string(r) +
                                           ": a box edge holds fewer than 3 cells of 1 + verlet_skin");
  if (t->nlist_width != c->tune.nlist_width) {
    HIPCHK(c, hipStreamSynchronize(c->st));
    const uint32_t cap64 = div_up(c->cap, 64) * 64;
    uint32_t* nl = nullptr;
    HIPCHK(c, hipMalloc(&nl, sizeof(uint32_t) * (size_t)cap64 * t->nlist_width));
    auto it = std::find(c->allocs.begin(), c->allocs.end(), (void*)c->d.nlist);
    if (it != c->allocs.end()) {
      hipFree(*it);
      *it = nl;
    }
    c->d.nlist = nl;
    c->d.W = t->nlist_width;
    c->have_density = false;
  }
  DevState& d = c->d;
  if (sk != c->tune.verlet_skin) {
    // new cell grid (cells only grow with the skin, so the cell arrays
    // allocated for the cutoff grid hold it)
    HIPCHK(c, hipStreamSynchronize(c->st));
    uint32_t cells = 0;
    for (int r = 0; r < c->n_rep; ++r) {
      set_cells(c->R[r], skin ? 1.0 + (double)sk : 1.0);
      c->R[r].cell_off = cells;
      c->cell_off[r] = cells;
      cells += c->R[r].n_cells;
    }
    c->cell_off[c->n_rep] = cells;
    HIPCHK(c, hipMemcpyAsync(d.reps, c->R.data(), sizeof(Replica) * c->n_rep,
                             hipMemcpyHostToDevice, c->st));
    HIPCHK(c, hipMemcpyAsync(d.rep_cell_off, c->cell_off.data(), sizeof(uint32_t) * (c->n_rep + 1),
                             hipMemcpyHostToDevice, c->st));
    // the step's cell histogram starts from zero
    launch_fill_u32(d.cell_cnt, c->ncell_alloc + 1, 0u, c->st);
    c->ncell = cells;
    d.ncell = cells;
    const double rv = 1.0 + (double)sk + 1e-3;
    d.vcut2 = (float)(rv * rv);
    d.disp2 = 0.25 * (double)sk * (double)sk;
    d.cell_slack = 0.0;
    c->have_cells = c->cells_fresh = c->have_density = false;
  }
  if (skin && (!d.vlist || t->verlet_width != d.Wv)) {
    HIPCHK(c, hipStreamSynchronize(c->st));
    const uint32_t cap64 = div_up(c->cap, 64) * 64;
    uint32_t* vl = nullptr;
    HIPCHK(c, hipMalloc(&vl, sizeof(uint32_t) * (size_t)cap64 * t->verlet_width));
    auto it = d.vlist ? std::find(c->allocs.begin(), c->allocs.end(), (void*)d.vlist)
                      : c->allocs.end();
    if (it != c->allocs.end()) {
      hipFree(*it);
      *it = vl;
    } else {
      c->allocs.push_back(vl);
    }
    d.vlist = vl;
    d.Wv = t->verlet_width;
    if (!d.vcount) {
      int rc;
      if ((rc = dalloc(c, &d.vcount, cap64)) != MDPD_OK) return rc;
      if ((rc = dalloc(c, &d.dref, 3 * (size_t)c->cap)) != MDPD_OK) return rc;
    }
  }
  c->tune = *t;
  invalidate_graphs(c);
  return MDPD_OK;
}

int mdpd_get_tuning(const mdpd_ctx* c, mdpd_tuning* t) {
  if (!c || !t) return MDPD_ERR_ARG;
  *t = c->tune;
  return MDPD_OK;
}

int mdpd_drift_predict(mdpd_ctx* c) {
  if (c) c->skin_valid = false;
  if (!c) return MDPD_ERR_ARG;
  launch_drift(c->d, false, false, c->st);
  c->drifted = true;
  c->cells_fresh = false;
  return MDPD_OK;
}

int mdpd_rebuild_cells(mdpd_ctx* c) {
  if (c) c->skin_valid = false;
  if (!c) return MDPD_ERR_ARG;
  const DevState& d = c->d;
  launch_fill_u32(d.cell_cnt, d.ncell + 1, 0u, c->st);
  launch_keys(d, c->st);
  launch_scan_cells(d, c->st);
  launch_scatter(d, c->st);
  launch_cell_sort(d, c->st);
  // the force uses the reference's vel: the predicted velocity after a
  // drift; without one, vel, acc and press are still live and move along
  launch_permute(d, c->drifted, true, c->st);
  swap_master(c);
  launch_topo_remap(c->d, c->st);
  c->d.cell_slack = 0.0;
  c->have_cells = c->cells_fresh = true;
  c->have_density = false;
  return MDPD_OK;
}

int mdpd_density(mdpd_ctx* c) {
  if (!c) return MDPD_ERR_ARG;
  if (!c->have_cells) return fail(c, MDPD_ERR_STATE, "density before rebuild_cells");
  launch_density(c->d, c->ns, c->nd, c->tune.block_scan, c->st);
  c->have_density = true;
  return check_flags(c);
}

int mdpd_force(mdpd_ctx* c) {
  if (!c) return MDPD_ERR_ARG;
  if (!c->have_density) return fail(c, MDPD_ERR_STATE, "force before density");
  HIPCHK(c, hipMemcpyAsync(c->d.d_time, &c->time, 4, hipMemcpyHostToDevice, c->st));
  launch_force(c->d, c->ns, c->nd, c->env, true, c->tune.block_force, c->st);
  return check_flags(c);
}

int mdpd_bonded(mdpd_ctx* c, double vir[3]) {
  if (!c) return MDPD_ERR_ARG;
  // positions may have changed since the last permutation (drift, upload)
  launch_chainpos_fill(c->d, c->st);
  launch_bonded_kick(c->d, true, false, c->vir_bonded, c->st);
  if (vir) {
    HIPCHK(c, hipMemcpyAsync(vir, c->vir_bonded, 24, hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
  }
  return MDPD_OK;
}

int mdpd_kick(mdpd_ctx* c) {
  if (!c) return MDPD_ERR_ARG;
  launch_kick(c->d, c->st);
  c->drifted = false;
  return MDPD_OK;
}

int mdpd_remove_cm_drift(mdpd_ctx* c) {
  if (!c) return MDPD_ERR_ARG;
  if (c->drifted) {
    // the reference's vel is the predicted velocity here; vel_buf stays
    for (int a = 0; a < 3; ++a)
      launch_copy_d(c->d.vel + (size_t)a * c->d.cap, c->d.vbuf + (size_t)a * c->d.cap, c->d.n, c->st);
    c->drifted = false;
  }
  launch_cm_drift(c->d, c->st);
  return MDPD_OK;
}

int mdpd_head_conversion(mdpd_ctx* c) {
  if (!c) return MDPD_ERR_ARG;
  HIPCHK(c, hipMemcpyAsync(c->d.d_time, &c->time, 4, hipMemcpyHostToDevice, c->st));
  launch_head_conversion(c->d, c->st);
  HIPCHK(c, hipStreamSynchronize(c->st));
  return MDPD_OK;
}

int mdpd_run(mdpd_ctx* c, uint32_t nsteps) {
  if (!c) return MDPD_ERR_ARG;
  if (!c->d.n) return fail(c, MDPD_ERR_STATE, "no beads uploaded");
  if (c->chem_on && c->chem_stopped)
    return fail(c, MDPD_ERR_STATE, "the chemistry stopped the run (max_amp_num reached); "
                                   "mdpd_set_chemistry starts it again");
  if (c->drifted) return fail(c, MDPD_ERR_STATE, "run called between drift and kick");
  int rc = run_steps(c, nsteps);
  if (rc) return rc;
  return check_flags(c);
}

int mdpd_prepare_run(mdpd_ctx* c, uint32_t nsteps) {
  if (!c) return MDPD_ERR_ARG;
  if (!c->d.n) return fail(c, MDPD_ERR_STATE, "no beads uploaded");
  if (c->drifted) return fail(c, MDPD_ERR_STATE, "prepare_run called between drift and kick");
  return prepare_graphs(c, nsteps);
}

int mdpd_run_timed(mdpd_ctx* c, uint32_t nsteps, double* gpu_ms) {
  if (!c || !gpu_ms) return MDPD_ERR_ARG;
  if (!c->d.n) return fail(c, MDPD_ERR_STATE, "no beads uploaded");
  if (c->chem_on && c->chem_stopped)
    return fail(c, MDPD_ERR_STATE, "the chemistry stopped the run (max_amp_num reached); "
                                   "mdpd_set_chemistry starts it again");
  if (c->drifted) return fail(c, MDPD_ERR_STATE, "run called between drift and kick");
  {
    // the step graphs are recorded before the first event
    const int rc = prepare_graphs(c, nsteps);
    if (rc) return rc;
  }
  hipEvent_t e0, e1;
  HIPCHK(c, hipEventCreate(&e0));
  HIPCHK(c, hipEventCreate(&e1));
  HIPCHK(c, hipEventRecord(e0, c->st));
  int rc = run_steps(c, nsteps);
  if (rc == MDPD_OK) {
    HIPCHK(c, hipEventRecord(e1, c->st));
    HIPCHK(c, hipEventSynchronize(e1));
    float ms = 0.f;
    HIPCHK(c, hipEventElapsedTime(&ms, e0, e1));
    *gpu_ms = ms;
  }
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  if (rc) return rc;
  return check_flags(c);
}

int mdpd_set_chemistry(mdpd_ctx* c, const mdpd_chem_params* cp, uint32_t beg_chem, int32_t enable) {
  if (!c) return MDPD_ERR_ARG;
  if (!enable) {
    c->chem_on = false;
    c->chem_stopped = false;
    return MDPD_OK;
  }
  if (!cp) return fail(c, MDPD_ERR_ARG, "null chemistry parameters");
  if (cp->n_core > kMaxCore) return fail(c, MDPD_ERR_ARG, "at most 256 core amphiphiles");
  if (cp->n_core && !cp->core_ptcl_id) return fail(c, MDPD_ERR_ARG, "core_ptcl_id is null");
  if (c->tune.verlet_skin > 0.0f)
    return fail(c, MDPD_ERR_UNSUPPORTED, "chemistry inside mdpd_run needs verlet_skin = 0");
  int rc = enter_grow(c);
  if (rc) return rc;
  ChemDev& h = c->chem_host[0];
  if (h.n_core != cp->n_core) invalidate_graphs(c);  // the launch grids depend on it
  h = ChemDev{};
  h.p_thresld = cp->p_thresld;
  h.eps = cp->eps;
  h.influ_rad = cp->influ_rad;
  h.influ_hei = cp->influ_hei;
  h.influ_dep = cp->influ_dep;
  h.influ_grd = cp->influ_grd;
  h.bond_leng = c->P[0].bond_leng;
  h.cf_b = c->P[0].cf_b;
  h.max_amp_num = cp->max_amp_num;
  h.n_core = cp->n_core;
  h.in_run = 1;
  for (uint32_t k = 0; k < cp->n_core; ++k) h.core_id[k] = cp->core_ptcl_id[k];
  HIPCHK(c, hipMemcpyAsync(c->chem_dev, &h, sizeof(ChemDev), hipMemcpyHostToDevice, c->st));
  if ((rc = chem_init_slots(c))) return rc;
  c->chem_on = true;
  c->chem_stopped = false;
  c->beg_chem = beg_chem;
  return MDPD_OK;
}

}  // extern "C"

namespace mdpd {
// the reference-MT mode on, from the stream state h (buffers allocated on
// the first use and kept)
int chem_mt_enable(mdpd_ctx* c, const MtState& h) {
  if (c->n_rep != 1 || !c->chem_dev)
    return fail(c, MDPD_ERR_UNSUPPORTED, "chemistry needs a single-system context");
  int rc;
  ChemMt& m = c->mt_store;
  if (!m.g) {
    const uint32_t amp_cap = c->cap / 4u + 1u;
    m.cone_words = (uint32_t)(((uint64_t)kMaxCore * amp_cap) / 32u + 2u);
    if ((rc = dalloc(c, &m.g, 1)) || (rc = dalloc(c, &m.cone_offs, (size_t)kMaxCore * chem_count_blocks(c->cap))) ||
        (rc = dalloc(c, &m.cone_bits, m.cone_words)) ||
        (rc = dalloc(c, &m.tdraws, (size_t)kMtDrawsPerTarget * amp_cap))) {
      m = ChemMt{};
      return rc;
    }
  }
  HIPCHK(c, hipMemcpyAsync(m.g, &h, sizeof(MtState), hipMemcpyHostToDevice, c->st));
  HIPCHK(c, hipStreamSynchronize(c->st));
  if (!c->chem_mt.g) invalidate_graphs(c);  // the captured events change their kernels
  c->chem_mt = m;
  return MDPD_OK;
}

int chem_mt_state(mdpd_ctx* c, MtState* h) {
  HIPCHK(c, hipMemcpyAsync(h, c->chem_mt.g, sizeof(MtState), hipMemcpyDeviceToHost, c->st));
  HIPCHK(c, hipStreamSynchronize(c->st));
  return MDPD_OK;
}
}  // namespace mdpd

extern "C" {

int mdpd_set_chem_mt(mdpd_ctx* c, uint32_t seed, int32_t enable) {
  if (!c) return MDPD_ERR_ARG;
  if (c->n_rep != 1 || !c->chem_dev)
    return fail(c, MDPD_ERR_UNSUPPORTED, "chemistry needs a single-system context");
  if (!enable) {
    if (c->chem_mt.g) invalidate_graphs(c);  // the captured events draw from the stream
    c->chem_mt = ChemMt{};
    return MDPD_OK;
  }
  // init_genrand (mt19937ar): the state ChemManager(seed) starts from
  MtState h;
  h.mt[0] = seed;
  for (uint32_t i = 1; i < kMtN; ++i) h.mt[i] = 1812433253u * (h.mt[i - 1] ^ (h.mt[i - 1] >> 30)) + i;
  h.mti = kMtN;
  h.pad = 0;
  h.draws = 0;
  return chem_mt_enable(c, h);
}

int mdpd_chem_mt_draws(mdpd_ctx* c, uint64_t* draws) {
  if (!c || !draws) return MDPD_ERR_ARG;
  *draws = 0;
  if (!c->chem_mt.g) return MDPD_OK;
  MtState h;
  const int rc = chem_mt_state(c, &h);
  if (rc) return rc;
  *draws = h.draws;
  return MDPD_OK;
}

int mdpd_chem_status(mdpd_ctx* c, int32_t* stopped, uint32_t* events, uint32_t* inserted) {
  if (!c || !stopped) return MDPD_ERR_ARG;
  *stopped = c->chem_stopped ? 1 : 0;
  uint32_t v[8] = {0};
  if (c->chem_state) {
    HIPCHK(c, hipMemcpyAsync(v, c->chem_state, sizeof(v), hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
  }
  if (events) *events = v[5];
  if (inserted) *inserted = v[6];
  return MDPD_OK;
}

int mdpd_skin_stats(mdpd_ctx* c, uint64_t* builds, uint64_t* entries) {
  if (!c || !builds) return MDPD_ERR_ARG;
  uint32_t v = 0;
  HIPCHK(c, hipMemcpyAsync(&v, c->d.rebuild + 1, 4, hipMemcpyDeviceToHost, c->st));
  HIPCHK(c, hipStreamSynchronize(c->st));
  *builds = v;
  if (entries) {
    *entries = 0;
    if (c->d.vcount && c->d.n) {
      std::vector<uint32_t> cnt(c->d.n);
      HIPCHK(c, hipMemcpy(cnt.data(), c->d.vcount, 4 * (size_t)c->d.n, hipMemcpyDeviceToHost));
      for (uint32_t k : cnt) *entries += std::min<uint32_t>(k, (uint32_t)c->d.Wv);
    }
  }
  return MDPD_OK;
}

int mdpd_sync(mdpd_ctx* c) {
  if (!c) return MDPD_ERR_ARG;
  HIPCHK(c, hipStreamSynchronize(c->st));
  return MDPD_OK;
}

static int observe_impl(mdpd_ctx* c, mdpd_observables* out, int all) {
  if (!c || !out) return MDPD_ERR_ARG;
  const DevState& d = c->d;
  launch_observe(d, c->d_amp_num, c->dl_load, c->obs_out, c->st);
  std::vector<double> q((size_t)c->n_rep * 13), bv((size_t)c->n_rep * 3);
  HIPCHK(c, hipMemcpyAsync(q.data(), c->obs_out, sizeof(double) * q.size(), hipMemcpyDeviceToHost, c->st));
  HIPCHK(c, hipMemcpyAsync(bv.data(), c->vir_bonded, sizeof(double) * bv.size(), hipMemcpyDeviceToHost, c->st));
  HIPCHK(c, hipStreamSynchronize(c->st));
  const int nr = all ? c->n_rep : 1;
  for (int r = 0; r < nr; ++r) {
    mdpd_observables& o = out[r];
    memset(&o, 0, sizeof(o));
    const double* s = &q[(size_t)r * 13];
    o.n_pairs = s[12];
    const uint32_t n_r = c->rep_off[r + 1] - c->rep_off[r];
    for (int a = 0; a < 3; ++a) o.kin_temp[a] = s[a] / n_r;
    o.kin_temp[3] = (o.kin_temp[0] + o.kin_temp[1] + o.kin_temp[2]) / 3.0;
    const double iv = (1.0 / c->R[r].box[0]) * (1.0 / c->R[r].box[1]) * (1.0 / c->R[r].box[2]);
    for (int a = 0; a < 3; ++a) {
      o.bonded_vir[a] = bv[(size_t)r * 3 + a];
      o.pressure[a] = (s[3 + a] + o.bonded_vir[a] * 0.5) * iv;
      o.pair_vir[a] = s[6 + a];
    }
    const uint32_t na = c->amp_num_rep[r], ns = c->sol_num_rep[r];
    o.diffusion[0] = s[9] / n_r;
    o.diffusion[1] = na ? s[11] / na : 0.0;
    o.diffusion[2] = ns ? s[10] / ns : 0.0;
    o.n_beads = n_r;
    o.n_amp = na;
    o.time = c->time;
  }
  return MDPD_OK;
}

int mdpd_observe(mdpd_ctx* c, mdpd_observables* out) { return observe_impl(c, out, 0); }
int mdpd_observe_replicas(mdpd_ctx* c, mdpd_observables* out) { return observe_impl(c, out, 1); }

int mdpd_density_profile(mdpd_ctx* c, int32_t axis, int32_t nbins, double* out) {
  if (!c || !out || axis < 0 || axis > 2 || nbins < 1) return fail(c, MDPD_ERR_ARG, "bad arguments");
  uint32_t* h = nullptr;
  const size_t cnt = (size_t)c->n_rep * nbins;
  HIPCHK(c, hipMalloc(&h, 4 * cnt));
  HIPCHK(c, hipMemsetAsync(h, 0, 4 * cnt, c->st));
  launch_density_profile(c->d, axis, nbins, h, c->st);
  std::vector<uint32_t> hv(cnt);
  HIPCHK(c, hipMemcpyAsync(hv.data(), h, 4 * cnt, hipMemcpyDeviceToHost, c->st));
  HIPCHK(c, hipStreamSynchronize(c->st));
  hipFree(h);
  for (size_t k = 0; k < cnt; ++k) out[k] = hv[k];
  return MDPD_OK;
}

int mdpd_neighbor_pairs(mdpd_ctx* c, double cutoff, uint32_t* out_i, uint32_t* out_j, uint64_t cap,
                        uint64_t* n_out) {
  if (!c || !n_out) return MDPD_ERR_ARG;
  if (!c->have_cells || !c->cells_fresh) return fail(c, MDPD_ERR_STATE, "no current cell list");
  if (!(cutoff > 0.0 && cutoff <= 1.0)) return fail(c, MDPD_ERR_ARG, "cutoff must be in (0, 1]");
  const uint32_t n = c->d.n;
  std::vector<uint64_t> keys;
  if (cutoff == 1.0 && c->have_density) {
    // the production pair list of the last density pass
    const uint32_t W = (uint32_t)c->d.W, cap64 = div_up(n, 64) * 64;
    std::vector<uint32_t> nl((size_t)cap64 * W), cnt(n), perm(n);
    HIPCHK(c, hipMemcpyAsync(nl.data(), c->d.nlist, 4 * nl.size(), hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipMemcpyAsync(cnt.data(), c->d.ncount, 4 * (size_t)n, hipMemcpyDeviceToHost, c->st));
    // report beads by their upload slot
    HIPCHK(c, hipMemcpyAsync(perm.data(), c->d.orig, 4 * (size_t)n, hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    for (uint32_t s = 0; s < n; ++s) {
      if (cnt[s] > W) return fail(c, MDPD_ERR_NLIST_OVERFLOW, "pair list overflow");
      for (uint32_t k = 0; k < cnt[s]; ++k) {
        // chunks of four entries: [tile][k / 4][lane][k % 4]
        const uint32_t e = nl[((size_t)(s >> 6) * W * 64 + (size_t)(k >> 2) * 256 +
                               (size_t)(s & 63) * 4) + (k & 3)];
        keys.push_back(((uint64_t)perm[s] << 32) | perm[e & kIndexMask]);
      }
    }
  } else {
    unsigned long long* ctr = nullptr;
    uint32_t *oi = nullptr, *oj = nullptr;
    const uint64_t lcap = (uint64_t)n * 128 + 1024;
    HIPCHK(c, hipMalloc(&ctr, 8));
    HIPCHK(c, hipMalloc(&oi, 4 * lcap));
    HIPCHK(c, hipMalloc(&oj, 4 * lcap));
    HIPCHK(c, hipMemsetAsync(ctr, 0, 8, c->st));
    launch_pairs_list(c->d, cutoff, oi, oj, ctr, lcap, c->st);
    unsigned long long m = 0;
    HIPCHK(c, hipMemcpyAsync(&m, ctr, 8, hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    if (m > lcap) return fail(c, MDPD_ERR_CAPACITY, "pair buffer too small");
    std::vector<uint32_t> hi(m), hj(m);
    HIPCHK(c, hipMemcpy(hi.data(), oi, 4 * m, hipMemcpyDeviceToHost));
    HIPCHK(c, hipMemcpy(hj.data(), oj, 4 * m, hipMemcpyDeviceToHost));
    hipFree(ctr);
    hipFree(oi);
    hipFree(oj);
    keys.resize(m);
    for (uint64_t k = 0; k < m; ++k) keys[k] = ((uint64_t)hi[k] << 32) | hj[k];
  }
  std::sort(keys.begin(), keys.end());
  *n_out = keys.size();
  if (out_i && out_j)
    for (uint64_t k = 0; k < keys.size() && k < cap; ++k) {
      out_i[k] = (uint32_t)(keys[k] >> 32);
      out_j[k] = (uint32_t)(keys[k] & 0xffffffffu);
    }
  return MDPD_OK;
}

int mdpd_saru_device(const uint32_t* s1, const uint32_t* s2, const uint32_t* s3, uint32_t n,
                     uint32_t* out_u32, float* out_nrml) {
  if (!s1 || !s2 || !s3 || !out_u32 || !out_nrml) return fail(nullptr, MDPD_ERR_ARG, "null argument");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail(nullptr, MDPD_ERR_NO_DEVICE, "no HIP device");
  uint32_t *a = nullptr, *b = nullptr, *cc = nullptr, *u = nullptr;
  float* f = nullptr;
  const size_t B = 4 * (size_t)std::max<uint32_t>(n, 1);
  if (hipMalloc(&a, B) || hipMalloc(&b, B) || hipMalloc(&cc, B) || hipMalloc(&u, 8 * B) || hipMalloc(&f, B))
    return fail(nullptr, MDPD_ERR_HIP, "hipMalloc failed");
  hipMemcpy(a, s1, 4 * (size_t)n, hipMemcpyHostToDevice);
  hipMemcpy(b, s2, 4 * (size_t)n, hipMemcpyHostToDevice);
  hipMemcpy(cc, s3, 4 * (size_t)n, hipMemcpyHostToDevice);
  launch_saru_test(a, b, cc, n, u, f, nullptr);
  hipError_t e = hipDeviceSynchronize();
  hipMemcpy(out_u32, u, 32 * (size_t)n, hipMemcpyDeviceToHost);
  hipMemcpy(out_nrml, f, 4 * (size_t)n, hipMemcpyDeviceToHost);
  hipFree(a); hipFree(b); hipFree(cc); hipFree(u); hipFree(f);
  if (e != hipSuccess) return fail(nullptr, MDPD_ERR_HIP, hipGetErrorString(e));
  return MDPD_OK;
}

// fold the recorded per-step events into the per-phase sums
static int collect_events(mdpd_ctx* c) {
  if (!c->ev_steps) return MDPD_OK;
  HIPCHK(c, hipEventSynchronize(c->evpool[(size_t)c->ev_steps * (kPhases + 1) - 1]));
  for (uint32_t k = 0; k < c->ev_steps; ++k) {
    hipEvent_t* ev = &c->evpool[(size_t)k * (kPhases + 1)];
    for (int p = 0; p < kPhases; ++p) {
      float ms = 0.f;
      HIPCHK(c, hipEventElapsedTime(&ms, ev[p], ev[p + 1]));
      c->ph_ms[p] += ms;
    }
  }
  // whole-region time: first mark of the first step to the last mark of the last
  float span = 0.f;
  HIPCHK(c, hipEventElapsedTime(&span, c->evpool[0],
                                c->evpool[(size_t)c->ev_steps * (kPhases + 1) - 1]));
  c->ph_ms[kPhases] += span;
  c->ph_steps += c->ev_steps;
  c->ev_steps = 0;
  return MDPD_OK;
}

int mdpd_set_profiling(mdpd_ctx* c, int32_t on) {
  if (!c) return MDPD_ERR_ARG;
  HIPCHK(c, hipStreamSynchronize(c->st));
  c->profiling = on != 0;
  for (double& x : c->ph_ms) x = 0.0;
  c->ph_steps = 0;
  c->ev_steps = 0;
  return MDPD_OK;
}

int mdpd_phase_times(mdpd_ctx* c, double* ms, int32_t cap, int32_t* n, char* names, size_t names_len) {
  if (!c || !n) return MDPD_ERR_ARG;
  int rc = collect_events(c);
  if (rc) return rc;
  *n = kPhases + 1;
  for (int p = 0; p <= kPhases && p < cap; ++p) ms[p] = c->ph_steps ? c->ph_ms[p] / c->ph_steps : 0.0;
  if (names && names_len) {
    size_t off = 0;
    for (int p = 0; p <= kPhases; ++p) {
      const char* nm = p < kPhases ? kPhaseNames[p] : "step";
      const size_t l = strlen(nm);
      if (off + l + 1 >= names_len) break;
      memcpy(names + off, nm, l);
      names[off + l] = '\0';
      off += l + 1;
    }
    if (off < names_len) names[off] = '\0';
  }
  return MDPD_OK;
}

int mdpd_bandwidth_probe(uint64_t bytes, int32_t reps, double* gbps) {
  if (!gbps || bytes < 1024 || reps < 1) return fail(nullptr, MDPD_ERR_ARG, "bad arguments");
  double *a = nullptr, *b = nullptr;
  const size_t nd = bytes / 8;
  if (hipMalloc(&a, nd * 8) != hipSuccess || hipMalloc(&b, nd * 8) != hipSuccess)
    return fail(nullptr, MDPD_ERR_HIP, "hipMalloc failed");
  hipMemset(a, 0, nd * 8);
  hipEvent_t e0, e1;
  hipEventCreate(&e0);
  hipEventCreate(&e1);
  launch_copy_d(b, a, nd, nullptr);
  hipEventRecord(e0, nullptr);
  for (int k = 0; k < reps; ++k) launch_copy_d(b, a, nd, nullptr);
  hipEventRecord(e1, nullptr);
  const hipError_t e = hipEventSynchronize(e1);
  float ms = 0.f;
  hipEventElapsedTime(&ms, e0, e1);
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  hipFree(a);
  hipFree(b);
  if (e != hipSuccess) return fail(nullptr, MDPD_ERR_HIP, hipGetErrorString(e));
  *gbps = 2.0 * (double)nd * 8.0 * reps / (ms * 1e-3) / 1e9;
  return MDPD_OK;
}

}  // extern "C"
