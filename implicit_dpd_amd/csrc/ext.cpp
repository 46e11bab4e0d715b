// Chemistry event and launch autotuner of the engine (host side).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "ctx.hpp"

using namespace mdpd;

extern "C" {

// ChemManager::RandomChemEvent, serial (chemmanager.hpp:244-281) with
// DetermineTargetIdWithCorePos (:116-164) and GetCorePos (:84-114): the
// whole event runs on the device (the kernels mdpd_run uses inside its
// steps), with the context's preallocated scratch; the host reads back the
// counts once.  `time` is the context clock.
int mdpd_chem_event(mdpd_ctx* c, const mdpd_chem_params* cp, int* stop) {
  if (c) c->skin_valid = false;
  if (!c || !cp || !stop) return fail(c, MDPD_ERR_ARG, "null argument");
  if (c->n_rep != 1) return fail(c, MDPD_ERR_UNSUPPORTED, "chemistry needs a single-system context");
  if (!c->cells_fresh) return fail(c, MDPD_ERR_STATE, "chemistry needs the cell list of the current positions");
  if (cp->n_core && !cp->core_ptcl_id) return fail(c, MDPD_ERR_ARG, "core_ptcl_id is null");
  if (cp->n_core > kMaxCore) return fail(c, MDPD_ERR_ARG, "at most 256 core amphiphiles");
  int rc = enter_grow(c);
  if (rc) return rc;
  DevState& d = c->d;
  ChemDev& h = c->chem_host[1];
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
  h.in_run = 0;
  for (uint32_t k = 0; k < cp->n_core; ++k) h.core_id[k] = cp->core_ptcl_id[k];
  HIPCHK(c, hipMemcpyAsync(c->chem_dev + 1, &h, sizeof(ChemDev), hipMemcpyHostToDevice, c->st));
  launch_fill_u32(d.d_time, 1, c->time, c->st);
  // head and unit-2 positions of every chain, from the current positions
  launch_chainpos_fill(d, c->st);
  launch_chem_event(d, c->chem_dev + 1, c->chem_state + 1, cp->n_core, c->chem_counts,
                    c->chem_masks, c->chem_targets, c->d_amp_num, false, c->chem_mt, c->st);
  HIPCHK(c, hipGetLastError());
  uint32_t res[5];
  HIPCHK(c, hipMemcpyAsync(res, c->chem_state + 1, sizeof(res), hipMemcpyDeviceToHost, c->st));
  if ((rc = pull_counts(c))) return rc;
  if (res[4]) {  // refused: nothing was inserted
    launch_fill_u32(d.flags, 1, 0u, c->st);
    hipStreamSynchronize(c->st);
    return fail(c, MDPD_ERR_CAPACITY, "inserted amphiphiles exceed the capacity");
  }
  if (res[0]) {
    c->cells_fresh = false;
    c->have_cells = false;
    c->have_density = false;
  }
  *stop = (c->n_amp >= cp->max_amp_num) ? 1 : 0;
  return MDPD_OK;
}

// ---- binary checkpoint (exact restart)
// Layout, little-endian: magic "MDPDCKP1", u32 version (2), n_rep, n, n_amp,
// time, flags; u32 rep_off[n_rep + 1], pair_seed[n_rep], chem_seed[n_rep],
// amp_num[n_rep], sol_num[n_rep]; u64 fingerprint[n_rep] of each replica's
// parameters; FP64 pos, vel, vel_buf, delta_sumr [3][n]; FP32 acc, press
// [3][n]; u32 id, prop, amp_id, unit, orig, chain [n]; u32 topo[4 n_amp];
// FP64 cf_bend[n_amp]; u32 amp_rep[n_amp]; with flags bit 0 (the chemistry
// draws from the reference's MT19937 stream) then u32 mt[624], mti, u64
// outputs drawn.  Arrays are in the device's
// storage order, so the next cell sort (ordered by previous slot) and every
// FP32 sum after it repeat the uninterrupted run.
namespace {
constexpr char kCkpMagic[8] = {'M', 'D', 'P', 'D', 'C', 'K', 'P', '1'};
constexpr uint32_t kCkpVersion = 2;

// FNV-1a over a replica's parameters with the seeds (restored from the
// file) and the padding cleared: a checkpoint only loads into a context of
// the same box and model
uint64_t param_fingerprint(const mdpd_params& p0) {
  mdpd_params p;
  std::memset(&p, 0, sizeof(p));
  std::memcpy(&p, &p0, sizeof(p));
  p.pair_seed = 0;
  p.chem_seed = 0;
  p.pad0 = 0;
  p.pad1 = 0;
  uint64_t h = 1469598103934665603ull;
  const unsigned char* b = reinterpret_cast<const unsigned char*>(&p);
  for (size_t k = 0; k < sizeof(p); ++k) {
    h ^= b[k];
    h *= 1099511628211ull;
  }
  return h;
}
}  // namespace

int mdpd_checkpoint_save(mdpd_ctx* c, const char* path) {
  if (!c || !path) return fail(c, MDPD_ERR_ARG, "null argument");
  if (c->drifted) return fail(c, MDPD_ERR_STATE, "checkpoint between drift and kick");
  const DevState& d = c->d;
  const uint32_t n = d.n, na = d.n_amp, nr = (uint32_t)c->n_rep;
  HIPCHK(c, hipStreamSynchronize(c->st));
  std::vector<double> f64(3 * (size_t)n);
  std::vector<float> f32(3 * (size_t)n);
  std::vector<uint32_t> u32((size_t)std::max(n, 4 * na));
  FILE* fp = fopen(path, "wb");
  if (!fp) return fail(c, MDPD_ERR_ARG, std::string("cannot open ") + path);
  bool ok = true;
  auto put = [&](const void* p, size_t bytes) { ok = ok && fwrite(p, 1, bytes, fp) == bytes; };
  const uint32_t head[6] = {kCkpVersion, nr, n, na, c->time, c->chem_mt.g ? 1u : 0u};
  put(kCkpMagic, 8);
  put(head, sizeof(head));
  put(c->rep_off.data(), 4 * (size_t)(nr + 1));
  for (uint32_t r = 0; r < nr; ++r) put(&c->R[r].pair_seed, 4);
  for (uint32_t r = 0; r < nr; ++r) put(&c->R[r].chem_seed, 4);
  put(c->amp_num_rep.data(), 4 * (size_t)nr);
  put(c->sol_num_rep.data(), 4 * (size_t)nr);
  for (uint32_t r = 0; r < nr; ++r) {
    const uint64_t fp = param_fingerprint(c->P[r]);
    put(&fp, 8);
  }
  int rc = MDPD_OK;
  auto get3d = [&](const double* src) -> int {
    for (int a = 0; a < 3; ++a)
      HIPCHK(c, hipMemcpy(f64.data() + (size_t)a * n, src + (size_t)a * d.cap, 8 * (size_t)n,
                          hipMemcpyDeviceToHost));
    put(f64.data(), 8 * f64.size());
    return MDPD_OK;
  };
  auto get3f = [&](const float* src) -> int {
    for (int a = 0; a < 3; ++a)
      HIPCHK(c, hipMemcpy(f32.data() + (size_t)a * n, src + (size_t)a * d.cap, 4 * (size_t)n,
                          hipMemcpyDeviceToHost));
    put(f32.data(), 4 * f32.size());
    return MDPD_OK;
  };
  auto getu = [&](const uint32_t* src, size_t cnt) -> int {
    HIPCHK(c, hipMemcpy(u32.data(), src, 4 * cnt, hipMemcpyDeviceToHost));
    put(u32.data(), 4 * cnt);
    return MDPD_OK;
  };
  if (!rc) rc = get3d(d.pos);
  if (!rc) rc = get3d(d.vel);
  if (!rc) rc = get3d(d.vbuf);
  if (!rc) rc = get3d(d.delta);
  if (!rc) rc = get3f(d.acc);
  if (!rc) rc = get3f(d.press);
  const uint32_t* const us[6] = {d.id, d.prop, d.amp_id, d.unit, d.orig, d.chain};
  for (int q = 0; q < 6 && !rc; ++q) rc = getu(us[q], n);
  if (!rc && na) {
    // the step keeps only the per-bead chain map current
    launch_topo_from_chain(d, c->st);
    rc = getu(d.topo, 4 * (size_t)na);
    if (!rc) {
      std::vector<double> cf(na);
      if (hipMemcpy(cf.data(), d.cf_bend, 8 * (size_t)na, hipMemcpyDeviceToHost) != hipSuccess)
        rc = fail(c, MDPD_ERR_HIP, "checkpoint: copy of cf_bend failed");
      put(cf.data(), 8 * (size_t)na);
    }
    if (!rc) rc = getu(d.amp_rep, na);
  }
  if (!rc && c->chem_mt.g) {
    MtState h;
    rc = chem_mt_state(c, &h);
    if (!rc) {
      put(h.mt, sizeof(h.mt));
      put(&h.mti, 4);
      put(&h.draws, 8);
    }
  }
  if (fclose(fp) != 0) ok = false;
  if (rc) return rc;
  if (!ok) return fail(c, MDPD_ERR_ARG, std::string("write failed: ") + path);
  return MDPD_OK;
}

int mdpd_checkpoint_load(mdpd_ctx* c, const char* path) {
  if (c) c->skin_valid = false;
  if (!c || !path) return fail(c, MDPD_ERR_ARG, "null argument");
  FILE* fp = fopen(path, "rb");
  if (!fp) return fail(c, MDPD_ERR_ARG, std::string("cannot open ") + path);
  bool ok = true;
  auto get = [&](void* p, size_t bytes) { ok = ok && fread(p, 1, bytes, fp) == bytes; };
  char magic[8];
  uint32_t head[6];
  get(magic, 8);
  get(head, sizeof(head));
  if (!ok || std::memcmp(magic, kCkpMagic, 8) != 0 || head[0] != kCkpVersion) {
    fclose(fp);
    return fail(c, MDPD_ERR_PARSE, std::string("not a checkpoint of this format (version 2): ") + path);
  }
  const uint32_t nr = head[1], n = head[2], na = head[3], flags = head[5];
  if (flags > 1u) {
    fclose(fp);
    return fail(c, MDPD_ERR_PARSE, std::string("unknown checkpoint flags: ") + path);
  }
  if ((flags & 1u) && (nr != 1u || !c->chem_dev)) {
    fclose(fp);
    return fail(c, MDPD_ERR_ARG, "checkpoint with a chemistry stream needs a single-system context");
  }
  if ((int)nr != c->n_rep || n > c->cap || na > c->amp_cap) {
    fclose(fp);
    return fail(c, MDPD_ERR_ARG, "checkpoint does not fit this context (replicas or capacity)");
  }
  std::vector<uint32_t> roff(nr + 1), seed(nr), cseed(nr), anum(nr), snum(nr);
  std::vector<uint64_t> fprint(nr);
  get(roff.data(), 4 * (size_t)(nr + 1));
  get(seed.data(), 4 * (size_t)nr);
  get(cseed.data(), 4 * (size_t)nr);
  get(anum.data(), 4 * (size_t)nr);
  get(snum.data(), 4 * (size_t)nr);
  get(fprint.data(), 8 * (size_t)nr);
  if (!ok) {
    fclose(fp);
    return fail(c, MDPD_ERR_PARSE, std::string("truncated checkpoint: ") + path);
  }
  for (uint32_t r = 0; r < nr; ++r)
    if (fprint[r] != param_fingerprint(c->P[r])) {
      fclose(fp);
      return fail(c, MDPD_ERR_ARG, "checkpoint of other parameters (box or model) than replica " +
                                       std::to_string(r) + " of this context");
    }
  if (nr > 1 ? roff != c->rep_off : (roff[0] != 0 || roff[1] != n)) {
    fclose(fp);
    return fail(c, MDPD_ERR_ARG, "checkpoint replica layout differs from this context");
  }
  std::vector<double> f64(12 * (size_t)n);
  std::vector<float> f32(6 * (size_t)n);
  std::vector<uint32_t> u32(6 * (size_t)n + 4 * (size_t)na + na);
  std::vector<double> cf(na);
  get(f64.data(), 8 * f64.size());
  get(f32.data(), 4 * f32.size());
  get(u32.data(), 4 * 6 * (size_t)n);
  if (na) {
    get(u32.data() + 6 * (size_t)n, 16 * (size_t)na);
    get(cf.data(), 8 * (size_t)na);
    get(u32.data() + 6 * (size_t)n + 4 * (size_t)na, 4 * (size_t)na);
  }
  MtState mts{};
  if (flags & 1u) {
    get(mts.mt, sizeof(mts.mt));
    get(&mts.mti, 4);
    get(&mts.draws, 8);
  }
  // nothing more may follow
  char extra;
  if (ok && fread(&extra, 1, 1, fp) == 1) ok = false;
  fclose(fp);
  if (!ok) return fail(c, MDPD_ERR_PARSE, std::string("truncated or oversized checkpoint: ") + path);
  // every word that a kernel uses as an index is checked here, before
  // anything reaches the device: a corrupt file is refused and the context
  // keeps its state
  {
    const uint32_t* prop = u32.data() + 1 * (size_t)n;
    const uint32_t* orig = u32.data() + 4 * (size_t)n;
    const uint32_t* chain = u32.data() + 5 * (size_t)n;
    const uint32_t* topo = u32.data() + 6 * (size_t)n;
    const uint32_t* arep = u32.data() + 6 * (size_t)n + 4 * (size_t)na;
    std::string why;
    uint64_t asum = 0;
    for (uint32_t r = 0; r < nr; ++r) asum += anum[r];
    if (asum != na) why = "amphiphile counts do not add up";
    for (uint32_t r = 0; r < nr && why.empty(); ++r)
      if (snum[r] > roff[r + 1] - roff[r]) why = "solvent count exceeds replica " + std::to_string(r);
    std::vector<uint8_t> seen_orig(n, 0), seen_chain(4 * (size_t)na, 0);
    for (uint32_t r = 0; r < nr && why.empty(); ++r)
      for (uint32_t i = roff[r]; i < roff[r + 1] && why.empty(); ++i) {
        if (prop[i] > 2u) { why = "bead type > 2 at slot " + std::to_string(i); break; }
        if (orig[i] >= n || seen_orig[orig[i]]++) { why = "slot map is not a permutation"; break; }
        if (chain[i] != 0xffffffffu) {
          if ((chain[i] >> 2) >= na || seen_chain[chain[i]]++) {
            why = "chain word out of range or repeated at slot " + std::to_string(i);
            break;
          }
        }
        for (int a = 0; a < 3; ++a) {
          const double x = f64[(size_t)a * n + i];
          if (!std::isfinite(x) || x < 0.0 || x > c->P[r].box[a]) {
            why = "position outside the box at slot " + std::to_string(i);
            break;
          }
          for (int q = 1; q < 4; ++q)
            if (!std::isfinite(f64[((size_t)q * 3 + a) * n + i])) why = "non-finite state at slot " + std::to_string(i);
          for (int q = 0; q < 2; ++q)
            if (!std::isfinite(f32[((size_t)q * 3 + a) * n + i])) why = "non-finite force at slot " + std::to_string(i);
        }
      }
    for (size_t k = 0; k < 4 * (size_t)na && why.empty(); ++k) {
      if (topo[k] >= n) why = "topology entry out of range";
      else if (!seen_chain[k]) why = "chain slot without its bead";
    }
    for (uint32_t a = 0; a < na && why.empty(); ++a)
      if (arep[a] >= nr || !std::isfinite(cf[a])) why = "chain replica or bending constant invalid";
    if ((flags & 1u) && mts.mti > kMtN) why = "chemistry stream index out of range";
    if (!why.empty()) return fail(c, MDPD_ERR_PARSE, "corrupt checkpoint (" + why + "): " + path);
  }
  DevState& d = c->d;
  double* const dd[4] = {d.pos, d.vel, d.vbuf, d.delta};
  for (int q = 0; q < 4; ++q)
    for (int a = 0; a < 3; ++a)
      HIPCHK(c, hipMemcpy(dd[q] + (size_t)a * d.cap, f64.data() + ((size_t)q * 3 + a) * n,
                          8 * (size_t)n, hipMemcpyHostToDevice));
  float* const df[2] = {d.acc, d.press};
  for (int q = 0; q < 2; ++q)
    for (int a = 0; a < 3; ++a)
      HIPCHK(c, hipMemcpy(df[q] + (size_t)a * d.cap, f32.data() + ((size_t)q * 3 + a) * n,
                          4 * (size_t)n, hipMemcpyHostToDevice));
  uint32_t* const du[6] = {d.id, d.prop, d.amp_id, d.unit, d.orig, d.chain};
  for (int q = 0; q < 6; ++q)
    HIPCHK(c, hipMemcpy(du[q], u32.data() + (size_t)q * n, 4 * (size_t)n, hipMemcpyHostToDevice));
  if (na) {
    HIPCHK(c, hipMemcpy(d.topo, u32.data() + 6 * (size_t)n, 16 * (size_t)na, hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(d.cf_bend, cf.data(), 8 * (size_t)na, hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(d.amp_rep, u32.data() + 6 * (size_t)n + 4 * (size_t)na, 4 * (size_t)na,
                        hipMemcpyHostToDevice));
  }
  // host mirrors: bead and chain counts, seeds, clock
  c->amp_num_rep = anum;
  c->sol_num_rep = snum;
  c->amp_off.assign(nr + 1, 0);
  for (uint32_t r = 0; r < nr; ++r) c->amp_off[r + 1] = c->amp_off[r] + anum[r];
  for (uint32_t r = 0; r < nr; ++r) {
    c->R[r].pair_seed = seed[r];
    c->P[r].pair_seed = seed[r];
    c->R[r].chem_seed = cseed[r];
    c->P[r].chem_seed = cseed[r];
  }
  if (nr == 1) {
    c->rep_off[1] = n;
    c->R[0].n_beads = n;
  }
  HIPCHK(c, hipMemcpy(d.reps, c->R.data(), sizeof(Replica) * nr, hipMemcpyHostToDevice));
  HIPCHK(c, hipMemcpy(d.rep_off, c->rep_off.data(), 4 * (size_t)(nr + 1), hipMemcpyHostToDevice));
  HIPCHK(c, hipMemcpy(d.amp_off, c->amp_off.data(), 4 * (size_t)(nr + 1), hipMemcpyHostToDevice));
  HIPCHK(c, hipMemcpy(c->d_amp_num, anum.data(), 4 * (size_t)nr, hipMemcpyHostToDevice));
  invalidate_graphs(c);
  d.n = n;
  d.n_amp = na;
  c->n_amp = na;
  set_launch_bound(c);
  c->time = head[4];
  HIPCHK(c, hipMemcpy(d.d_time, &c->time, 4, hipMemcpyHostToDevice));
  c->drifted = false;
  c->have_cells = c->cells_fresh = c->have_density = false;
  // the reference-MT chemistry stream where the saving context had it on
  // (a file without one leaves the context's draw mode as it is)
  if (flags & 1u) {
    const int rc = chem_mt_enable(c, mts);
    if (rc) return rc;
  }
  return MDPD_OK;
}

// ---- observables beyond the Observer (k_obs.hip)
namespace {
std::vector<double> sq_edges(int nb, double width) {
  std::vector<double> e(nb + 1);
  for (int k = 0; k <= nb; ++k) {
    const double x = k * width;
    e[k] = x * x;
  }
  return e;
}
}  // namespace

int mdpd_rdf(mdpd_ctx* c, double rmax, int32_t nbins, double* out) {
  if (!c || !out || !(rmax > 0.0 && rmax <= 1.0) || nbins < 1 || nbins > 1024)
    return fail(c, MDPD_ERR_ARG, "rdf: 0 < rmax <= 1 and 1 <= nbins <= 1024");
  if (!c->have_cells || !c->cells_fresh) return fail(c, MDPD_ERR_STATE, "no current cell list");
  const std::vector<double> e2 = sq_edges(nbins, rmax / nbins);
  const size_t cnt = (size_t)c->n_rep * nbins;
  double* de = nullptr;
  unsigned long long* h = nullptr;
  HIPCHK(c, hipMalloc(&de, 8 * e2.size()));
  HIPCHK(c, hipMalloc(&h, 8 * cnt));
  HIPCHK(c, hipMemcpyAsync(de, e2.data(), 8 * e2.size(), hipMemcpyHostToDevice, c->st));
  HIPCHK(c, hipMemsetAsync(h, 0, 8 * cnt, c->st));
  launch_rdf(c->d, de, nbins, rmax, h, c->st);
  std::vector<unsigned long long> hv(cnt);
  HIPCHK(c, hipMemcpyAsync(hv.data(), h, 8 * cnt, hipMemcpyDeviceToHost, c->st));
  HIPCHK(c, hipStreamSynchronize(c->st));
  hipFree(de);
  hipFree(h);
  for (size_t k = 0; k < cnt; ++k) out[k] = (double)hv[k];
  return MDPD_OK;
}

int mdpd_chain_histograms(mdpd_ctx* c, int32_t nbins, double bond_max, double* out) {
  if (!c || !out || nbins < 1 || !(bond_max > 0.0))
    return fail(c, MDPD_ERR_ARG, "chain histograms: nbins >= 1 and bond_max > 0");
  if (c->drifted) return fail(c, MDPD_ERR_STATE, "chain histograms between drift and kick");
  const std::vector<double> eb2 = sq_edges(nbins, bond_max / nbins);
  const std::vector<double> ee2 = sq_edges(nbins, 3.0 * bond_max / nbins);
  const size_t cnt = (size_t)c->n_rep * 4 * nbins;
  double* de = nullptr;
  unsigned long long* h = nullptr;
  HIPCHK(c, hipMalloc(&de, 16 * eb2.size()));
  HIPCHK(c, hipMalloc(&h, 8 * cnt));
  HIPCHK(c, hipMemcpyAsync(de, eb2.data(), 8 * eb2.size(), hipMemcpyHostToDevice, c->st));
  HIPCHK(c, hipMemcpyAsync(de + eb2.size(), ee2.data(), 8 * ee2.size(), hipMemcpyHostToDevice,
                           c->st));
  HIPCHK(c, hipMemsetAsync(h, 0, 8 * cnt, c->st));
  // the step keeps only the per-bead chain map current
  launch_topo_from_chain(c->d, c->st);
  launch_chain_hist(c->d, de, de + eb2.size(), nbins, h, c->st);
  std::vector<unsigned long long> hv(cnt);
  HIPCHK(c, hipMemcpyAsync(hv.data(), h, 8 * cnt, hipMemcpyDeviceToHost, c->st));
  HIPCHK(c, hipStreamSynchronize(c->st));
  hipFree(de);
  hipFree(h);
  for (size_t k = 0; k < cnt; ++k) out[k] = (double)hv[k];
  return MDPD_OK;
}

int mdpd_count_map(mdpd_ctx* c, int32_t gx, int32_t gy, int32_t gz, uint32_t type_mask,
                   double* out) {
  if (!c || !out || gx < 1 || gy < 1 || gz < 1 || (int64_t)gx * gy * gz > (1 << 26))
    return fail(c, MDPD_ERR_ARG, "count map: 1 <= g per axis, at most 2^26 voxels");
  const size_t cnt = (size_t)c->n_rep * gx * gy * gz;
  unsigned int* h = nullptr;
  HIPCHK(c, hipMalloc(&h, 4 * cnt));
  HIPCHK(c, hipMemsetAsync(h, 0, 4 * cnt, c->st));
  launch_count_map(c->d, gx, gy, gz, type_mask, h, c->st);
  std::vector<unsigned int> hv(cnt);
  HIPCHK(c, hipMemcpyAsync(hv.data(), h, 4 * cnt, hipMemcpyDeviceToHost, c->st));
  HIPCHK(c, hipStreamSynchronize(c->st));
  hipFree(h);
  for (size_t k = 0; k < cnt; ++k) out[k] = (double)hv[k];
  return MDPD_OK;
}

// Times every (block_scan, block_force) pair on the current state for
// `nsteps` steps and keeps the fastest.  The state is restored after every
// candidate; per-bead sums run in a fixed order, so every candidate gives
// the same bits.  With MDPD_TUNE_SKIN in flags it then also times the
// neighbour-list skins 0.1, 0.2 and 0.3 (cells of 1 + skin; those the boxes
// allow) against the every-step cell list, with the best block sizes.
// With MDPD_TUNE_CACHE set, results are kept in that file.
int mdpd_autotune_ex(mdpd_ctx* c, uint32_t nsteps, uint32_t flags, mdpd_tuning* out) {
  if (!c) return MDPD_ERR_ARG;
  if (!c->d.n) return fail(c, MDPD_ERR_STATE, "no beads uploaded");
  if (c->drifted) return fail(c, MDPD_ERR_STATE, "autotune between drift and kick");
  if (flags & ~(uint32_t)MDPD_TUNE_SKIN) return fail(c, MDPD_ERR_ARG, "unknown autotune flags");
  DevState& d = c->d;
  char keybuf[256];
  snprintf(keybuf, sizeof(keybuf),
           "n=%u,ncell=%u,nrep=%d,ns=%d,nd=%d,env=%d,namp=%u,W=%d,skin=%g,flags=%u", d.n, d.ncell,
           c->n_rep, c->ns, c->nd, c->env, d.n_amp, d.W, (double)c->tune.verlet_skin, flags);
  const std::string key(keybuf);
  const char* cache = std::getenv("MDPD_TUNE_CACHE");
  if (cache) {
    std::ifstream f(cache);
    std::string line;
    while (std::getline(f, line)) {
      std::istringstream ss(line);
      std::string k;
      int bs = 0, bf = 0;
      float sk = 0.0f;
      if ((ss >> k >> bs >> bf >> sk) && k == key) {
        mdpd_tuning t = c->tune;
        t.block_scan = bs;
        t.block_force = bf;
        t.verlet_skin = sk;
        int rc = mdpd_set_tuning(c, &t);
        if (rc) return rc;
        if (out) *out = c->tune;
        return MDPD_OK;
      }
    }
  }
  if (nsteps == 0) nsteps = 1;
  // snapshot of the evolving state
  const size_t n3 = 3 * (size_t)d.cap;
  double* snap = nullptr;
  float* snapf = nullptr;
  uint32_t* snapu = nullptr;
  HIPCHK(c, hipMalloc(&snap, 4 * n3 * sizeof(double)));
  HIPCHK(c, hipMalloc(&snapf, 2 * n3 * sizeof(float)));
  HIPCHK(c, hipMalloc(&snapu, (6 * (size_t)d.cap + 4 * (size_t)c->amp_cap) * sizeof(uint32_t)));
  const uint32_t t0 = c->time;
  // (d is re-read through c->d inside: the storage parity may flip)
  auto save = [&]() -> int {
    DevState& e = c->d;
    HIPCHK(c, hipMemcpyAsync(snap, e.pos, n3 * 8, hipMemcpyDeviceToDevice, c->st));
    HIPCHK(c, hipMemcpyAsync(snap + n3, e.vel, n3 * 8, hipMemcpyDeviceToDevice, c->st));
    HIPCHK(c, hipMemcpyAsync(snap + 2 * n3, e.vbuf, n3 * 8, hipMemcpyDeviceToDevice, c->st));
    HIPCHK(c, hipMemcpyAsync(snap + 3 * n3, e.delta, n3 * 8, hipMemcpyDeviceToDevice, c->st));
    HIPCHK(c, hipMemcpyAsync(snapf, e.acc, n3 * 4, hipMemcpyDeviceToDevice, c->st));
    HIPCHK(c, hipMemcpyAsync(snapf + n3, e.press, n3 * 4, hipMemcpyDeviceToDevice, c->st));
    {
      // every per-slot array the storage permutation moves, chain included
      uint32_t* const src[6] = {e.prop, e.id, e.amp_id, e.unit, e.orig, e.chain};
      for (int q = 0; q < 6; ++q)
        HIPCHK(c, hipMemcpyAsync(snapu + (size_t)q * e.cap, src[q], (size_t)e.cap * 4, hipMemcpyDeviceToDevice, c->st));
      HIPCHK(c, hipMemcpyAsync(snapu + 6 * (size_t)e.cap, e.topo, 16 * (size_t)c->amp_cap, hipMemcpyDeviceToDevice, c->st));
    }
    return MDPD_OK;
  };
  auto restore = [&]() -> int {
    DevState& e = c->d;
    HIPCHK(c, hipMemcpyAsync(e.pos, snap, n3 * 8, hipMemcpyDeviceToDevice, c->st));
    HIPCHK(c, hipMemcpyAsync(e.vel, snap + n3, n3 * 8, hipMemcpyDeviceToDevice, c->st));
    HIPCHK(c, hipMemcpyAsync(e.vbuf, snap + 2 * n3, n3 * 8, hipMemcpyDeviceToDevice, c->st));
    HIPCHK(c, hipMemcpyAsync(e.delta, snap + 3 * n3, n3 * 8, hipMemcpyDeviceToDevice, c->st));
    HIPCHK(c, hipMemcpyAsync(e.acc, snapf, n3 * 4, hipMemcpyDeviceToDevice, c->st));
    HIPCHK(c, hipMemcpyAsync(e.press, snapf + n3, n3 * 4, hipMemcpyDeviceToDevice, c->st));
    {
      uint32_t* const dst[6] = {e.prop, e.id, e.amp_id, e.unit, e.orig, e.chain};
      for (int q = 0; q < 6; ++q)
        HIPCHK(c, hipMemcpyAsync(dst[q], snapu + (size_t)q * e.cap, (size_t)e.cap * 4, hipMemcpyDeviceToDevice, c->st));
      HIPCHK(c, hipMemcpyAsync(e.topo, snapu + 6 * (size_t)e.cap, 16 * (size_t)c->amp_cap, hipMemcpyDeviceToDevice, c->st));
    }
    // the cell list and pair list of the trial steps are stale now
    c->cells_fresh = false;
    c->skin_valid = false;
    c->have_density = false;
    c->time = t0;
    HIPCHK(c, hipMemcpyAsync(e.d_time, &c->time, 4, hipMemcpyHostToDevice, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    return MDPD_OK;
  };
  int rc = save();
  if (rc) return rc;
  hipEvent_t e0, e1;
  hipEventCreate(&e0);
  hipEventCreate(&e1);
  const bool prof = c->profiling;
  c->profiling = false;
  // one candidate: MDPD_OK and its time in *ms, or a status (set_tuning
  // refusals of a skin the boxes do not allow come back as MDPD_ERR_ARG)
  auto trial = [&](const mdpd_tuning& t, float* ms) -> int {
    int r;
    if ((r = mdpd_set_tuning(c, &t))) return r;
    if ((r = run_steps(c, 1))) return r;  // warm-up (graph capture)
    if ((r = restore())) return r;
    hipEventRecord(e0, c->st);
    if ((r = run_steps(c, nsteps))) return r;
    hipEventRecord(e1, c->st);
    hipEventSynchronize(e1);
    hipEventElapsedTime(ms, e0, e1);
    if ((r = check_flags(c))) return r;
    return restore();
  };
  const int sizes[3] = {64, 128, 256};
  float best = 1e30f;
  mdpd_tuning bt = c->tune;
  bt.verlet_skin = 0.0f;
  for (int bs : sizes) {
    for (int bf : sizes) {
      mdpd_tuning t = c->tune;
      t.block_scan = bs;
      t.block_force = bf;
      if (flags & MDPD_TUNE_SKIN) t.verlet_skin = 0.0f;
      float ms = 0.f;
      if ((rc = trial(t, &ms))) break;
      if (ms < best) {
        best = ms;
        bt = t;
      }
    }
    if (rc) break;
  }
  if (!rc && (flags & MDPD_TUNE_SKIN)) {
    const float skins[3] = {0.1f, 0.2f, 0.3f};
    for (float sk : skins) {
      mdpd_tuning t = bt;
      t.verlet_skin = sk;
      float ms = 0.f;
      const int r = trial(t, &ms);
      if (r == MDPD_ERR_ARG) continue;  // the boxes hold too few cells of 1 + sk
      if ((rc = r)) break;
      if (ms < best) {
        best = ms;
        bt = t;
      }
    }
  }
  c->profiling = prof;
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  hipFree(snap);
  hipFree(snapf);
  hipFree(snapu);
  if (rc) return rc;
  rc = mdpd_set_tuning(c, &bt);
  if (rc) return rc;
  c->have_cells = c->cells_fresh = c->have_density = false;
  if (cache) {
    std::ofstream f(cache, std::ios::app);
    f << key << " " << bt.block_scan << " " << bt.block_force << " " << bt.verlet_skin << "\n";
  }
  if (out) *out = c->tune;
  return MDPD_OK;
}

int mdpd_autotune(mdpd_ctx* c, uint32_t nsteps, mdpd_tuning* out) {
  return mdpd_autotune_ex(c, nsteps, 0u, out);
}

}  // extern "C"
