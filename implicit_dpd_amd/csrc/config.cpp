// Initial configurations, host side: config_param.txt input, the ring
// layout of the sphere and cylinder modes, the reference's validity checks,
// and the C ABI (mdpd_read_config_param, mdpd_config_count,
// mdpd_make_config_host, mdpd_make_config).  The beads themselves are made
// on the device by k_config.hip.
//
// Follows src/config_maker.cpp: MatchingTagValues (:386-418), GenConfig
// (:342-384) with MakeFlatSheet (:103-138), MakeSphSheet (:178-213),
// MakeCylindSheet (:229-262), MakeRandomConfig (:264-280), the solvent
// makers (:282-321), GenVeloc (:60-72) and CheckParticleConfigIsValid
// (:420-452).
#include <hip/hip_runtime.h>
#include <strings.h>

#include <algorithm>
#include <cmath>
#include <fstream>
#include <string>
#include <vector>

#include "config.hpp"
#include "ctx.hpp"
#include "tags.hpp"

using namespace mdpd;

namespace {

struct Plan {
  CfgLaunch L{};
  std::vector<CfgRing> rings;
  uint32_t n = 0, amp_beads = 0;
  double temperature = 0.0;
};

// appends the rings of one MakeLineForEachTheta / MakeLineForEachAxis pass;
// `placed` counts amphiphiles laid so far, the last ring is cut at amp_num
struct RingBuilder {
  const mdpd_config_spec& s;
  std::vector<CfgRing>& out;
  uint64_t placed = 0;
  bool overshoot = false;

  // MakeSphLine :139-161 without the placement: the ring's element count
  // phi_elem = (int)(2 pi / (lip_len / (rad sin the)))
  bool ring(double the, double rad, double sign, double cz, std::string& err) {
    const double sin_t = std::sin(the), cos_t = std::cos(the);
    const double prj_rad = rad * sin_t;
    double d_phi = s.lip_len / prj_rad;
    const double q = 2.0 * M_PI / d_phi;
    if (!(q < 2147483647.0)) {
      err = "ring holds more than 2^31 amphiphiles (lip_len too small)";
      return false;
    }
    const int32_t phi_elem = q > 0.0 ? (int32_t)q : 0;  // the loop runs for phi_elem > 0 only
    if (phi_elem <= 0) return true;
    d_phi = 2.0 * M_PI / phi_elem;
    if (placed >= s.amp_num) {
      // only reachable in the axis passes, which have no guard on entry:
      // the reference lays one amphiphile past the end (then fails)
      overshoot = true;
      return true;
    }
    const uint32_t cnt = (uint32_t)std::min<uint64_t>((uint64_t)phi_elem, s.amp_num - placed);
    CfgRing r{sin_t, cos_t, rad, sign, d_phi, cz, (uint32_t)placed, cnt};
    out.push_back(r);
    placed += cnt;
    return true;
  }
  bool full() const { return overshoot || placed >= s.amp_num; }
};

int plan_config(const mdpd_config_spec* s, Plan& P, std::string& err) {
  if (!s) {
    err = "null argument";
    return MDPD_ERR_ARG;
  }
  if ((uint64_t)s->head_unit + s->tail_unit > kIndexMask) {
    err = "head_unit + tail_unit is too large";
    return MDPD_ERR_ARG;
  }
  const uint32_t all_unit = s->head_unit + s->tail_unit;
  const uint64_t amp_beads = (uint64_t)all_unit * s->amp_num;
  const uint64_t n = amp_beads + (s->with_solvent ? s->sol_num : 0u);
  if (n > kIndexMask) {
    err = "more than " + std::to_string(kIndexMask) + " beads";
    return MDPD_ERR_ARG;
  }
  if (n == 0) {
    err = "the configuration holds no beads (amp_num and sol_num are 0)";
    return MDPD_ERR_ARG;
  }
  if (s->amp_num && all_unit == 0) {
    err = "head_unit + tail_unit must be positive";
    return MDPD_ERR_ARG;
  }
  for (int a = 0; a < 3; ++a)
    if (!(s->box[a] > 0.0) || !std::isfinite(s->box[a])) {
      err = "box_leng must be positive";
      return MDPD_ERR_ARG;
    }
  if (!(s->temperature >= 0.0) || !std::isfinite(s->temperature)) {
    err = "Temperature must be finite and >= 0";
    return MDPD_ERR_ARG;
  }
  if (s->mode < MDPD_CFG_FLAT || s->mode > MDPD_CFG_RANDOM) {
    err = "Unknown mode";
    return MDPD_ERR_ARG;
  }
  if (s->with_solvent && (s->mode == MDPD_CFG_RANDOM || s->mode == MDPD_CFG_ASYM_FLAT)) {
    err = "with_solvent is not supported in this mode.";
    return MDPD_ERR_CONFIG;
  }
  CfgLaunch& L = P.L;
  L.mode = s->mode;
  L.n = (uint32_t)n;
  L.amp_num = s->amp_num;
  L.all_unit = all_unit;
  L.head_unit = s->head_unit;
  L.seed = s->seed;
  for (int a = 0; a < 3; ++a) L.box[a] = s->box[a];
  L.bond_leng = s->bond_leng;
  L.q_thick = 0.5 * all_unit * s->bond_leng;
  L.sd = std::sqrt(s->temperature);
  L.sol_kind = kSolNone;
  P.n = (uint32_t)n;
  P.amp_beads = (uint32_t)amp_beads;
  P.temperature = s->temperature;
  const double eps = 1.0e-5;  // MakeFlatSheet :117
  RingBuilder rb{*s, P.rings};
  switch (s->mode) {
    case MDPD_CFG_FLAT:
    case MDPD_CFG_ASYM_FLAT:
      L.span = (all_unit - 1) * s->bond_leng + eps;
      if (s->mode == MDPD_CFG_ASYM_FLAT) L.rat = (uint32_t)(100 * s->in_out_rat);
      if (s->with_solvent) {  // lower half of y (GenConfig :345-350)
        L.sol_kind = kSolCuboid;
        for (int a = 0; a < 3; ++a) {
          L.sol_lo[a] = 0.0;
          L.sol_hi[a] = s->box[a];
        }
        L.sol_hi[1] *= 0.5;
      }
      break;
    case MDPD_CFG_SPHERE: {
      const double min_box = std::min(std::min(s->box[0], s->box[1]), s->box[2]);
      if (!(s->sph_rad >= 0.0 && s->sph_rad <= 0.5 * min_box)) {
        err = "sphere radius is too large.";
        return MDPD_ERR_CONFIG;
      }
      const double up_the = s->upper_the * M_PI / 180.0;  // degrees
      const double q = L.q_thick;
      const double out_rad = s->sph_rad + q, in_rad = s->sph_rad - q;
      double d_the_out = s->lip_len / out_rad, d_the_in = s->lip_len / in_rad;
      const double qo = up_the / d_the_out, qi = up_the / d_the_in;
      const int32_t out_elem = qo > 0.0 && qo < 2147483647.0 ? (int32_t)qo : 0;
      const int32_t in_elem = qi > 0.0 && qi < 2147483647.0 ? (int32_t)qi : 0;
      d_the_out = up_the / out_elem;
      d_the_in = up_the / in_elem;
      const double cz = 0.5 * s->box[2];
      // MakeLineForEachTheta: returns at once when full, stops at the ring
      // that fills the last amphiphile
      auto pass = [&](int32_t elem, double d_the, double off, double rad, double sign) {
        for (int32_t i = 0; i < elem && !rb.full(); ++i)
          if (!rb.ring(d_the * (i + off), rad, sign, cz, err)) return false;
        return true;
      };
      if (!pass(out_elem, d_the_out, 0.0, out_rad, -1.0)) return MDPD_ERR_CONFIG;
      if (!pass(in_elem, d_the_in, 0.0, in_rad, 1.0)) return MDPD_ERR_CONFIG;
      if (!rb.full() && !pass(in_elem, d_the_in, 0.5, in_rad, 1.0)) return MDPD_ERR_CONFIG;
      if (!rb.full()) {
        err = "Too many particles are specified.";
        return MDPD_ERR_CONFIG;
      }
      if (s->with_solvent) {
        L.sol_kind = kSolSphere;
        L.sol_rad = s->sph_rad - 0.5 * all_unit * s->bond_leng;
      }
      break;
    }
    case MDPD_CFG_CYLIND: {
      const int axis = 2;
      if (!(s->cyl_l >= 0.0 && s->cyl_l <= s->box[axis])) {
        err = "cyl_l must lie in [0, box_leng.z]";
        return MDPD_ERR_CONFIG;
      }
      if (!(s->cyl_r >= 0.0 && (2.0 * s->cyl_r + all_unit * s->bond_leng) < s->box[axis])) {
        err = "cyl_r must satisfy 2 cyl_r + all_unit bond_leng < box_leng.z";
        return MDPD_ERR_CONFIG;
      }
      const double q = L.q_thick;
      const double out_rad = s->cyl_r + q, in_rad = s->cyl_r - q;
      const double d_the_in = s->lip_len / in_rad;
      const double qi = M_PI / d_the_in;
      const int32_t in_elem = qi > 0.0 && qi < 2147483647.0 ? (int32_t)qi : 0;
      const double qz = s->box[axis] / s->lip_len;
      const int32_t z_elem = qz > 0.0 && qz < 2147483647.0 ? (int32_t)qz : 0;
      const double h_pi = std::acos(0.0);
      // MakeLineForEachAxis: no guard on entry, stops at the ring that
      // fills the last amphiphile
      auto pass = [&](int32_t elem, double rad, double sign) {
        for (int32_t i = 0; i < elem; ++i) {
          const bool was_full = rb.full();
          if (!rb.ring(h_pi, rad, sign, i * s->lip_len, err)) return false;
          if (was_full && rb.overshoot) return true;
          if (rb.full()) return true;
        }
        return true;
      };
      if (!pass(z_elem, out_rad, -1.0)) return MDPD_ERR_CONFIG;
      if (!pass(z_elem, in_rad, 1.0)) return MDPD_ERR_CONFIG;
      // the fill pass (residue > 0) reuses the outer radius and in_the_elem
      // rings, as the reference does (:254)
      if ((rb.overshoot || !rb.full()) && !pass(in_elem, out_rad, -1.0)) return MDPD_ERR_CONFIG;
      if (rb.overshoot || !rb.full()) {
        err = "Too many particles are specified.";
        return MDPD_ERR_CONFIG;
      }
      if (s->with_solvent) {
        L.sol_kind = kSolCylinder;
        L.sol_axis = axis;
        L.sol_rad = s->cyl_r - 0.5 * all_unit * s->bond_leng;
      }
      break;
    }
    case MDPD_CFG_RANDOM:
      break;
  }
  L.center[0] = 0.5 * s->box[0];
  L.center[1] = 0.5 * s->box[1];
  L.n_rings = (uint32_t)P.rings.size();
  if ((s->mode == MDPD_CFG_SPHERE || s->mode == MDPD_CFG_CYLIND) && s->amp_num && P.rings.empty()) {
    err = "Too many particles are specified.";
    return MDPD_ERR_CONFIG;
  }
  return MDPD_OK;
}

// Runs the kernels for plan P into `out` on stream st; device scratch is
// allocated here.  Returns the status and the kinetic temperature.
int generate(Plan& P, const CfgOut& out, hipStream_t st, double* kin_temp, std::string& err) {
  CfgRing* d_rings = nullptr;
  uint32_t* d_flags = nullptr;
  double* d_red = nullptr;
  int rc = MDPD_OK;
  auto hip = [&](hipError_t e, const char* what) {
    if (e != hipSuccess && rc == MDPD_OK) {
      rc = MDPD_ERR_HIP;
      err = std::string(what) + ": " + hipGetErrorString(e);
    }
    return e == hipSuccess;
  };
  uint32_t flags = 0;
  double red[4] = {0, 0, 0, 0};
  const size_t rbytes = std::max<size_t>(P.rings.size(), 1) * sizeof(CfgRing);
  if (hip(hipMalloc(&d_rings, rbytes), "hipMalloc") && hip(hipMalloc(&d_flags, 4), "hipMalloc") &&
      hip(hipMalloc(&d_red, cfg_scratch_doubles(P.n) * sizeof(double)), "hipMalloc") &&
      hip(hipMemsetAsync(d_flags, 0, 4, st), "hipMemsetAsync") &&
      (P.rings.empty() ||
       hip(hipMemcpyAsync(d_rings, P.rings.data(), P.rings.size() * sizeof(CfgRing),
                          hipMemcpyHostToDevice, st), "hipMemcpyAsync"))) {
    P.L.rings = d_rings;
    P.L.flags = d_flags;
    P.L.out = out;
    launch_cfg_generate(P.L, d_red, st);
    hip(hipGetLastError(), "config kernels");
    hip(hipMemcpyAsync(&flags, d_flags, 4, hipMemcpyDeviceToHost, st), "hipMemcpyAsync");
    hip(hipMemcpyAsync(red, d_red, 4 * sizeof(double), hipMemcpyDeviceToHost, st), "hipMemcpyAsync");
    hip(hipStreamSynchronize(st), "hipStreamSynchronize");
  }
  if (d_rings) (void)hipFree(d_rings);
  if (d_flags) (void)hipFree(d_flags);
  if (d_red) (void)hipFree(d_red);
  if (rc) return rc;
  // CheckParticleConfigIsValid :420-452
  if (flags & kCfgFlagSolvent) {
    err = "no solvent position inside the region after " + std::to_string(kSolMaxAttempts) +
          " draws (is the region empty?)";
    return MDPD_ERR_CONFIG;
  }
  if (flags & kCfgFlagOutOfBox) {
    err = "There is a particle in outside range.";
    return MDPD_ERR_CONFIG;
  }
  const double kt = red[3] / (3.0 * P.n);
  if (kin_temp) *kin_temp = kt;
  if (std::fabs(kt - P.temperature) >= 1.0e-2) {
    err = "Please check kinetic temperature. Kinetic temperature is " + std::to_string(kt) + "K_BT.";
    return MDPD_ERR_CONFIG;
  }
  return MDPD_OK;
}

}  // namespace

extern "C" {

int mdpd_read_config_param(const char* dir, mdpd_config_spec* s) {
  if (!dir || !s) return fail(nullptr, MDPD_ERR_ARG, "null argument");
  const std::string fname = std::string(dir) + "/config_param.txt";
  std::ifstream fin(fname.c_str());
  if (!fin) return fail(nullptr, MDPD_ERR_PARSE, "Cannot find file " + fname);
  TagMap tv;
  read_tags(fin, tv);
  Parse P{tv, ""};
  std::memset(s, 0, sizeof(*s));
  bool ok = true;
  ok &= P.f64("Temperature", &s->temperature, 1);
  ok &= P.f64("box_leng", s->box, 3);
  ok &= P.u32("amp_num", &s->amp_num, 1);
  std::string mode;
  if (P.need("mode", 1)) mode = tv["mode"][0];
  else ok = false;
  ok &= P.u32("head_unit", &s->head_unit, 1);
  ok &= P.u32("tail_unit", &s->tail_unit, 1);
  ok &= P.f64("bond_leng", &s->bond_leng, 1);
  s->with_solvent = tv.count("with_solvent") ? 1 : 0;
  auto eq = [&](const char* m) { return strcasecmp(mode.c_str(), m) == 0; };
  if (eq("flat")) s->mode = MDPD_CFG_FLAT;
  else if (eq("asym_flat")) s->mode = MDPD_CFG_ASYM_FLAT;
  else if (eq("sphere")) s->mode = MDPD_CFG_SPHERE;
  else if (eq("cylind")) s->mode = MDPD_CFG_CYLIND;
  else if (eq("random")) s->mode = MDPD_CFG_RANDOM;
  else if (ok) return fail(nullptr, MDPD_ERR_PARSE, mode.empty() ? "Mode is not specified" : mode + ": Unknown mode");
  if (ok && s->mode == MDPD_CFG_SPHERE) {
    ok &= P.f64("sph_rad", &s->sph_rad, 1);
    ok &= P.f64("upper_the", &s->upper_the, 1);
  }
  if (ok && s->mode == MDPD_CFG_CYLIND) {
    ok &= P.f64("cyl_l", &s->cyl_l, 1);
    ok &= P.f64("cyl_r", &s->cyl_r, 1);
  }
  if (ok && s->mode == MDPD_CFG_ASYM_FLAT) ok &= P.f64("in_out_rat", &s->in_out_rat, 1);
  if (ok && s->with_solvent) ok &= P.u32("sol_num", &s->sol_num, 1);
  ok &= P.f64("lip_len", &s->lip_len, 1);
  if (!ok) return fail(nullptr, MDPD_ERR_PARSE, P.err);
  return MDPD_OK;
}

int mdpd_config_count(const mdpd_config_spec* s, uint32_t* n_beads) {
  if (!s || !n_beads) return fail(nullptr, MDPD_ERR_ARG, "null argument");
  Plan P;
  std::string err;
  const int rc = plan_config(s, P, err);
  if (rc) return fail(nullptr, rc, err);
  *n_beads = P.n;
  return MDPD_OK;
}

int mdpd_make_config_host(const mdpd_config_spec* s, uint32_t n, double* pos, double* vel,
                          uint32_t* id, uint32_t* prop, uint32_t* amp_id, uint32_t* unit,
                          double* kin_temp) {
  if (!pos || !vel || !id || !prop || !amp_id || !unit) return fail(nullptr, MDPD_ERR_ARG, "null argument");
  Plan P;
  std::string err;
  int rc = plan_config(s, P, err);
  if (rc) return fail(nullptr, rc, err);
  if (n != P.n) return fail(nullptr, MDPD_ERR_ARG, "n must be " + std::to_string(P.n) + " (mdpd_config_count)");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail(nullptr, MDPD_ERR_NO_DEVICE, "no HIP device");
  hipStream_t st = nullptr;
  double* dpv = nullptr;
  uint32_t* du = nullptr;
  if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess)
    return fail(nullptr, MDPD_ERR_HIP, "hipStreamCreate failed");
  const size_t nn = n;
  if (hipMalloc(&dpv, 6 * nn * sizeof(double)) != hipSuccess || hipMalloc(&du, 4 * nn * sizeof(uint32_t)) != hipSuccess) {
    rc = MDPD_ERR_HIP;
    err = "hipMalloc failed";
  } else {
    // AoS [n][3] like the host arrays
    CfgOut o{dpv, dpv + 3 * nn, nullptr, du, du + nn, du + 2 * nn, du + 3 * nn, 3, 1};
    rc = generate(P, o, st, kin_temp, err);
    if (rc == MDPD_OK) {
      if (hipMemcpy(pos, dpv, 3 * nn * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess ||
          hipMemcpy(vel, dpv + 3 * nn, 3 * nn * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess ||
          hipMemcpy(id, du, nn * 4, hipMemcpyDeviceToHost) != hipSuccess ||
          hipMemcpy(prop, du + nn, nn * 4, hipMemcpyDeviceToHost) != hipSuccess ||
          hipMemcpy(amp_id, du + 2 * nn, nn * 4, hipMemcpyDeviceToHost) != hipSuccess ||
          hipMemcpy(unit, du + 3 * nn, nn * 4, hipMemcpyDeviceToHost) != hipSuccess) {
        rc = MDPD_ERR_HIP;
        err = "hipMemcpy failed";
      }
    }
  }
  if (dpv) (void)hipFree(dpv);
  if (du) (void)hipFree(du);
  (void)hipStreamDestroy(st);
  if (rc) return fail(nullptr, rc, err);
  return MDPD_OK;
}

int mdpd_make_config(mdpd_ctx* c, const mdpd_config_spec* s, double* kin_temp) {
  if (!c || !s) return fail(c, MDPD_ERR_ARG, "null argument");
  if (c->n_rep != 1) return fail(c, MDPD_ERR_UNSUPPORTED, "mdpd_make_config needs a single-system context");
  Plan P;
  std::string err;
  int rc = plan_config(s, P, err);
  if (rc) return fail(c, rc, err);
  for (int a = 0; a < 3; ++a)
    if (s->box[a] != c->P[0].box[a]) return fail(c, MDPD_ERR_ARG, "box_leng differs from the context's box");
  if (P.n > c->cap) return fail(c, MDPD_ERR_CAPACITY, "more beads than the capacity");
  if (s->amp_num && s->head_unit + s->tail_unit != 4)
    return fail(c, MDPD_ERR_UNSUPPORTED, "the engine's chains are 4-bead amphiphiles (head_unit + tail_unit = 4)");
  if (s->amp_num > c->amp_cap) return fail(c, MDPD_ERR_CAPACITY, "too many amphiphiles");
  c->skin_valid = false;
  DevState& d = c->d;
  const size_t cap = d.cap;
  CfgOut o{d.pos, d.vel, d.vbuf, d.id, d.prop, d.amp_id, d.unit, 1, cap};
  rc = generate(P, o, c->st, kin_temp, err);
  if (rc) {
    // the kernels may have overwritten the old beads: leave no beads rather
    // than a mix (the reference's config_maker exits here)
    invalidate_graphs(c);
    d.n = 0;
    d.n_amp = 0;
    c->n_amp = 0;
    set_launch_bound(c);
    c->rep_off[1] = 0;
    c->R[0].n_beads = 0;
    c->have_cells = c->cells_fresh = c->have_density = false;
    return fail(c, rc, err);
  }
  const uint32_t n = P.n;
  for (int a = 0; a < 3; ++a) {
    HIPCHK(c, hipMemsetAsync(d.delta + a * cap, 0, 8 * (size_t)n, c->st));
    HIPCHK(c, hipMemsetAsync(d.acc + a * cap, 0, 4 * (size_t)n, c->st));
  }
  const uint32_t na = s->amp_num;
  launch_cfg_chains(n, na, c->P[0].cf_b, d.chain, d.orig, d.topo, d.cf_bend, d.amp_rep, c->st);
  HIPCHK(c, hipGetLastError());
  // the bookkeeping of an upload (mdpd_upload_soa) and of a topology with
  // amp_num chains and no core amphiphiles (mdpd_set_topology)
  c->sol_num_rep[0] = n - P.amp_beads;
  c->amp_num_rep[0] = na;
  c->amp_off.assign(2, 0);
  c->amp_off[1] = na;
  if (n != d.n || na != d.n_amp) invalidate_graphs(c);
  d.n = n;
  d.n_amp = na;
  c->n_amp = na;
  set_launch_bound(c);
  c->rep_off[1] = n;
  c->R[0].n_beads = n;
  HIPCHK(c, hipMemcpyAsync(d.rep_off, c->rep_off.data(), sizeof(uint32_t) * 2, hipMemcpyHostToDevice, c->st));
  HIPCHK(c, hipMemcpyAsync(d.amp_off, c->amp_off.data(), sizeof(uint32_t) * 2, hipMemcpyHostToDevice, c->st));
  HIPCHK(c, hipMemcpyAsync(c->d_amp_num, c->amp_num_rep.data(), sizeof(uint32_t), hipMemcpyHostToDevice, c->st));
  HIPCHK(c, hipStreamSynchronize(c->st));
  c->drifted = false;
  c->have_cells = c->cells_fresh = c->have_density = false;
  return MDPD_OK;
}

}  // extern "C"
