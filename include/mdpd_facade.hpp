// C++ host facade over the C ABI (mdpd_capi.h) that keeps the names and
// argument meaning of the reference's driver surface, so that a driver
// written against the reference (src/main.cpp:77-209) reads the same:
//
//   reference                                   facade
//   PS::ParticleSystem<FPDPD> + DomainInfo      mdpd::ParticleSystem
//   system.exchangeParticle(dinfo)              system.exchangeParticle()
//   TreeForForceShort<...>::Gather              mdpd::TreeForForceShort
//     .calcForceAllAndWriteBack(CalcDensity(), system, dinfo)
//     .calcForceAllAndWriteBack(CalcForceEpEpDPD(), system, dinfo)
//   drift_and_predict / kick / remove_cmdrift_global   (driftkick.hpp:5-54)
//   ForceBonded<Tpsys>::CalcListedForce(sys, vir)       (bdf_calculator.hpp:191)
//   ChemManager<FPDPD>::RandomChemEvent(sys, topol, param) (chemmanager.hpp:244)
//   Observer<Tpsys> (observer.hpp:7-178): the same files and columns
//   ConfigMaker (src/config_maker.cpp:13-533)   mdpd::ConfigMaker
//
// Errors behave as in the reference: a failed call prints the engine's
// message and aborts (PS::Abort).  Header-only; link with -lmdpd.
#pragma once

#include <array>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <ctime>
#include <string>
#include <vector>

#include "mdpd_capi.h"

namespace mdpd {

// The reference's host particle record (ptcl_class.hpp:29-36), 232 bytes.
struct FPDPD {
  uint32_t id, prop, amp_id, unit;
  double pos[3], delta_sumr[3];
  double vel[3], vel_buf[3];
  double acc[3];
  double press[3];
  double nei_cm_pos[2][3];
  double density[3];
};
static_assert(sizeof(FPDPD) == 232, "FPDPD must match the reference layout");

// Parameter::atom_type (parameter.hpp:207-209)
constexpr char kAtomType[] = "ONCSPZXONCSPZXONCSPZX";
constexpr uint32_t kAllUnit = 4;  // Parameter::all_unit (H1T3)

[[noreturn]] inline void Abort(const std::string& msg) {
  std::fprintf(stderr, "%s\n", msg.c_str());
  std::fflush(stderr);
  std::abort();
}

inline void Check(int rc, const mdpd_ctx* c, const char* call) {
  if (rc != MDPD_OK)
    Abort(std::string(call) + " failed (" + std::to_string(rc) + "): " +
          (c ? mdpd_last_error(c) : mdpd_global_error()));
}

// One system on one GPU: the particle array, the domain and both trees of
// the reference live in the engine context.
class ParticleSystem {
 public:
  ParticleSystem(const mdpd_params& p, uint32_t capacity) : p_(p) {
    Check(mdpd_create(&p, capacity, &ctx_), nullptr, "mdpd_create");
  }
  ~ParticleSystem() {
    if (ctx_) mdpd_destroy(ctx_);
  }
  ParticleSystem(const ParticleSystem&) = delete;
  ParticleSystem& operator=(const ParticleSystem&) = delete;

  mdpd_ctx* ctx() const { return ctx_; }
  const mdpd_params& params() const { return p_; }

  // host records in, host records out (in upload order)
  void setParticles(const std::vector<FPDPD>& v) {
    Check(mdpd_upload_fpdpd(ctx_, v.data(), (uint32_t)v.size(), sizeof(FPDPD)), ctx_,
          "mdpd_upload_fpdpd");
  }
  std::vector<FPDPD> getParticles() const {
    std::vector<FPDPD> v(getNumberOfParticleLocal());
    Check(mdpd_download_fpdpd(ctx_, v.data(), (uint32_t)v.size(), sizeof(FPDPD)), ctx_,
          "mdpd_download_fpdpd");
    return v;
  }
  uint32_t getNumberOfParticleLocal() const {
    uint32_t n = 0;
    Check(mdpd_num_beads(ctx_, &n), ctx_, "mdpd_num_beads");
    return n;
  }
  // ParticleSystem::exchangeParticle + the tree build of the next pass
  void exchangeParticle() { Check(mdpd_rebuild_cells(ctx_), ctx_, "mdpd_rebuild_cells"); }

 private:
  mdpd_params p_;
  mdpd_ctx* ctx_ = nullptr;
};

// Functor tags of nbdf_calculator.hpp: the work runs on the device.
struct CalcDensity {};
struct CalcForceEpEpDPD {
  static inline uint32_t m_seed = 0;  // nbdf_calculator.hpp:65
};

// Parameter::time (parameter.hpp:220) lives on the device; the facade keeps
// the host copy in step with it.
struct Clock {
  static inline uint32_t time = 0;
  static void Push(ParticleSystem& s) {
    Check(mdpd_set_clock(s.ctx(), time, CalcForceEpEpDPD::m_seed), s.ctx(), "mdpd_set_clock");
  }
};

class TreeForForceShort {
 public:
  void calcForceAllAndWriteBack(CalcDensity, ParticleSystem& s) {
    Check(mdpd_density(s.ctx()), s.ctx(), "mdpd_density");
  }
  void calcForceAllAndWriteBack(CalcForceEpEpDPD, ParticleSystem& s) {
    Clock::Push(s);
    Check(mdpd_force(s.ctx()), s.ctx(), "mdpd_force");
  }
};

inline void drift_and_predict(ParticleSystem& s) {
  Check(mdpd_drift_predict(s.ctx()), s.ctx(), "mdpd_drift_predict");
}
inline void kick(ParticleSystem& s) { Check(mdpd_kick(s.ctx()), s.ctx(), "mdpd_kick"); }
inline void remove_cmdrift_global(ParticleSystem& s) {
  Check(mdpd_remove_cm_drift(s.ctx()), s.ctx(), "mdpd_remove_cm_drift");
}

// ForceBonded (bdf_calculator.hpp:11-243): the global topology is built from
// the particles' (amp_id, unit) as MakeGlobalBondedList does (:59-90).
class ForceBonded {
 public:
  ForceBonded(ParticleSystem& s, const std::vector<FPDPD>& ptcl, uint32_t amp_num,
              const std::vector<uint32_t>& core_amp_id)
      : glob_topol_(amp_num * kAllUnit, 0xffffffffu) {
    uint32_t cnt = 0;
    for (uint32_t i = 0; i < ptcl.size(); ++i) {
      const uint32_t aid = ptcl[i].amp_id, unit = ptcl[i].unit;
      if (aid != 0xffffffffu && unit != 0xffffffffu) {
        if (aid >= amp_num || unit >= kAllUnit) Abort("ForceBonded: amp_id/unit out of range");
        glob_topol_[aid * kAllUnit + unit] = i;
        ++cnt;
      } else if (ptcl[i].prop != MDPD_SOLVENT) {
        Abort("ForceBonded: a non-solvent particle without amp_id/unit");
      }
    }
    if (cnt != glob_topol_.size()) Abort("ForceBonded: incomplete amphiphiles");
    Check(mdpd_set_topology(s.ctx(), glob_topol_.data(), amp_num, core_amp_id.data(),
                            (uint32_t)core_amp_id.size()),
          s.ctx(), "mdpd_set_topology");
  }
  void CalcListedForce(ParticleSystem& s, double vir[3]) {
    Check(mdpd_bonded(s.ctx(), vir), s.ctx(), "mdpd_bonded");
  }
  const std::vector<uint32_t>& glob_topol() const { return glob_topol_; }

 private:
  std::vector<uint32_t> glob_topol_;
};

// ChemManager (chemmanager.hpp:18-281).  ChemManager(cp): keyed draws on
// the chem_seed of the system's parameters.  ChemManager(cp, sys, seed): the
// reference's draws, one MT19937 stream started by init_genrand(seed) as
// ChemManager(seed) does (chemmanager.hpp:176-179).
class ChemManager {
 public:
  explicit ChemManager(const mdpd_chem_params& cp) : cp_(cp) {}
  ChemManager(const mdpd_chem_params& cp, ParticleSystem& s, uint32_t seed) : cp_(cp) {
    Check(mdpd_set_chem_mt(s.ctx(), seed, 1), s.ctx(), "mdpd_set_chem_mt");
  }
  // false once amp_num has reached max_amp_num (the reference's stop signal)
  bool RandomChemEvent(ParticleSystem& s) {
    Clock::Push(s);
    int stop = 0;
    Check(mdpd_chem_event(s.ctx(), &cp_, &stop), s.ctx(), "mdpd_chem_event");
    return stop == 0;
  }

 private:
  mdpd_chem_params cp_;
};

// io_util::WriteXYZForm + FPDPD::writeAscii (io_util.hpp:25-31, ptcl_class.hpp:67-73)
inline void WriteXYZForm(const std::vector<FPDPD>& v, uint32_t time, FILE* fp) {
  std::fprintf(fp, "%u\n", (unsigned)v.size());
  std::fprintf(fp, "time %u\n", time);
  for (const FPDPD& p : v)
    std::fprintf(fp,
                 "%c %.15g %.15g %.15g %u %u %u %u %.15g %.15g %.15g %.15g %.15g %.15g %.15g "
                 "%.15g %.15g\n",
                 kAtomType[p.prop], p.pos[0], p.pos[1], p.pos[2], p.id, p.prop, p.amp_id, p.unit,
                 p.vel[0], p.vel[1], p.vel[2], p.vel_buf[0], p.vel_buf[1], p.vel_buf[2], p.acc[0],
                 p.acc[1], p.acc[2]);
}

// io_util::ReadXYZForm + FPDPD::readAscii; returns the time of the header.
// The library reads the file once and parses its lines in parallel.
inline uint32_t ReadXYZForm(const std::string& fname, std::vector<FPDPD>& v) {
  uint32_t n = 0, t = 0;
  if (mdpd_read_xyz_header(fname.c_str(), &n, &t) != MDPD_OK) Abort(mdpd_global_error());
  v.assign(n, FPDPD{});
  if (mdpd_read_xyz(fname.c_str(), v.data(), n, sizeof(FPDPD), &n, &t) != MDPD_OK)
    Abort(mdpd_global_error());
  return t;
}

// ConfigMaker (src/config_maker.cpp:13-533): config_param.txt in the target
// directory -> init_config.xyz, the beads generated on the device.  As in the
// reference, Initialize() seeds from time(NULL) (:452-454); pass a seed to
// reproduce a configuration.  GenParticles(system) generates straight into
// an engine context instead of host records.
class ConfigMaker {
 public:
  explicit ConfigMaker(std::string cdir) : cdir_(std::move(cdir)) {}
  void Initialize(uint32_t seed = (uint32_t)std::time(nullptr)) { seed_ = seed; }
  void LoadParam() {
    Check(mdpd_read_config_param(cdir_.c_str(), &spec_), nullptr, "mdpd_read_config_param");
    spec_.seed = seed_;
  }
  void GenParticles() {
    uint32_t n = 0;
    Check(mdpd_config_count(&spec_, &n), nullptr, "mdpd_config_count");
    std::vector<double> pos(3 * (size_t)n), vel(3 * (size_t)n);
    std::vector<uint32_t> id(n), prop(n), amp(n), unit(n);
    Check(mdpd_make_config_host(&spec_, n, pos.data(), vel.data(), id.data(), prop.data(),
                                amp.data(), unit.data(), &kin_temp_),
          nullptr, "mdpd_make_config_host");
    prtcls_.assign(n, FPDPD{});
    for (uint32_t i = 0; i < n; ++i) {
      FPDPD& q = prtcls_[i];
      q.id = id[i];
      q.prop = prop[i];
      q.amp_id = amp[i];
      q.unit = unit[i];
      for (int a = 0; a < 3; ++a) {
        q.pos[a] = pos[3 * (size_t)i + a];
        q.vel[a] = q.vel_buf[a] = vel[3 * (size_t)i + a];
      }
    }
  }
  void GenParticles(ParticleSystem& system) {
    Check(mdpd_make_config(system.ctx(), &spec_, &kin_temp_), system.ctx(), "mdpd_make_config");
  }
  void DumpParticleConfig() const {
    const std::string f = cdir_ + "/init_config.xyz";
    FILE* fp = std::fopen(f.c_str(), "w");
    if (!fp) Abort("cannot open " + f);
    WriteXYZForm(prtcls_, 0, fp);
    std::fclose(fp);
  }
  void ShowSystemParameter() const {
    static const char* modes[] = {"flat", "asym_flat", "sphere", "cylind", "random"};
    std::printf("Show ConfigMaker Paramter\nTempera %g\nbox_leng %g %g %g\namp_num %u\n"
                "head_unit %u\ntail_unit %u\nall_unit %u\nbond_leng %g\nmode %s\n",
                spec_.temperature, spec_.box[0], spec_.box[1], spec_.box[2], spec_.amp_num,
                spec_.head_unit, spec_.tail_unit, spec_.head_unit + spec_.tail_unit,
                spec_.bond_leng, modes[spec_.mode]);
  }
  const std::vector<FPDPD>& particles() const { return prtcls_; }
  const mdpd_config_spec& spec() const { return spec_; }
  double kinetic_temperature() const { return kin_temp_; }

 private:
  std::string cdir_;
  uint32_t seed_ = 0;
  mdpd_config_spec spec_{};
  std::vector<FPDPD> prtcls_;
  double kin_temp_ = 0.0;
};

// Observer (observer.hpp:7-178): same file names, columns and formats.  One
// device reduction (mdpd_observe) serves every quantity of a step.
class Observer {
 public:
  explicit Observer(std::string cdir) : cdir_(std::move(cdir)) {}
  ~Observer() { CleanUp(); }
  void Initialize() {
    static const char* names[kNum] = {"kin_temp.txt", "pressure.txt", "diffus.txt",
                                      "num_amp.txt", "traject.xyz", "fin_config.xyz"};
    for (int i = 0; i < kNum; ++i) {
      const std::string f = cdir_ + "/" + names[i];
      fp_[i] = std::fopen(f.c_str(), "w");
      if (!fp_[i]) Abort("cannot open " + f);
    }
    // the engine's writer appends the trajectory frames and writes the final
    // configuration itself: these two stay empty files until it does
    for (int i : {kTraj, kFin}) {
      std::fclose(fp_[i]);
      fp_[i] = nullptr;
    }
  }
  void KineticTempera(ParticleSystem& s) {
    const mdpd_observables& o = Observe(s);
    std::fprintf(fp_[kKin], "%.15g %.15g %.15g %.15g\n", o.kin_temp[0], o.kin_temp[1],
                 o.kin_temp[2], o.kin_temp[3]);
  }
  // P = [sum_i (press_i / 2 + v_i^2) + bonded_vir / 2] / V (observer.hpp:108-130)
  void Pressure(ParticleSystem& s, const double bonded_vir[3]) {
    const mdpd_observables& o = Observe(s);
    const double* L = s.params().box;
    const double iv = (1.0 / L[0]) * (1.0 / L[1]) * (1.0 / L[2]);
    double p[3];
    for (int a = 0; a < 3; ++a)
      p[a] = (0.5 * o.pair_vir[a] + o.kin_temp[a] * o.n_beads + 0.5 * bonded_vir[a]) * iv;
    std::fprintf(fp_[kPress], "%.15g %.15g %.15g\n", p[0], p[1], p[2]);
  }
  // MSD per bead, per amphiphile CM and per solvent bead (observer.hpp:132-178);
  // the engine tracks amp_num and sol_num itself, the arguments are kept for
  // the reference's signature
  void Diffusion(ParticleSystem& s, uint32_t amp_num, uint32_t sol_num) {
    (void)amp_num;
    (void)sol_num;
    const mdpd_observables& o = Observe(s);
    if (first_diff_) {
      std::fprintf(fp_[kDiff], "#D_ptcl D_lipid D_sol\n");
      first_diff_ = false;
    }
    std::fprintf(fp_[kDiff], "%.15g %.15g %.15g\n", o.diffusion[0], o.diffusion[1],
                 o.diffusion[2]);
  }
  void NumAmp(uint32_t amp_num) { std::fprintf(fp_[kNumAmp], "%u\n", amp_num); }
  // a frame of traject.xyz, formatted and written by a host thread while the
  // run goes on (mdpd_write_xyz_async)
  void Trajectory(ParticleSystem& s) {
    Check(mdpd_write_xyz_async(s.ctx(), (cdir_ + "/traject.xyz").c_str(), Clock::time, 1), s.ctx(),
          "mdpd_write_xyz_async");
  }
  void FinConfig(ParticleSystem& s) {
    Check(mdpd_write_xyz_async(s.ctx(), (cdir_ + "/fin_config.xyz").c_str(), Clock::time, 0), s.ctx(),
          "mdpd_write_xyz_async");
    Check(mdpd_io_wait(s.ctx()), s.ctx(), "mdpd_io_wait");
  }
  void FlushAll() {
    for (FILE* f : fp_)
      if (f) std::fflush(f);
  }
  void CleanUp() {
    for (FILE*& f : fp_)
      if (f) {
        std::fclose(f);
        f = nullptr;
      }
  }

 private:
  enum { kKin = 0, kPress, kDiff, kNumAmp, kTraj, kFin, kNum };
  // one device reduction per state: the quantities of one observation share
  // it (keyed on the engine's clock, which every step advances)
  const mdpd_observables& Observe(ParticleSystem& s) {
    uint32_t t = 0;
    Check(mdpd_get_clock(s.ctx(), &t), s.ctx(), "mdpd_get_clock");
    if (!have_ || obs_.time != t) {
      Check(mdpd_observe(s.ctx(), &obs_), s.ctx(), "mdpd_observe");
      have_ = true;
    }
    return obs_;
  }
  std::string cdir_;
  FILE* fp_[kNum] = {};
  mdpd_observables obs_{};
  bool have_ = false;
  bool first_diff_ = true;
};

}  // namespace mdpd
