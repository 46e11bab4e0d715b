// Stand-alone host driver of the engine with the control flow of the
// reference's src/main.cpp:77-209 (single process, no CALC_HEIGHT), written
// against the C++ facade so that the calls read like the reference's:
//
//   mdpd_main <target_dir> [--seed S] [--chem] [--chem-mt] [--beg-chem T]
//             [--nlist-width W]
//
// Reads <target_dir>/run_param.txt and init_config.xyz, runs all_time steps
// and writes kin_temp.txt, pressure.txt, diffus.txt, num_amp.txt,
// traject.xyz and fin_config.xyz in the reference's formats.  --seed fixes
// CalcForceEpEpDPD::m_seed (the reference takes time(nullptr), main.cpp:133);
// --chem enables the RandomChemEvent branch (the reference's CHEM_MODE) for
// steps with time >= T (--beg-chem, default 100000 = Parameter::beg_chem,
// parameter.hpp:254).  --chem-mt takes the chemistry's draws from the
// reference's MT19937 stream in the reference's order, seeded as
// ChemManager(seed) is on one process (time(nullptr) * rank = 0,
// main.cpp:151), instead of the keyed draws.  --nlist-width sets the engine's per-bead pair-list
// capacity (default 96) for dense systems.
//
// The main loop runs as the engine's production path: the steps between two
// observation points (time % step_mac == 0 or time % step_mic == 0) are one
// mdpd_run, replayed from captured graphs, with the chemistry event inside
// each step; the observations then read the state after the step whose time
// they print, as main.cpp:198-208 does.
#include <chrono>
#include <cstring>
#include <ctime>
#include <iostream>
#include <string>
#include <vector>

#include "../../include/mdpd_facade.hpp"

namespace {

constexpr uint32_t kBegChem = 100000;  // Parameter::beg_chem (parameter.hpp:254)

// Parameter::RemoveCMDriftLocal (parameter.hpp:390-398)
void RemoveCMDriftLocal(std::vector<mdpd::FPDPD>& v) {
  double cm[3] = {0.0, 0.0, 0.0};
  for (const auto& p : v)
    for (int a = 0; a < 3; ++a) cm[a] += p.vel[a];
  for (int a = 0; a < 3; ++a) cm[a] /= (double)v.size();
  for (auto& p : v)
    for (int a = 0; a < 3; ++a) p.vel[a] -= cm[a];
}

// Parameter::AdjustCMToBoxCenter + ApplyPBC (parameter.hpp:400-418)
void AdjustCMToBoxCenter(std::vector<mdpd::FPDPD>& v, const double L[3]) {
  double cm[3] = {0.0, 0.0, 0.0};
  for (const auto& p : v)
    for (int a = 0; a < 3; ++a) cm[a] += p.pos[a];
  for (int a = 0; a < 3; ++a) cm[a] /= (double)v.size();
  for (auto& p : v)
    for (int a = 0; a < 3; ++a) {
      p.pos[a] += L[a] * 0.5 - cm[a];
      p.pos[a] -= std::floor(p.pos[a] * (1.0 / L[a])) * L[a];
    }
}

// Parameter::CheckParticleConfigIsValid (parameter.hpp:436-470)
void CheckParticleConfigIsValid(const std::vector<mdpd::FPDPD>& v, uint32_t init_amp_num,
                                const double L[3]) {
  const uint32_t n = (uint32_t)v.size();
  for (const auto& p : v) {
    if (p.id >= n || p.prop >= 3) mdpd::Abort("invalid particle id or prop");
    if (p.prop != MDPD_SOLVENT) {
      if (p.amp_id >= init_amp_num || p.unit >= mdpd::kAllUnit)
        mdpd::Abort("invalid amp_id or unit");
    } else if (p.amp_id != 0xffffffffu || p.unit != 0xffffffffu) {
      mdpd::Abort("solvent with amp_id or unit");
    }
    for (int a = 0; a < 3; ++a)
      if (!(p.pos[a] <= L[a] && p.pos[a] >= 0.0))
        mdpd::Abort("There is a particle in outside range.");
  }
}

void do_observe_macro(mdpd::Observer& ob, mdpd::ParticleSystem& sys, uint32_t amp_num,
                      uint32_t sol_num, const double bonded_vir[3]) {
  ob.KineticTempera(sys);
  ob.Pressure(sys, bonded_vir);
  ob.Diffusion(sys, amp_num, sol_num);
  ob.NumAmp(amp_num);
}

}  // namespace

int main(int argc, char* argv[]) {
  if (argc < 2) {
    std::cerr << "argv[1] = target directory.\n";
    std::abort();
  }
  const std::string cdir = argv[1];
  bool chem = false, chem_mt = false, have_seed = false;
  uint32_t seed = 0, beg_chem = kBegChem, nlist_width = 0;
  for (int i = 2; i < argc; ++i) {
    if (!std::strcmp(argv[i], "--chem")) chem = true;
    else if (!std::strcmp(argv[i], "--chem-mt")) chem_mt = true;
    else if (!std::strcmp(argv[i], "--beg-chem") && i + 1 < argc)
      beg_chem = (uint32_t)std::strtoul(argv[++i], nullptr, 10);
    else if (!std::strcmp(argv[i], "--nlist-width") && i + 1 < argc)
      nlist_width = (uint32_t)std::strtoul(argv[++i], nullptr, 10);
    else if (!std::strcmp(argv[i], "--seed") && i + 1 < argc) {
      seed = (uint32_t)std::strtoul(argv[++i], nullptr, 10);
      have_seed = true;
    } else {
      std::cerr << "unknown argument " << argv[i] << "\n";
      return 2;
    }
  }

  // run input and coefficients (Parameter::LoadParam + CalcInterCoef)
  mdpd_params p;
  mdpd_run_input in;
  mdpd::Check(mdpd_read_run_param(cdir.c_str(), &p, &in), nullptr, "mdpd_read_run_param");

  // particle configuration (Parameter::LoadParticleConfig, parameter.hpp:420-434)
  std::vector<mdpd::FPDPD> ptcl;
  mdpd::Clock::time = mdpd::ReadXYZForm(cdir + "/init_config.xyz", ptcl);
  if (ptcl.size() != (size_t)in.init_amp_num * mdpd::kAllUnit + in.sol_num)
    mdpd::Abort("# of lines is not equal to the run input parameter information.");
  RemoveCMDriftLocal(ptcl);
  AdjustCMToBoxCenter(ptcl, p.box);
  CheckParticleConfigIsValid(ptcl, in.init_amp_num, p.box);

  // Parameter::CalcCorePtclId (parameter.hpp:472-482)
  std::vector<uint32_t> core_amp_id(in.core_amp_id, in.core_amp_id + in.n_core);
  std::vector<uint32_t> core_ptcl_id(in.n_core, 0xffffffffu);
  for (const auto& q : ptcl)
    for (uint32_t j = 0; j < in.n_core; ++j)
      if (q.amp_id == core_amp_id[j] && q.unit == 0) core_ptcl_id[j] = q.id;

  const uint32_t est_max_ptcl_num =
      chem ? in.sol_num + (uint32_t)(in.max_amp_num * mdpd::kAllUnit * 1.3) : (uint32_t)ptcl.size();
  mdpd::ParticleSystem system(p, est_max_ptcl_num > ptcl.size() ? est_max_ptcl_num
                                                                 : (uint32_t)ptcl.size());
  system.setParticles(ptcl);
  if (nlist_width) {
    mdpd_tuning tu;
    mdpd::Check(mdpd_get_tuning(system.ctx(), &tu), system.ctx(), "mdpd_get_tuning");
    tu.nlist_width = (int32_t)nlist_width;
    mdpd::Check(mdpd_set_tuning(system.ctx(), &tu), system.ctx(), "mdpd_set_tuning");
  }

  mdpd::CalcForceEpEpDPD::m_seed = have_seed ? seed : (uint32_t)std::time(nullptr);
  mdpd::Clock::Push(system);

  // Initial step
  mdpd::TreeForForceShort dens_tree, force_tree;
  mdpd::drift_and_predict(system);
  system.exchangeParticle();
  dens_tree.calcForceAllAndWriteBack(mdpd::CalcDensity(), system);
  force_tree.calcForceAllAndWriteBack(mdpd::CalcForceEpEpDPD(), system);

  double bonded_vir[3] = {0.0, 0.0, 0.0};
  mdpd::ForceBonded fbonded(system, ptcl, in.init_amp_num, core_amp_id);
  fbonded.CalcListedForce(system, bonded_vir);

  mdpd::Observer observer(cdir);
  observer.Initialize();
  uint32_t amp_num = in.init_amp_num;
  do_observe_macro(observer, system, amp_num, in.sol_num, bonded_vir);
  observer.Trajectory(system);

  mdpd_chem_params cp{};
  cp.p_thresld = in.p_thresld;
  cp.eps = in.eps;
  cp.influ_rad = in.influ_rad;
  cp.influ_hei = in.influ_hei;
  cp.influ_dep = in.influ_dep;
  cp.influ_grd = in.influ_grd;
  cp.max_amp_num = in.max_amp_num;
  cp.n_core = in.n_core;
  cp.core_ptcl_id = core_ptcl_id.data();

  mdpd::kick(system);
  mdpd::Clock::time++;
  // End of initial step.

  // the chemistry branch (main.cpp:186-196) runs inside every step of the
  // runs below once time >= beg_chem
  if (chem)
    mdpd::Check(mdpd_set_chemistry(system.ctx(), &cp, beg_chem, 1), system.ctx(),
                "mdpd_set_chemistry");
  if (chem && chem_mt)
    mdpd::Check(mdpd_set_chem_mt(system.ctx(), 0u, 1), system.ctx(), "mdpd_set_chem_mt");
  const uint32_t atime = mdpd::Clock::time + in.all_time - 1;
  mdpd::Clock::Push(system);
  const uint32_t t_begin = mdpd::Clock::time;
  const auto wall0 = std::chrono::steady_clock::now();  // timer_start (main.cpp:45-56)
  while (mdpd::Clock::time < atime) {
    // steps up to the next observation point, then the observations
    uint32_t t_obs = mdpd::Clock::time;
    while (t_obs + 1 < atime && t_obs % in.step_mac != 0 && t_obs % in.step_mic != 0) ++t_obs;
    mdpd::Check(mdpd_run(system.ctx(), t_obs - mdpd::Clock::time + 1), system.ctx(), "mdpd_run");
    if (chem) {
      int32_t stopped = 0;
      mdpd::Check(mdpd_chem_status(system.ctx(), &stopped, nullptr, nullptr), system.ctx(),
                  "mdpd_chem_status");
      if (stopped) {
        // RandomChemEvent reached max_amp_num: the loop ends at that step's
        // time, before its observations (main.cpp:188)
        mdpd::Check(mdpd_get_clock(system.ctx(), &mdpd::Clock::time), system.ctx(), "mdpd_get_clock");
        std::cout << "Now param.amp_num >= param.max_amp_num.\nSo, simulation will stop.\n";
        break;
      }
    }
    mdpd::Clock::time = t_obs;
    if (t_obs % in.step_mac == 0) {
      mdpd_observables o;
      mdpd::Check(mdpd_observe(system.ctx(), &o), system.ctx(), "mdpd_observe");
      do_observe_macro(observer, system, o.n_amp, in.sol_num, o.bonded_vir);
    }
    if (t_obs % in.step_mic == 0) observer.Trajectory(system);
    mdpd::Clock::time = t_obs + 1;
  }

  mdpd::Check(mdpd_sync(system.ctx()), system.ctx(), "mdpd_sync");
  // timer_stop + show_duration, with the loop's steps and seconds for tools
  const double wall =
      std::chrono::duration<double>(std::chrono::steady_clock::now() - wall0).count();
  int all_sec = (int)wall;
  const int dd = all_sec / 86400, hh = (all_sec % 86400) / 3600, mm = (all_sec % 3600) / 60,
            ss = all_sec % 60;
  std::cout << "Simulation time is \n" << dd << "d " << hh << "h " << mm << "m " << ss << "s\n";
  std::printf("main loop: %u steps in %.6f s\n", mdpd::Clock::time - t_begin, wall);

  observer.FinConfig(system);
  observer.CleanUp();
  std::printf("m_seed %u\n", mdpd::CalcForceEpEpDPD::m_seed);
  return 0;
}
