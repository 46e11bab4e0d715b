// run_param.txt input and the reference's coefficient derivation, host C++
// (the tag reader is in tags.hpp).
//
// ReadTagValues / Matching* follow parameter.hpp:16-30,300-376 (';'
// comments, head/tail blanks stripped, later tags override earlier ones,
// floating values rounded through std::stof as at :339); the coefficient
// tables follow CalcInterCoef (parameter.hpp:131-186) and the truncated-cone
// helpers CalcInfluGrdAndRad / CalcInfluVol (:42-72).
#include <cmath>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>
#include <string>
#include <vector>

#include "ctx.hpp"
#include "tags.hpp"

using mdpd::Parse;
using mdpd::TagMap;
using mdpd::read_tags;

extern "C" int mdpd_read_run_param(const char* dir, mdpd_params* p, mdpd_run_input* in) {
  if (!dir || !p || !in) return mdpd::fail(nullptr, MDPD_ERR_ARG, "null argument");
  const std::string fname = std::string(dir) + "/run_param.txt";
  std::ifstream fin(fname.c_str());
  if (!fin) return mdpd::fail(nullptr, MDPD_ERR_PARSE, fname + ": Cannot open file");
  TagMap tv;
  read_tags(fin, tv);
  Parse P{tv, ""};
  std::memset(p, 0, sizeof(*p));
  std::memset(in, 0, sizeof(*in));
  bool ok = true;
  ok &= P.f64("box_leng", p->box, 3);
  ok &= P.u32("init_amp_num", &in->init_amp_num, 1);
  ok &= P.u32("sol_num", &in->sol_num, 1);
  ok &= P.f64("dt", &p->dt, 1);
  ok &= P.f64("cf_r", in->cf_r_in, 6);
  ok &= P.f64("cf_s", &p->cf_s, 1);
  ok &= P.f64("cf_b", &p->cf_b, 1);
  ok &= P.f64("cf_b_rigid", &p->cf_b_rigid, 1);
  in->vbb = 0.1;  // Parameter::vbb default (parameter.hpp:228)
  if (tv.count("vbb")) ok &= P.f64("vbb", &in->vbb, 1);
  ok &= P.f64("chi", &in->chi, 1);
  ok &= P.f64("kappa", &in->kappa, 1);
  ok &= P.f64("rho_co", &in->rho_co, 1);
  ok &= P.u32("all_time", &in->all_time, 1);
  ok &= P.u32("step_mic", &in->step_mic, 1);
  ok &= P.u32("step_mac", &in->step_mac, 1);
  ok &= P.f64("p_thresld", &in->p_thresld, 1);
  ok &= P.f64("eps", &in->eps, 1);
  ok &= P.u32("max_amp_num", &in->max_amp_num, 1);
  ok &= P.f64("influ_hei", &in->influ_hei, 1);
  ok &= P.f64("influ_dep", &in->influ_dep, 1);
  if (tv.count("influ_vol")) {
    ok &= P.f64("influ_vol", &in->influ_vol, 1);
    ok &= P.f64("upside_rad", &in->upside_rad, 1);
    if (ok) {  // CalcInfluGrdAndRad, parameter.hpp:42-56
      const double hs = in->influ_hei + in->influ_dep;
      const double vs = 12.0 * in->influ_vol / (M_PI * hs) - 3.0 * in->upside_rad * in->upside_rad;
      if (vs > 0.0) {
        const double down = (std::sqrt(vs) - in->upside_rad) * 0.5;
        in->influ_grd = (down - in->upside_rad) / hs;
      } else {
        const double hn = 3.0 * in->influ_vol / (M_PI * in->upside_rad * in->upside_rad);
        in->influ_grd = -in->upside_rad / hn;
      }
      in->influ_rad = in->upside_rad + in->influ_grd * in->influ_hei;
      if (!(in->influ_rad > 0.0)) return mdpd::fail(nullptr, MDPD_ERR_PARSE, "influ_rad <= 0");
    }
  } else {
    ok &= P.f64("influ_rad", &in->influ_rad, 1);
    ok &= P.f64("influ_grd", &in->influ_grd, 1);
    if (ok) {  // CalcInfluVol, parameter.hpp:58-72
      const double hs = in->influ_hei + in->influ_dep;
      const double up = in->influ_rad - in->influ_grd * in->influ_hei;
      const double down = in->influ_rad + in->influ_grd * in->influ_dep;
      if (down >= 0.0) {
        const double ua = M_PI * up * up, da = M_PI * down * down;
        in->influ_vol = (ua + da + std::sqrt(da * ua)) * hs / 3.0;
      } else {
        const double hn = -up / in->influ_grd;
        in->influ_vol = M_PI * up * up * hn / 3.0;
      }
    }
  }
  if (!tv.count("core_amp_id")) {
    return mdpd::fail(nullptr, MDPD_ERR_PARSE, "core_amp_id is not specified in run input paramter.");
  }
  const size_t ncore = tv["core_amp_id"].size();
  if (ncore > 64) return mdpd::fail(nullptr, MDPD_ERR_PARSE, "more than 64 core_amp_id values");
  in->n_core = (uint32_t)ncore;
  ok &= P.u32("core_amp_id", in->core_amp_id, ncore);
  if (!ok) return mdpd::fail(nullptr, MDPD_ERR_PARSE, P.err);

  // CalcInterCoef, parameter.hpp:138-186 (Reo = 1.4, arc = 0.9, H1T3)
  enum { PHIL = 0, PHOB = 1, SOLV = 2, NP = 3 };
  const double reo = 1.4, arc = 0.9, tempera = 1.0;
  const int all_unit = 4;
  double cf_c[3][3], cf_g[3][3], cf_r[3][3], cf_m[3][3][3];
  int cnt = 0;
  for (int i = 0; i < NP; ++i)
    for (int j = i; j < NP; ++j) cf_r[i][j] = in->cf_r_in[cnt++];
  const double kappa = in->kappa, rho_co = in->rho_co, chi = in->chi, vbb = in->vbb;
  cf_c[PHOB][PHOB] = -2.0 * (kappa + 3.0) / rho_co;
  cf_c[PHIL][PHIL] = cf_c[SOLV][SOLV] = vbb;
  cf_c[PHIL][PHOB] = cf_c[PHOB][PHIL] = chi / (rho_co) + 0.5 * (cf_c[PHIL][PHIL] + cf_c[PHOB][PHOB]);
  cf_c[SOLV][PHOB] = cf_c[PHOB][SOLV] = cf_c[PHIL][PHOB];
  cf_c[SOLV][PHIL] = cf_c[PHIL][SOLV] = cf_c[PHIL][PHIL];
  cf_m[PHOB][PHOB][PHOB] = 1.5 * (kappa + 2.0) / (rho_co * rho_co);
  for (int i = 0; i < NP; ++i)
    for (int j = 0; j < NP; ++j)
      for (int k = 0; k < NP; ++k) cf_m[i][j][k] = cf_m[PHOB][PHOB][PHOB];
  cf_m[PHIL][PHIL][PHIL] = 0.0;
  cf_m[SOLV][SOLV][SOLV] = 0.0;
  for (int i = 0; i < NP; ++i)
    for (int j = i + 1; j < NP; ++j) cf_r[j][i] = cf_r[i][j];
  for (int i = 0; i < NP; ++i)
    for (int j = 0; j < NP; ++j) cf_g[i][j] = 0.5 * cf_r[i][j] * cf_r[i][j] / tempera;
  for (int i = 0; i < NP; ++i)
    for (int j = i + 1; j < NP; ++j)
      if (cf_r[i][j] < 0.0) {
        cf_g[i][j] = cf_g[j][i] = 2.0 / ((1.0 / cf_g[i][i]) + (1.0 / cf_g[j][j]));
        cf_r[i][j] = cf_r[j][i] = std::sqrt(2.0 * cf_g[i][j] * tempera);
      }
  for (int i = 0; i < NP; ++i)
    for (int j = 0; j < NP; ++j) cf_r[i][j] /= std::sqrt(p->dt);
  const double mol_energy = (reo * reo * reo) / (all_unit * all_unit);
  const double arc5 = arc * arc * arc * arc * arc;
  const double w2 = -45.0 / ((arc5 * (2.0 * arc - 3.0) + (3.0 * arc - 2.0)) * M_PI);
  const double w3 = 15.0 / M_PI;
  const double dens_unit = ((reo * reo * reo * 15.0) / (2.0 * M_PI * all_unit));
  for (int i = 0; i < NP; ++i)
    for (int j = 0; j < NP; ++j) cf_c[i][j] *= mol_energy * w2;
  for (int i = 0; i < NP; ++i)
    for (int j = 0; j < NP; ++j)
      for (int k = 0; k < NP; ++k) cf_m[i][j][k] *= mol_energy * (2.0 / 3.0) * w3 * dens_unit;

  p->n_species = 3;
  p->n_dens = 3;
  p->envelope = 0;
  p->rd = 1.0;
  p->crho = 1.0;
  p->rdm = 1.0;
  p->arc = arc;
  for (int i = 0; i < NP; ++i)
    for (int j = 0; j < NP; ++j) {
      p->a0[i][j] = 0.0;
      p->c1[i][j] = cf_c[i][j];
      p->sigma[i][j] = cf_r[i][j];
      p->gamma[i][j] = cf_g[i][j];
      for (int k = 0; k < NP; ++k) p->mb[i][j][k] = cf_m[i][j][k];
    }
  p->bond_leng = 0.5;
  p->cm_corr = 100;
  p->conv_p = 0.0;
  return MDPD_OK;
}
