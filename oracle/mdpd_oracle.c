/*
 * mdpd_oracle.c -- CPU oracle of the MDPD step path (FP64, OpenMP).
 * TEST INFRASTRUCTURE ONLY: only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py load it, to check the engine against.  The
 * engine never links or calls it.  See mdpd_oracle.h for what pins each function.
 *
 * Every function restates one piece of kohnakagawa/implicit_dpd (cited as
 * file:line of that tree) with the same FP64 operation order, so that on
 * the reference model (rd = rdm = crho = 1, n_dens = 3, envelope 0) the
 * arithmetic is the reference's own.  Build with -ffp-contract=off.
 */
#include "mdpd_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

/* stream tags of the keyed chemistry draws (shared with the engine) */
#define TAG_CONE 0x9E3779B9u
#define TAG_TANG 0x5bd1e995u
#define TAG_VEL 0x68e31da4u

/* ======================================================================
 * Saru PRNG -- saruprng.hpp:208-255 (3-seed constructor), 322-334 (u32
 * with one LCG step and one Weyl step, specialisation at :199-202),
 * 427-440 (uni_open), 442-445 (nrml: first draw feeds log, second cos).
 * ====================================================================== */
typedef struct { uint32_t state, wstate; } o_saru;

static inline uint32_t o_asr(uint32_t x, int s) { return (uint32_t)(((int32_t)x) >> s); }

static inline o_saru o_saru3(uint32_t seed1, uint32_t seed2, uint32_t seed3) {
  o_saru s;
  seed3 ^= (seed1 << 7) ^ (seed2 >> 6);
  seed2 += (seed1 >> 4) ^ (seed3 >> 15);
  seed1 ^= (seed2 << 9) + (seed3 << 8);
  seed3 ^= 0xA5366B4Du * ((seed2 >> 11) ^ (seed1 << 1));
  seed2 += 0x72BE1579u * ((seed1 << 4) ^ (seed3 >> 16));
  seed1 ^= 0x3F38A6EDu * ((seed3 >> 5) ^ o_asr(seed2, 22));
  seed2 += seed1 * seed3;
  seed1 += seed3 ^ (seed2 >> 2);
  seed2 ^= o_asr(seed2, 17);
  s.state = 0x79dedea3u * (seed1 ^ o_asr(seed1, 14));
  s.wstate = (s.state + seed2) ^ o_asr(s.state, 8);
  s.state = s.state + (s.wstate * (s.wstate ^ 0xdddf97f5u));
  s.wstate = 0xABCB96F7u + (s.wstate >> 1);
  return s;
}

static inline uint32_t o_saru_u32(o_saru* s) {
  s->state = 0x4beb5d59u * s->state + 0x2600e1f7u;                 /* LCG */
  s->wstate = s->wstate + 0x8009d14bu + (o_asr(s->wstate, 31) & 0xda879addu); /* Weyl */
  uint32_t v = (s->state ^ (s->state >> 26)) + s->wstate;
  return (v ^ (v >> 20)) * 0x6957f5a7u;
}

static inline double o_saru_uni(o_saru* s) {
  return ((double)(int32_t)(o_saru_u32(s) >> 1) + 0.5) * (1.0 / 2147483648.0);
}

static inline double o_saru_nrml(o_saru* s) {
  const double u1 = o_saru_uni(s);
  const double u2 = o_saru_uni(s);
  return sqrt(-2.0 * log(u1)) * cos(2.0 * M_PI * u2);
}

void oracle_saru_batch(uint32_t n, const uint32_t* s1, const uint32_t* s2,
                       const uint32_t* s3, uint32_t* out_u32, double* out_uni,
                       double* out_nrml) {
  for (uint32_t k = 0; k < n; ++k) {
    o_saru a = o_saru3(s1[k], s2[k], s3[k]);
    if (out_u32)
      for (int q = 0; q < 8; ++q) out_u32[8 * (size_t)k + q] = o_saru_u32(&a);
    o_saru b = o_saru3(s1[k], s2[k], s3[k]);
    if (out_uni) {
      out_uni[2 * (size_t)k] = o_saru_uni(&b);
      out_uni[2 * (size_t)k + 1] = o_saru_uni(&b);
    }
    o_saru c = o_saru3(s1[k], s2[k], s3[k]);
    if (out_nrml) out_nrml[k] = o_saru_nrml(&c);
  }
}

/* pair noise: nbdf_calculator.hpp:99-105 */
double oracle_pair_nrml(uint32_t id_i, uint32_t id_j, uint32_t seed3) {
  uint32_t m_i = id_i, m_j = id_j;
  if (id_i > id_j) { m_i = id_j; m_j = id_i; }
  o_saru s = o_saru3(m_i, m_j, seed3);
  return o_saru_nrml(&s);
}

/* ======================================================================
 * CalcInterCoef -- parameter.hpp:131-186.  cf_r_upper holds the six
 * upper-triangle values of the run_param.txt tag (FillUpperTri, :32-40).
 * ====================================================================== */
void oracle_calc_inter_coef(double chi, double kappa, double rho_co,
                            double vbb, const double cf_r_upper[6], double dt,
                            double reo, uint32_t all_unit, double arc,
                            mdpd_params* o) {
  enum { PHIL = 0, PHOB = 1, SOLV = 2, NP = 3 };
  double cf_c[3][3], cf_g[3][3], cf_r[3][3], cf_m[3][3][3];
  const double tempera = 1.0;
  int cnt = 0;
  for (int i = 0; i < NP; ++i)
    for (int j = i; j < NP; ++j) cf_r[i][j] = cf_r_upper[cnt++];

  cf_c[PHOB][PHOB] = -2.0 * (kappa + 3.0) / rho_co;
  cf_c[PHIL][PHIL] = cf_c[SOLV][SOLV] = vbb;
  cf_c[PHIL][PHOB] = cf_c[PHOB][PHIL] =
      chi / (rho_co) + 0.5 * (cf_c[PHIL][PHIL] + cf_c[PHOB][PHOB]);
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
    for (int j = 0; j < NP; ++j)
      cf_g[i][j] = 0.5 * cf_r[i][j] * cf_r[i][j] / tempera;
  for (int i = 0; i < NP; ++i)
    for (int j = i + 1; j < NP; ++j)
      if (cf_r[i][j] < 0.0) { /* CalcGammaWithHarmonicMean, :131-136 */
        cf_g[i][j] = cf_g[j][i] = 2.0 / ((1.0 / cf_g[i][i]) + (1.0 / cf_g[j][j]));
        cf_r[i][j] = cf_r[j][i] = sqrt(2.0 * cf_g[i][j] * tempera);
      }
  for (int i = 0; i < NP; ++i)
    for (int j = 0; j < NP; ++j) cf_r[i][j] /= sqrt(dt);

  const double mol_energy = (reo * reo * reo) / (double)(all_unit * all_unit);
  const double arc5 = arc * arc * arc * arc * arc;
  const double w2 = -45.0 / ((arc5 * (2.0 * arc - 3.0) + (3.0 * arc - 2.0)) * M_PI);
  const double w3 = 15.0 / M_PI;
  const double dens_unit = ((reo * reo * reo * 15.0) / (2.0 * M_PI * (double)all_unit));
  for (int i = 0; i < NP; ++i)
    for (int j = 0; j < NP; ++j) cf_c[i][j] *= mol_energy * w2;
  for (int i = 0; i < NP; ++i)
    for (int j = 0; j < NP; ++j)
      for (int k = 0; k < NP; ++k)
        cf_m[i][j][k] *= mol_energy * (2.0 / 3.0) * w3 * dens_unit;

  o->n_species = 3;
  o->n_dens = 3;
  o->envelope = 0;
  o->rd = 1.0;
  o->crho = 1.0;
  o->rdm = 1.0;
  o->arc = arc;
  o->dt = dt;
  for (int i = 0; i < NP; ++i)
    for (int j = 0; j < NP; ++j) {
      o->a0[i][j] = 0.0;
      o->c1[i][j] = cf_c[i][j];
      o->sigma[i][j] = cf_r[i][j];
      o->gamma[i][j] = cf_g[i][j];
      for (int k = 0; k < NP; ++k) o->mb[i][j][k] = cf_m[i][j][k];
    }
}

/* ======================================================================
 * Neighbour enumeration.  The FDPS Gather of the reference hands each
 * functor image-shifted j positions within the search radius
 * (main.cpp:169-174, ptcl_class.hpp:39-41); here the minimum image of
 * r_i - r_j is taken per axis and r^2 = dx*dx + dy*dy + dz*dz.
 * ====================================================================== */
static inline void o_min_image(const double* box, double d[3]) {
  for (int a = 0; a < 3; ++a) {
    if (d[a] > 0.5 * box[a]) d[a] -= box[a];
    else if (d[a] < -0.5 * box[a]) d[a] += box[a];
  }
}

static inline double o_disp(const double* box, const double* pos, uint32_t i,
                            uint32_t j, double d[3]) {
  d[0] = pos[3 * (size_t)i] - pos[3 * (size_t)j];
  d[1] = pos[3 * (size_t)i + 1] - pos[3 * (size_t)j + 1];
  d[2] = pos[3 * (size_t)i + 2] - pos[3 * (size_t)j + 2];
  o_min_image(box, d);
  return d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
}

typedef struct {
  int nc[3];
  int brute;
  uint32_t* start; /* [ncell+1] */
  uint32_t* idx;   /* [n] bead indices, ascending within a cell */
  uint32_t* cell;  /* [n] cell of each bead */
} o_cells;

static int o_cell_coord(double x, double L, int nc) {
  int c = (int)floor(x / L * (double)nc);
  if (c < 0) c += nc * (1 + (-c) / nc);
  if (c >= nc) c %= nc;
  return c;
}

static void o_cells_build(const double* box, double min_edge, uint32_t n,
                          const double* pos, o_cells* c) {
  c->brute = 0;
  for (int a = 0; a < 3; ++a) {
    c->nc[a] = (int)floor(box[a] / min_edge);
    if (c->nc[a] < 3) c->brute = 1;
  }
  c->start = NULL; c->idx = NULL; c->cell = NULL;
  if (c->brute) return;
  const size_t ncell = (size_t)c->nc[0] * c->nc[1] * c->nc[2];
  c->start = (uint32_t*)calloc(ncell + 1, sizeof(uint32_t));
  c->idx = (uint32_t*)malloc((n ? n : 1) * sizeof(uint32_t));
  c->cell = (uint32_t*)malloc((n ? n : 1) * sizeof(uint32_t));
  for (uint32_t i = 0; i < n; ++i) {
    const int cx = o_cell_coord(pos[3 * (size_t)i], box[0], c->nc[0]);
    const int cy = o_cell_coord(pos[3 * (size_t)i + 1], box[1], c->nc[1]);
    const int cz = o_cell_coord(pos[3 * (size_t)i + 2], box[2], c->nc[2]);
    c->cell[i] = (uint32_t)((cz * c->nc[1] + cy) * c->nc[0] + cx);
    c->start[c->cell[i] + 1]++;
  }
  for (size_t k = 0; k < ncell; ++k) c->start[k + 1] += c->start[k];
  uint32_t* fill = (uint32_t*)malloc((ncell + 1) * sizeof(uint32_t));
  memcpy(fill, c->start, (ncell + 1) * sizeof(uint32_t));
  for (uint32_t i = 0; i < n; ++i) c->idx[fill[c->cell[i]]++] = i;
  free(fill);
}

static void o_cells_free(o_cells* c) {
  free(c->start); free(c->idx); free(c->cell);
}

/* Calls fn(ctx, i, j, d, r2) for every j != i with r2 < cut2, in a fixed
   order (27 cells z,y,x-major from -1, beads ascending; or j ascending in
   the brute-force case). */
#define O_FOR_NEIGHBOURS(cells, box, pos, n, i, cut2, ...)                   \
  do {                                                                       \
    if ((cells)->brute) {                                                    \
      for (uint32_t j = 0; j < (n); ++j) {                                   \
        if (j == (i)) continue;                                              \
        double d[3];                                                         \
        const double r2 = o_disp((box), (pos), (i), j, d);                   \
        if (r2 < (cut2)) { __VA_ARGS__ }                                            \
      }                                                                      \
    } else {                                                                 \
      const uint32_t ci_ = (cells)->cell[(i)];                               \
      const int nx_ = (cells)->nc[0], ny_ = (cells)->nc[1], nz_ = (cells)->nc[2]; \
      const int cx_ = (int)(ci_ % (uint32_t)nx_);                            \
      const int cy_ = (int)((ci_ / (uint32_t)nx_) % (uint32_t)ny_);          \
      const int cz_ = (int)(ci_ / ((uint32_t)nx_ * (uint32_t)ny_));          \
      for (int oz_ = -1; oz_ <= 1; ++oz_)                                    \
        for (int oy_ = -1; oy_ <= 1; ++oy_)                                  \
          for (int ox_ = -1; ox_ <= 1; ++ox_) {                              \
            const int qx_ = (cx_ + ox_ + nx_) % nx_;                         \
            const int qy_ = (cy_ + oy_ + ny_) % ny_;                         \
            const int qz_ = (cz_ + oz_ + nz_) % nz_;                         \
            const uint32_t q_ = (uint32_t)((qz_ * ny_ + qy_) * nx_ + qx_);   \
            for (uint32_t s_ = (cells)->start[q_]; s_ < (cells)->start[q_ + 1]; ++s_) { \
              const uint32_t j = (cells)->idx[s_];                           \
              if (j == (i)) continue;                                        \
              double d[3];                                                   \
              const double r2 = o_disp((box), (pos), (i), j, d);             \
              if (r2 < (cut2)) { __VA_ARGS__ }                                      \
            }                                                                \
          }                                                                  \
    }                                                                        \
  } while (0)

static int cmp_u64(const void* a, const void* b) {
  const uint64_t x = *(const uint64_t*)a, y = *(const uint64_t*)b;
  return (x > y) - (x < y);
}

uint64_t oracle_pairs(const mdpd_params* p, uint32_t n, const double* pos,
                      double cutoff, int32_t brute, uint32_t* out_i,
                      uint32_t* out_j, uint64_t cap) {
  o_cells cells;
  o_cells_build(p->box, cutoff, n, pos, &cells);
  if (brute) cells.brute = 1;
  const double cut2 = cutoff * cutoff;
  uint64_t cnt = 0;
  uint64_t* keys = NULL;
  uint64_t kcap = 0;
  for (uint32_t i = 0; i < n; ++i) {
    O_FOR_NEIGHBOURS(&cells, p->box, pos, n, i, cut2, {
      (void)d; (void)r2;
      if (cnt == kcap) {
        kcap = kcap ? 2 * kcap : 1024;
        keys = (uint64_t*)realloc(keys, kcap * sizeof(uint64_t));
      }
      keys[cnt++] = ((uint64_t)i << 32) | j;
    });
  }
  qsort(keys, cnt, sizeof(uint64_t), cmp_u64);
  for (uint64_t k = 0; k < cnt && k < cap; ++k) {
    out_i[k] = (uint32_t)(keys[k] >> 32);
    out_j[k] = (uint32_t)(keys[k] & 0xffffffffu);
  }
  free(keys);
  o_cells_free(&cells);
  return cnt;
}

/* ======================================================================
 * CalcDensity -- nbdf_calculator.hpp:18-59 with write-back
 * ptcl_class.hpp:53-57.  dens[3i+k]: k = type of j (n_dens 3) or 0.
 * cm[6i..6i+2] = hydrophilic CM (self + Hyphil neighbours, r^2 < 1),
 * cm[6i+3..] = hydrophobic CM (Hyphob neighbours, r^2 < rn_c^2 = 1.44);
 * j enters with its image position r_i - d.
 * ====================================================================== */
void oracle_density(const mdpd_params* p, uint32_t n, const double* pos,
                    const uint32_t* prop, double* dens, double* cm,
                    uint32_t* cm_cnt) {
  const double rn_c2 = 1.2 * 1.2;
  const double search = cm ? 1.2 : (p->rd > 1.0 ? p->rd : 1.0);
  const double rd2 = p->rd * p->rd;
  o_cells cells;
  o_cells_build(p->box, search, n, pos, &cells);
  const double cut2 = cm ? rn_c2 : (rd2 > 1.0 ? rd2 : 1.0);
#pragma omp parallel for schedule(dynamic, 256)
  for (int64_t ii = 0; ii < (int64_t)n; ++ii) {
    const uint32_t i = (uint32_t)ii;
    double d_sum[3] = {0.0, 0.0, 0.0};
    const double* ri = &pos[3 * (size_t)i];
    double hypb[3] = {0.0, 0.0, 0.0}, hypl[3] = {ri[0], ri[1], ri[2]};
    uint32_t hypb_cnt = 0, hypl_cnt = 1;
    O_FOR_NEIGHBOURS(&cells, p->box, pos, n, i, cut2, {
      if (r2 != 0.0) {
        const uint32_t pj = prop[j];
        if (cm && r2 < rn_c2 && pj == MDPD_HYPHOB) {
          for (int a = 0; a < 3; ++a) hypb[a] += ri[a] - d[a];
          hypb_cnt++;
        }
        if (r2 < rd2) {
          const double dr = sqrt(r2);
          const double w = 1.0 - dr / p->rd;
          d_sum[p->n_dens == 1 ? 0 : pj] += p->crho * w * w;
        }
        if (cm && r2 < 1.0 && pj == MDPD_HYPHIL) {
          for (int a = 0; a < 3; ++a) hypl[a] += ri[a] - d[a];
          hypl_cnt++;
        }
      }
    });
    for (int k = 0; k < 3; ++k) dens[3 * (size_t)i + k] = d_sum[k];
    if (cm) {
      for (int a = 0; a < 3; ++a) {
        cm[6 * (size_t)i + a] = hypl[a] / (double)hypl_cnt;
        cm[6 * (size_t)i + 3 + a] = hypb[a] / (double)hypb_cnt;
      }
      if (cm_cnt) {
        cm_cnt[2 * (size_t)i] = hypl_cnt;
        cm_cnt[2 * (size_t)i + 1] = hypb_cnt;
      }
    }
  }
  o_cells_free(&cells);
}

/* ======================================================================
 * CalcForceEpEpDPD -- nbdf_calculator.hpp:73-143, write-back
 * ptcl_class.hpp:48-51 (acc and press overwritten).
 * ====================================================================== */
static inline double o_pair_coef(const mdpd_params* p, uint32_t pi, uint32_t pj,
                                 const double* di, const double* dj,
                                 double dr, double drij_dvij, double rnd,
                                 double* fabs_out) {
  const double inv_dr = 1.0 / dr;
  const double rc_m_dr = 1.0 - dr;
  const double cf_pair = p->a0[pi][pj] + p->c1[pi][pj] * (dr - p->arc) * (double)(dr >= p->arc);
  double cf_mbd = 0.0;
  for (int k = 0; k < p->n_dens; ++k) cf_mbd += p->mb[pi][pj][k] * (di[k] + dj[k]);
  double wr, wd;
  if (p->envelope == 0) { wd = rc_m_dr; wr = sqrt(wd); }
  else { wr = rc_m_dr; wd = rc_m_dr * rc_m_dr; }
  double cons;
  if (p->rdm == 1.0) cons = (cf_pair + cf_mbd) * rc_m_dr;
  else {
    double wm = 1.0 - dr / p->rdm;
    if (wm < 0.0) wm = 0.0;
    cons = cf_pair * rc_m_dr + cf_mbd * wm;
  }
  const double rand_t = p->sigma[pi][pj] * wr * rnd;
  const double diss_t = p->gamma[pi][pj] * wd * drij_dvij * inv_dr;
  if (fabs_out) *fabs_out = fabs(cons) + fabs(rand_t) + fabs(diss_t);
  return (cons + rand_t - diss_t) * inv_dr;
}

static void o_force_impl(const mdpd_params* p, uint32_t n, const double* pos,
                         const double* vel, const uint32_t* id,
                         const uint32_t* prop, const double* dens,
                         uint32_t seed3, double* acc, double* press,
                         double* fabs_sum) {
  o_cells cells;
  o_cells_build(p->box, 1.0, n, pos, &cells);
#pragma omp parallel for schedule(dynamic, 256)
  for (int64_t ii = 0; ii < (int64_t)n; ++ii) {
    const uint32_t i = (uint32_t)ii;
    const double* vi = &vel[3 * (size_t)i];
    const uint32_t idi = id[i], pi = prop[i];
    double fsum[3] = {0.0, 0.0, 0.0}, psum[3] = {0.0, 0.0, 0.0}, fa = 0.0;
    O_FOR_NEIGHBOURS(&cells, p->box, pos, n, i, 1.0, {
      if (id[j] != idi) {
        const double* vj = &vel[3 * (size_t)j];
        const double dv[3] = {vi[0] - vj[0], vi[1] - vj[1], vi[2] - vj[2]};
        uint32_t m_i = idi, m_j = id[j];
        if (idi > id[j]) { m_i = id[j]; m_j = idi; }
        o_saru s = o_saru3(m_i, m_j, seed3);
        const double rnd = o_saru_nrml(&s);
        const double dr = sqrt(r2);
        const double drij_dvij = d[0] * dv[0] + d[1] * dv[1] + d[2] * dv[2];
        double fa_ij = 0.0;
        const double all_cf = o_pair_coef(p, pi, prop[j], &dens[3 * (size_t)i],
                                          &dens[3 * (size_t)j], dr, drij_dvij,
                                          rnd, fabs_sum ? &fa_ij : NULL);
        const double dF[3] = {all_cf * d[0], all_cf * d[1], all_cf * d[2]};
        const double dP[3] = {dF[0] * d[0], dF[1] * d[1], dF[2] * d[2]};
        for (int a = 0; a < 3; ++a) { fsum[a] += dF[a]; psum[a] += dP[a]; }
        fa += fa_ij;
      }
    });
    if (acc) for (int a = 0; a < 3; ++a) acc[3 * (size_t)i + a] = fsum[a];
    if (press) for (int a = 0; a < 3; ++a) press[3 * (size_t)i + a] = psum[a];
    if (fabs_sum) fabs_sum[i] = fa;
  }
  o_cells_free(&cells);
}

void oracle_force(const mdpd_params* p, uint32_t n, const double* pos,
                  const double* vel, const uint32_t* id, const uint32_t* prop,
                  const double* dens, uint32_t seed3, double* acc,
                  double* press) {
  o_force_impl(p, n, pos, vel, id, prop, dens, seed3, acc, press, NULL);
}

void oracle_force_abs(const mdpd_params* p, uint32_t n, const double* pos,
                      const double* vel, const uint32_t* id,
                      const uint32_t* prop, const double* dens, uint32_t seed3,
                      double* fabs_sum) {
  o_force_impl(p, n, pos, vel, id, prop, dens, seed3, NULL, NULL, fabs_sum);
}

/* ======================================================================
 * ForceBonded -- bdf_calculator.hpp:55-59 (MinImage with round),
 * :95-115 (bond), :117-150 (bend), :153-189 (chain of 4), :211-217
 * (serial virial sum).
 * ====================================================================== */
static inline void o_min_image_round(const double* box, double d[3]) {
  for (int a = 0; a < 3; ++a) d[a] -= box[a] * round(d[a] * (1.0 / box[a]));
}

static inline double o_dot(const double* a, const double* b) {
  return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

static void o_store_bond(double cf_s, double ibond, const double dr[3],
                         double inv_dr, double vir[3], double F[][3]) {
  const double cf_bond = cf_s * (inv_dr - ibond);
  const double Fb[3] = {cf_bond * dr[0], cf_bond * dr[1], cf_bond * dr[2]};
  for (int a = 0; a < 3; ++a) vir[a] += 2.0 * dr[a] * Fb[a];
  for (int a = 0; a < 3; ++a) { F[0][a] -= Fb[a]; F[1][a] += Fb[a]; }
}

static void o_store_bend(const double dr[][3], const double* inv_dr,
                         double vir[3], double F[][3], double cf_b) {
  const double inv_dr_prod = inv_dr[0] * inv_dr[1];
  const double inv_dist[2] = {inv_dr[0] * inv_dr[0], inv_dr[1] * inv_dr[1]};
  const double in_prod = o_dot(dr[0], dr[1]);
  const double cf_bd = cf_b * inv_dr_prod;
  const double cf_crs[2] = {in_prod * inv_dist[0], in_prod * inv_dist[1]};
  double Ftb0[3], Ftb1[3];
  for (int a = 0; a < 3; ++a) {
    Ftb0[a] = cf_bd * (dr[1][a] - cf_crs[0] * dr[0][a]);
    Ftb1[a] = cf_bd * (dr[0][a] - cf_crs[1] * dr[1][a]);
  }
  for (int a = 0; a < 3; ++a) vir[a] += 2.0 * (dr[0][a] * Ftb0[a] + dr[1][a] * Ftb1[a]);
  for (int a = 0; a < 3; ++a) {
    F[0][a] -= Ftb0[a];
    F[1][a] += Ftb0[a] - Ftb1[a];
    F[2][a] += Ftb1[a];
  }
}

void oracle_bonded(const mdpd_params* p, uint32_t n_amp, const uint32_t* topo,
                   const double* cf_bend, const double* pos, double* acc,
                   double vir[3]) {
  const double ibond = (p->bond_leng != 0.0) ? 1.0 / p->bond_leng : 0.0;
  double v[3] = {0.0, 0.0, 0.0};
  for (uint32_t a = 0; a < n_amp; ++a) {
    const uint32_t* l = &topo[4 * (size_t)a];
    double Fbb[4][3] = {{0}}, dr[3][3], inv_dr[3];
    for (int u = 1; u < 4; ++u) {
      for (int c = 0; c < 3; ++c)
        dr[u - 1][c] = pos[3 * (size_t)l[u] + c] - pos[3 * (size_t)l[u - 1] + c];
      o_min_image_round(p->box, dr[u - 1]);
      const double dist2 = o_dot(dr[u - 1], dr[u - 1]);
      inv_dr[u - 1] = 1.0 / sqrt(dist2);
      o_store_bond(p->cf_s, ibond, dr[u - 1], inv_dr[u - 1], v, &Fbb[u - 1]);
      if (u >= 2) o_store_bend(&dr[u - 2], &inv_dr[u - 2], v, &Fbb[u - 2], cf_bend[a]);
    }
    for (int u = 0; u < 4; ++u)
      for (int c = 0; c < 3; ++c) acc[3 * (size_t)l[u] + c] += Fbb[u][c];
  }
  if (vir) for (int c = 0; c < 3; ++c) vir[c] = v[c];
}

/* ======================================================================
 * Integrator -- driftkick.hpp:5-30 (drift_and_predict), :32-40 (kick),
 * :42-54 (remove_cmdrift_global).
 * ====================================================================== */
void oracle_drift(const mdpd_params* p, uint32_t n, double* pos, double* vel,
                  double* vel_buf, const double* acc, double* delta) {
  const double dt = p->dt;
  const double ib[3] = {1.0 / p->box[0], 1.0 / p->box[1], 1.0 / p->box[2]};
#pragma omp parallel for schedule(static)
  for (int64_t ii = 0; ii < (int64_t)n; ++ii) {
    const size_t i = (size_t)ii;
    for (int a = 0; a < 3; ++a) {
      const double dr = dt * (vel[3 * i + a] + 0.5 * dt * acc[3 * i + a]);
      const double temp_r = pos[3 * i + a] + dr;
      delta[3 * i + a] += dr;
      pos[3 * i + a] = temp_r - floor(temp_r * ib[a]) * p->box[a];
      vel_buf[3 * i + a] = vel[3 * i + a] + 0.5 * dt * acc[3 * i + a];
      vel[3 * i + a] += 0.5 * dt * acc[3 * i + a];
    }
  }
}

void oracle_kick(const mdpd_params* p, uint32_t n, double* vel,
                 const double* vel_buf, const double* acc) {
  const double dt = p->dt;
#pragma omp parallel for schedule(static)
  for (int64_t ii = 0; ii < (int64_t)n; ++ii)
    for (int a = 0; a < 3; ++a)
      vel[3 * ii + a] = vel_buf[3 * ii + a] + dt * 0.5 * acc[3 * ii + a];
}

void oracle_remove_cm_drift(uint32_t n, double* vel) {
  double s[3] = {0.0, 0.0, 0.0};
  for (uint32_t i = 0; i < n; ++i)
    for (int a = 0; a < 3; ++a) s[a] += vel[3 * (size_t)i + a];
  for (int a = 0; a < 3; ++a) s[a] = s[a] / (double)n;
  for (uint32_t i = 0; i < n; ++i)
    for (int a = 0; a < 3; ++a) vel[3 * (size_t)i + a] -= s[a];
}

/* ======================================================================
 * Observer reductions -- observer.hpp:82-102 (KineticTempera, serial),
 * :108-130 (Pressure), :132-178 (Diffusion, serial; the solvent sum
 * starts at bead index amp_num, as at :171).
 * ====================================================================== */
void oracle_observe(const mdpd_params* p, uint32_t n, uint32_t amp_num,
                    uint32_t sol_num, const double* vel, const double* press,
                    const double* delta, const double bonded_vir[3],
                    mdpd_observables* o) {
  memset(o, 0, sizeof(*o));
  double k[3] = {0.0, 0.0, 0.0};
  for (uint32_t i = 0; i < n; ++i)
    for (int a = 0; a < 3; ++a) k[a] += vel[3 * (size_t)i + a] * vel[3 * (size_t)i + a];
  for (int a = 0; a < 3; ++a) k[a] /= n;
  for (int a = 0; a < 3; ++a) o->kin_temp[a] = k[a];
  o->kin_temp[3] = (k[0] + k[1] + k[2]) / 3.0;

  double ps[3] = {0.0, 0.0, 0.0}, pv[3] = {0.0, 0.0, 0.0};
  for (uint32_t i = 0; i < n; ++i) {
    for (int a = 0; a < 3; ++a) {
      ps[a] += press[3 * (size_t)i + a] * 0.5;
      pv[a] += press[3 * (size_t)i + a];
    }
    for (int a = 0; a < 3; ++a) ps[a] += vel[3 * (size_t)i + a] * vel[3 * (size_t)i + a];
  }
  for (int a = 0; a < 3; ++a) ps[a] += bonded_vir[a] * 0.5;
  const double iv = (1.0 / p->box[0]) * (1.0 / p->box[1]) * (1.0 / p->box[2]);
  for (int a = 0; a < 3; ++a) {
    o->pressure[a] = ps[a] * iv;
    o->bonded_vir[a] = bonded_vir[a];
    o->pair_vir[a] = pv[a];
  }

  double dl = 0.0;
  for (uint32_t i = 0; i < n; ++i) dl += o_dot(&delta[3 * (size_t)i], &delta[3 * (size_t)i]);
  dl /= n;
  double dm = 0.0;
  for (uint32_t a = 0; a < amp_num; ++a) {
    double s[3] = {0.0, 0.0, 0.0};
    for (int u = 0; u < 4; ++u)
      for (int c = 0; c < 3; ++c) s[c] += delta[3 * (4 * (size_t)a + u) + c];
    for (int c = 0; c < 3; ++c) s[c] /= 4;
    dm += o_dot(s, s);
  }
  dm = amp_num ? dm / amp_num : 0.0;
  double ds = 0.0;
  for (uint32_t i = amp_num; i < n; ++i) ds += o_dot(&delta[3 * (size_t)i], &delta[3 * (size_t)i]);
  ds = sol_num ? ds / sol_num : 0.0;
  o->diffusion[0] = dl;
  o->diffusion[1] = dm;
  o->diffusion[2] = ds;
  o->n_beads = n;
  o->n_amp = amp_num;
}

/* ======================================================================
 * Head conversion (no reference counterpart): a unit-0 bead of type
 * conv_from becomes conv_to when Saru(amp_id, time, chem_seed).uni_open()
 * <= conv_p.
 * ====================================================================== */
uint32_t oracle_head_conversion(const mdpd_params* p, uint32_t n,
                                const uint32_t* amp_id, const uint32_t* unit,
                                uint32_t* prop, uint32_t time) {
  uint32_t cnt = 0;
  for (uint32_t i = 0; i < n; ++i) {
    if (unit[i] != 0 || prop[i] != (uint32_t)p->conv_from) continue;
    o_saru s = o_saru3(amp_id[i], time, p->chem_seed);
    if (o_saru_uni(&s) <= p->conv_p) { prop[i] = (uint32_t)p->conv_to; cnt++; }
  }
  return cnt;
}

/* ======================================================================
 * Whole system.
 * ====================================================================== */
/* ----------------------------------------------------------------------
 * MT19937, the generator behind the reference's PS::MT (FDPS wraps the
 * mt19937ar functions init_genrand / genrand_int32 / genrand_real1 /
 * genrand_real3, chemmanager.hpp:19,74,154,167,178).  Restated from the
 * published algorithm (M. Matsumoto and T. Nishimura, "Mersenne Twister: a
 * 623-dimensionally equidistributed uniform pseudo-random number
 * generator", ACM TOMACS 8 (1998) 3-30, with the 2002 initialisation):
 * state of 624 words, twist with the matrix word 0x9908b0df, tempering by
 * (11, 7 & 0x9d2c5680, 15 & 0xefc60000, 18).
 * ---------------------------------------------------------------------- */
typedef struct {
  uint32_t mt[624];
  int mti;
} o_mt;

static void o_mt_init(o_mt* g, uint32_t seed) {
  g->mt[0] = seed;
  for (int i = 1; i < 624; ++i)
    g->mt[i] = 1812433253u * (g->mt[i - 1] ^ (g->mt[i - 1] >> 30)) + (uint32_t)i;
  g->mti = 624;
}

static uint32_t o_mt_u32(o_mt* g) {
  if (g->mti >= 624) {
    for (int k = 0; k < 624; ++k) {
      const uint32_t y = (g->mt[k] & 0x80000000u) | (g->mt[(k + 1) % 624] & 0x7fffffffu);
      g->mt[k] = g->mt[(k + 397) % 624] ^ (y >> 1) ^ ((y & 1u) ? 0x9908b0dfu : 0u);
    }
    g->mti = 0;
  }
  uint32_t y = g->mt[g->mti++];
  y ^= y >> 11;
  y ^= (y << 7) & 0x9d2c5680u;
  y ^= (y << 15) & 0xefc60000u;
  y ^= y >> 18;
  return y;
}
/* genrand_real1: [0, 1]; genrand_real3: (0, 1) */
static double o_mt_real1(o_mt* g) { return (double)o_mt_u32(g) * (1.0 / 4294967295.0); }
static double o_mt_real3(o_mt* g) { return ((double)o_mt_u32(g) + 0.5) * (1.0 / 4294967296.0); }

void oracle_mt_u32(uint32_t seed, uint32_t n, uint32_t* out) {
  o_mt g;
  o_mt_init(&g, seed);
  for (uint32_t k = 0; k < n; ++k) out[k] = o_mt_u32(&g);
}

struct oracle_sys {
  mdpd_params p;
  uint32_t n, cap, n_amp, amp_cap;
  double *pos, *vel, *vel_buf, *acc, *press, *dens, *delta, *cm;
  uint32_t *id, *prop, *amp_id, *unit, *topo;
  double* cf_bend;
  uint32_t time;
  double bonded_vir[3];
  /* reference-MT mode of the chemistry: one MT19937 stream consumed in the
     reference's order (ChemManager(seed)); mt_draws counts its outputs */
  int mt_on;
  o_mt mt;
  uint64_t mt_draws;
};

void oracle_sys_set_mt(oracle_sys* s, uint32_t seed, int32_t enable) {
  s->mt_on = enable != 0;
  o_mt_init(&s->mt, seed);
  s->mt_draws = 0;
}
uint64_t oracle_sys_mt_draws(const oracle_sys* s) { return s->mt_draws; }

oracle_sys* oracle_sys_new(const mdpd_params* p, uint32_t cap, uint32_t amp_cap) {
  oracle_sys* s = (oracle_sys*)calloc(1, sizeof(oracle_sys));
  s->p = *p;
  s->cap = cap;
  s->amp_cap = amp_cap;
  const size_t v3 = 3 * (size_t)(cap ? cap : 1) * sizeof(double);
  s->pos = (double*)calloc(1, v3); s->vel = (double*)calloc(1, v3);
  s->vel_buf = (double*)calloc(1, v3); s->acc = (double*)calloc(1, v3);
  s->press = (double*)calloc(1, v3); s->dens = (double*)calloc(1, v3);
  s->delta = (double*)calloc(1, v3); s->cm = (double*)calloc(2, v3);
  const size_t u1 = (size_t)(cap ? cap : 1) * sizeof(uint32_t);
  s->id = (uint32_t*)calloc(1, u1); s->prop = (uint32_t*)calloc(1, u1);
  s->amp_id = (uint32_t*)calloc(1, u1); s->unit = (uint32_t*)calloc(1, u1);
  s->topo = (uint32_t*)calloc(4 * (size_t)(amp_cap ? amp_cap : 1), sizeof(uint32_t));
  s->cf_bend = (double*)calloc((size_t)(amp_cap ? amp_cap : 1), sizeof(double));
  return s;
}

void oracle_sys_free(oracle_sys* s) {
  if (!s) return;
  free(s->pos); free(s->vel); free(s->vel_buf); free(s->acc); free(s->press);
  free(s->dens); free(s->delta); free(s->cm); free(s->id); free(s->prop);
  free(s->amp_id); free(s->unit); free(s->topo); free(s->cf_bend);
  free(s);
}

int oracle_sys_set(oracle_sys* s, uint32_t n, const double* pos,
                   const double* vel, const double* vel_buf, const double* acc,
                   const uint32_t* id, const uint32_t* prop,
                   const uint32_t* amp_id, const uint32_t* unit) {
  if (n > s->cap) return -1;
  s->n = n;
  const size_t v3 = 3 * (size_t)n * sizeof(double), u1 = (size_t)n * sizeof(uint32_t);
  memcpy(s->pos, pos, v3);
  memcpy(s->vel, vel, v3);
  if (vel_buf) memcpy(s->vel_buf, vel_buf, v3); else memcpy(s->vel_buf, vel, v3);
  if (acc) memcpy(s->acc, acc, v3); else memset(s->acc, 0, v3);
  memset(s->delta, 0, v3);
  memset(s->press, 0, v3);
  memset(s->dens, 0, v3);
  memcpy(s->id, id, u1);
  memcpy(s->prop, prop, u1);
  if (amp_id) memcpy(s->amp_id, amp_id, u1); else memset(s->amp_id, 0xff, u1);
  if (unit) memcpy(s->unit, unit, u1); else memset(s->unit, 0xff, u1);
  return 0;
}

int oracle_sys_set_topology(oracle_sys* s, uint32_t n_amp, const uint32_t* topo,
                            const uint32_t* core_amp_id, uint32_t n_core) {
  if (n_amp > s->amp_cap) return -1;
  s->n_amp = n_amp;
  memcpy(s->topo, topo, 4 * (size_t)n_amp * sizeof(uint32_t));
  for (uint32_t a = 0; a < n_amp; ++a) {  /* bdf_calculator.hpp:77-86 */
    s->cf_bend[a] = s->p.cf_b;
    for (uint32_t c = 0; c < n_core; ++c)
      if (a == core_amp_id[c]) s->cf_bend[a] = s->p.cf_b_rigid;
  }
  return 0;
}

void oracle_sys_set_clock(oracle_sys* s, uint32_t time) { s->time = time; }

uint32_t oracle_sys_get(const oracle_sys* s, double* pos, double* vel,
                        double* vel_buf, double* acc, double* press,
                        double* dens, double* delta, uint32_t* id,
                        uint32_t* prop, uint32_t* amp_id, uint32_t* unit) {
  const size_t v3 = 3 * (size_t)s->n * sizeof(double), u1 = (size_t)s->n * sizeof(uint32_t);
  if (pos) memcpy(pos, s->pos, v3);
  if (vel) memcpy(vel, s->vel, v3);
  if (vel_buf) memcpy(vel_buf, s->vel_buf, v3);
  if (acc) memcpy(acc, s->acc, v3);
  if (press) memcpy(press, s->press, v3);
  if (dens) memcpy(dens, s->dens, v3);
  if (delta) memcpy(delta, s->delta, v3);
  if (id) memcpy(id, s->id, u1);
  if (prop) memcpy(prop, s->prop, u1);
  if (amp_id) memcpy(amp_id, s->amp_id, u1);
  if (unit) memcpy(unit, s->unit, u1);
  return s->n;
}

uint32_t oracle_sys_num_amp(const oracle_sys* s) { return s->n_amp; }
void oracle_sys_get_topology(const oracle_sys* s, uint32_t* topo) {
  memcpy(topo, s->topo, 4 * (size_t)s->n_amp * sizeof(uint32_t));
}

void oracle_sys_drift(oracle_sys* s) {
  oracle_drift(&s->p, s->n, s->pos, s->vel, s->vel_buf, s->acc, s->delta);
}
void oracle_sys_density(oracle_sys* s) {
  oracle_density(&s->p, s->n, s->pos, s->prop, s->dens, NULL, NULL);
}
void oracle_sys_force(oracle_sys* s) {
  oracle_force(&s->p, s->n, s->pos, s->vel, s->id, s->prop, s->dens,
               s->p.pair_seed + s->time, s->acc, s->press);
}
void oracle_sys_bonded(oracle_sys* s, double vir[3]) {
  oracle_bonded(&s->p, s->n_amp, s->topo, s->cf_bend, s->pos, s->acc, s->bonded_vir);
  if (vir) for (int a = 0; a < 3; ++a) vir[a] = s->bonded_vir[a];
}
void oracle_sys_kick(oracle_sys* s) { oracle_kick(&s->p, s->n, s->vel, s->vel_buf, s->acc); }
void oracle_sys_remove_cm_drift(oracle_sys* s) { oracle_remove_cm_drift(s->n, s->vel); }
uint32_t oracle_sys_head_conversion(oracle_sys* s) {
  return oracle_head_conversion(&s->p, s->n, s->amp_id, s->unit, s->prop, s->time);
}

void oracle_sys_observe(oracle_sys* s, uint32_t sol_num, mdpd_observables* o) {
  oracle_observe(&s->p, s->n, s->n_amp, sol_num, s->vel, s->press, s->delta,
                 s->bonded_vir, o);
  o->time = s->time;
}

/* ======================================================================
 * ChemManager::RandomChemEvent, serial (chemmanager.hpp:244-281) with
 * DetermineTargetIdWithCorePos (:116-164), GetCorePos (:84-114),
 * GetMembraneTangRandomVec (:166-173), CreateAmpPart (:35-67) and
 * GenThermalVeloc (:18-20).  PS::MT draws are replaced by keyed Saru
 * streams (cone test: (amp j, time, chem_seed + TAG_CONE*(core+1)); tangent
 * angle: (new amp id, time, chem_seed ^ TAG_TANG); velocities: one stream
 * (new bead id, time, chem_seed ^ TAG_VEL), nrml() per component).  New
 * amphiphiles are appended to the topology (the MPI path's semantics,
 * bdf_calculator.hpp:442-543).
 * ====================================================================== */
static inline void o_normalize(double v[3]) {
  const double l = sqrt(o_dot(v, v));
  for (int a = 0; a < 3; ++a) v[a] /= l;
}
static inline void o_apply_pbc(const double* box, double v[3]) {
  for (int a = 0; a < 3; ++a) v[a] -= floor(v[a] * (1.0 / box[a])) * box[a];
}

int oracle_sys_chem_event(oracle_sys* s, const mdpd_chem_params* cp, int32_t* stop) {
  /* neighbour CMs of the core heads from the current positions (the density
     pass of the same step ran on them) */
  const double* box = s->p.box;
  uint32_t ntarget = 0, tcap = 64;
  uint32_t* tgt = (uint32_t*)malloc(2 * tcap * sizeof(uint32_t));
  double* cmall = NULL;
  if (cp->n_core) {
    cmall = (double*)malloc(6 * (size_t)s->n * sizeof(double));
    double* dtmp = (double*)malloc(3 * (size_t)s->n * sizeof(double));
    oracle_density(&s->p, s->n, s->pos, s->prop, dtmp, cmall, NULL);
    free(dtmp);
  }
  for (uint32_t c = 0; c < cp->n_core; ++c) {
    double ch[3] = {-1.0, -1.0, -1.0}, ct[3] = {-1.0, -1.0, -1.0}, ccm[3] = {-1.0, -1.0, -1.0};
    for (uint32_t i = 0; i < s->n; ++i)
      if (s->id[i] == cp->core_ptcl_id[c]) {
        for (int a = 0; a < 3; ++a) {
          ch[a] = s->pos[3 * (size_t)i + a];
          ct[a] = cmall[6 * (size_t)i + 3 + a];
          ccm[a] = cmall[6 * (size_t)i + a];
        }
        break;
      }
    double h2t[3] = {ct[0] - ch[0], ct[1] - ch[1], ct[2] - ch[2]};
    o_min_image_round(box, h2t);
    o_normalize(h2t);
    double c2cm[3] = {ccm[0] - ch[0], ccm[1] - ch[1], ccm[2] - ch[2]};
    o_min_image_round(box, c2cm);
    const double prj = o_dot(c2cm, h2t);
    for (int a = 0; a < 3; ++a) ch[a] += prj * h2t[a];
    o_apply_pbc(box, ch);
    for (uint32_t j = 0; j < s->n_amp; ++j) {
      const uint32_t head = s->topo[4 * (size_t)j];
      double rnd = 2.0;
      double c2p[3] = {s->pos[3 * (size_t)head] - ch[0], s->pos[3 * (size_t)head + 1] - ch[1],
                       s->pos[3 * (size_t)head + 2] - ch[2]};
      o_min_image_round(box, c2p);
      const double depth = o_dot(h2t, c2p);
      double tang[3];
      for (int a = 0; a < 3; ++a) tang[a] = c2p[a] - depth * h2t[a];
      const double rad = sqrt(o_dot(tang, tang));
      const double rad_thr = cp->influ_grd * depth + cp->influ_rad;
      if ((rad <= rad_thr) && (depth >= -cp->influ_hei) && (depth <= cp->influ_dep)) {
        if (s->mt_on) {
          rnd = o_mt_real1(&s->mt);  /* chemmanager.hpp:153-154 */
          s->mt_draws++;
        } else {
          o_saru sr = o_saru3(j, s->time, s->p.chem_seed + TAG_CONE * (c + 1));
          rnd = o_saru_uni(&sr);
        }
      }
      if (rnd <= cp->p_thresld) {
        if (ntarget == tcap) { tcap *= 2; tgt = (uint32_t*)realloc(tgt, 2 * tcap * sizeof(uint32_t)); }
        tgt[2 * ntarget] = head;
        tgt[2 * ntarget + 1] = s->topo[4 * (size_t)j + 2];
        ntarget++;
      }
    }
  }
  free(cmall);
  if (s->n + 4 * ntarget > s->cap || s->n_amp + ntarget > s->amp_cap) {
    free(tgt);
    return -1;
  }
  uint32_t new_id = s->n;
  const uint32_t amp0 = s->n_amp;
  for (uint32_t t = 0; t < ntarget; ++t) {
    const uint32_t hd = tgt[2 * t], tl = tgt[2 * t + 1];
    double h2e[3] = {s->pos[3 * (size_t)tl] - s->pos[3 * (size_t)hd],
                     s->pos[3 * (size_t)tl + 1] - s->pos[3 * (size_t)hd + 1],
                     s->pos[3 * (size_t)tl + 2] - s->pos[3 * (size_t)hd + 2]};
    o_min_image_round(box, h2e);
    o_normalize(h2e);
    const uint32_t new_amp = amp0 + t;
    double theta;
    if (s->mt_on) {
      theta = 2.0 * M_PI * o_mt_real1(&s->mt);  /* GetMembraneTangRandomVec, :167 */
      s->mt_draws++;
    } else {
      o_saru st = o_saru3(new_amp, s->time, s->p.chem_seed ^ TAG_TANG);
      theta = 2.0 * M_PI * o_saru_uni(&st);
    }
    double base[3] = {cos(theta), sin(theta), 0.0};
    const double bd = o_dot(base, h2e);
    double tv[3];
    for (int a = 0; a < 3; ++a) tv[a] = base[a] - h2e[a] * bd;
    o_normalize(tv);
    for (int a = 0; a < 3; ++a) { tv[a] *= cp->eps; h2e[a] *= s->p.bond_leng; }
    double cpos[3];
    for (int a = 0; a < 3; ++a) cpos[a] = tv[a] + s->pos[3 * (size_t)hd + a];
    for (uint32_t u = 0; u < 4; ++u) {
      const size_t i = s->n;
      s->id[i] = new_id;
      s->prop[i] = (u == 0) ? MDPD_HYPHIL : MDPD_HYPHOB;
      s->amp_id[i] = new_amp;
      s->unit[i] = u;
      double np_[3];
      for (int a = 0; a < 3; ++a) np_[a] = cpos[a] + h2e[a] * (double)u;
      o_apply_pbc(box, np_);
      o_saru sv = o_saru3(new_id, s->time, s->p.chem_seed ^ TAG_VEL);
      for (int a = 0; a < 3; ++a) {
        s->pos[3 * i + a] = np_[a];
        if (s->mt_on) {
          /* GenThermalVeloc (:18-20), T = 1: sqrt(-2 log u) cos(2 pi u') with
             u, u' two genrand_real3 draws; C++ leaves the order of the two
             calls in that expression open, and g++ (the reference's
             compiler, at its Makefile flags) takes the log's first, as here */
          const double u = o_mt_real3(&s->mt), u2 = o_mt_real3(&s->mt);
          s->mt_draws += 2;
          s->vel[3 * i + a] = sqrt(-2.0 * log(u)) * cos(2.0 * M_PI * u2);
        } else {
          s->vel[3 * i + a] = o_saru_nrml(&sv);  /* sqrt(-2 T log u1) cos(2 pi u2), T = 1 */
        }
        s->vel_buf[3 * i + a] = 0.0;
        s->acc[3 * i + a] = 0.0;
        s->press[3 * i + a] = 0.0;
        s->dens[3 * i + a] = 0.0;
        s->delta[3 * i + a] = 0.0;
      }
      s->topo[4 * (size_t)new_amp + u] = (uint32_t)i;
      s->n++;
      new_id++;
    }
    s->cf_bend[new_amp] = s->p.cf_b;  /* AppendCfBend, bdf_calculator.hpp:21-26 */
  }
  s->n_amp += ntarget;
  free(tgt);
  *stop = (s->n_amp >= cp->max_amp_num) ? 1 : 0;
  return (int)ntarget;
}

int oracle_sys_run(oracle_sys* s, uint32_t nsteps) {
  const uint32_t cm_corr = s->p.cm_corr ? s->p.cm_corr : 100u;
  for (uint32_t k = 0; k < nsteps; ++k) {
    oracle_sys_drift(s);
    oracle_sys_density(s);
    oracle_sys_force(s);
    oracle_sys_bonded(s, NULL);
    oracle_sys_kick(s);
    if (s->time % cm_corr == 0) oracle_sys_remove_cm_drift(s);
    if (s->p.conv_p > 0.0) oracle_sys_head_conversion(s);
    s->time++;
  }
  return 0;
}

int oracle_num_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

void oracle_set_num_threads(int n) {
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
#else
  (void)n;
#endif
}

/* ======================================================================
 * Initial configurations -- src/config_maker.cpp, walked sequentially as
 * the reference does: GenConfig (:342-384), MakeFlatSheet (:103-138, flap
 * toggled amphiphile by amphiphile), MakeSphSheet / MakeLineForEachTheta /
 * MakeSphLine (:139-213), MakeCylindSheet / MakeLineForEachAxis
 * (:215-262, including their missing guard: a pass entered when full lays
 * one amphiphile past the end and the run fails), MakeRandomConfig
 * (:264-280), MakeCuboidSolvent / MakeSphericalSolvent (:282-321), GenVeloc
 * with RemoveCMDrift (:51-72, plain sequential sums) and
 * CheckParticleConfigIsValid (:420-452).  The MT19937 draws become keyed
 * Saru streams, one per amphiphile / solvent bead / velocity:
 * Saru(index, tag, seed), a uniform from the top 26 bits of two draws.
 * ====================================================================== */
#define CFG_TAG_AMP 0x414D5048u
#define CFG_TAG_SOL 0x534F4C56u
#define CFG_TAG_VEL 0x56454C4Fu
#define CFG_SOL_MAX_ATTEMPTS (1u << 16)

static inline double o_cfg_uni(o_saru* s, int open) {
  const uint32_t a = o_saru_u32(s) >> 6, b = o_saru_u32(s) >> 6;
  const double k = (double)a * 67108864.0 + (double)b;
  return (open ? k + 0.5 : k) * (1.0 / 4503599627370496.0);
}

typedef struct {
  const mdpd_config_spec* s;
  uint32_t all_unit, end_idx, idx, n;
  double *pos;
  uint32_t *id, *prop, *amp_id, *unit;
} o_cfg;

static void o_cfg_pbc(const double* box, double* p) {
  for (int a = 0; a < 3; ++a) p[a] -= floor(p[a] / box[a]) * box[a];
}

/* SetAmphilPartPos :79-94; beads past the array (the axis passes'
   overshoot) are counted, not written */
static void o_cfg_amph(o_cfg* c, const double base[3], const double nv[3]) {
  for (uint32_t u = 0; u < c->all_unit; ++u) {
    double p[3];
    for (int a = 0; a < 3; ++a) p[a] = base[a] + (c->s->bond_leng * u) * nv[a];
    o_cfg_pbc(c->s->box, p);
    const uint32_t i = c->idx;
    if (i < c->n) {
      memcpy(c->pos + 3 * (size_t)i, p, sizeof(p));
      c->id[i] = i;
      c->unit[i] = u;
      c->amp_id[i] = i / c->all_unit;
      c->prop[i] = u < c->s->head_unit ? MDPD_HYPHIL : MDPD_HYPHOB;
    }
    c->idx++;
  }
}

/* MakeSphLine :139-161; returns 0 once the amphiphiles are all laid */
static int o_cfg_sph_line(o_cfg* c, double the, const double center[3], double rad, double sign) {
  const double q_thick = 0.5 * c->all_unit * c->s->bond_leng;
  const double prj_rad = rad * sin(the);
  double d_phi = c->s->lip_len / prj_rad;
  const int32_t phi_elem = (int32_t)(2.0 * M_PI / d_phi);
  d_phi = 2.0 * M_PI / phi_elem;
  for (int32_t i = 0; i < phi_elem; i++) {
    const double phi = d_phi * i;
    const double nv[3] = {sign * sin(the) * cos(phi), sign * sin(the) * sin(phi), sign * cos(the)};
    double base[3];
    for (int a = 0; a < 3; ++a) base[a] = sign * nv[a] * (rad - sign * q_thick) + center[a];
    o_cfg_amph(c, base, nv);
    if (c->idx >= c->end_idx) return 0;
  }
  return 1;
}

static void o_cfg_theta_pass(o_cfg* c, int32_t elem, double d_the, double offset, double rad,
                             double sign) {
  if (c->idx >= c->end_idx) return;
  const double cent[3] = {0.5 * c->s->box[0], 0.5 * c->s->box[1], 0.5 * c->s->box[2]};
  for (int32_t i = 0; i < elem; i++)
    if (!o_cfg_sph_line(c, d_the * (i + offset), cent, rad, sign)) return;
}

static void o_cfg_axis_pass(o_cfg* c, int32_t elem, double rad, double sign) {
  const double h_pi = acos(0.0);
  double cent[3] = {0.5 * c->s->box[0], 0.5 * c->s->box[1], 0.5 * c->s->box[2]};
  for (int32_t i = 0; i < elem; i++) {
    cent[2] = i * c->s->lip_len;
    if (!o_cfg_sph_line(c, h_pi, cent, rad, sign)) return;
  }
}

/* the count check of the ring passes; elem from a double as the reference's
   static_cast<S32> (finite, in range by the checks below) */
static int32_t o_cfg_elem(double q) { return (q > 0.0 && q < 2147483647.0) ? (int32_t)q : 0; }

int oracle_config_count(const mdpd_config_spec* s, uint32_t* n) {
  const uint64_t all_unit = (uint64_t)s->head_unit + s->tail_unit;
  const uint64_t tot = all_unit * s->amp_num + (s->with_solvent ? s->sol_num : 0u);
  if (tot == 0 || tot > 0xffffffffu) return MDPD_ERR_ARG;
  *n = (uint32_t)tot;
  return MDPD_OK;
}

int oracle_make_config(const mdpd_config_spec* s, uint32_t n, double* pos, double* vel,
                       uint32_t* id, uint32_t* prop, uint32_t* amp_id, uint32_t* unit,
                       double* kin_temp) {
  uint32_t nn = 0;
  if (oracle_config_count(s, &nn) != MDPD_OK || nn != n) return MDPD_ERR_ARG;
  o_cfg c = {s, s->head_unit + s->tail_unit, 0, 0, n, pos, id, prop, amp_id, unit};
  const uint32_t amp_ptcl = c.all_unit * s->amp_num;
  c.end_idx = amp_ptcl;
  const double* L = s->box;
  if (s->with_solvent && (s->mode == MDPD_CFG_RANDOM || s->mode == MDPD_CFG_ASYM_FLAT))
    return MDPD_ERR_CONFIG;
  int32_t sol_dim = 0;
  double sol_rad = 0.0;
  if (s->mode == MDPD_CFG_FLAT || s->mode == MDPD_CFG_ASYM_FLAT) {
    /* MakeFlatSheet, axis = 1 */
    const int axis = 1;
    const uint32_t rat = s->mode == MDPD_CFG_ASYM_FLAT ? (uint32_t)(100 * s->in_out_rat) : 0xffffffffu;
    const double eps = 1.0e-5;
    double nv[3] = {0.0, 0.0, 0.0};
    int flap = 1;
    uint32_t a = 0;
    while (c.idx < c.end_idx) {
      o_saru r = o_saru3(a, CFG_TAG_AMP, s->seed);
      double base[3];
      for (int k = 0; k < 3; ++k) base[k] = L[k] * o_cfg_uni(&r, 1);
      const double sign = flap ? 1.0 : -1.0;
      base[axis] = 0.5 * L[axis] + ((c.all_unit - 1) * s->bond_leng + eps) * sign;
      nv[axis] = -sign;
      o_cfg_amph(&c, base, nv);
      a++;
      if (s->mode == MDPD_CFG_ASYM_FLAT) {
        const uint32_t next = c.idx / c.all_unit;
        if (((next % 100) == rat) || ((next % 100) == 0 && (next != 0))) flap ^= 1;
      } else {
        flap ^= 1;
      }
    }
  } else if (s->mode == MDPD_CFG_RANDOM) {
    uint32_t a = 0;
    while (c.idx < c.end_idx) {
      o_saru r = o_saru3(a, CFG_TAG_AMP, s->seed);
      double base[3], nv[3];
      for (int k = 0; k < 3; ++k) base[k] = L[k] * o_cfg_uni(&r, 0);
      for (int k = 0; k < 3; ++k) nv[k] = o_cfg_uni(&r, 0);
      const double norm_nv = sqrt(nv[0] * nv[0] + nv[1] * nv[1] + nv[2] * nv[2]);
      for (int k = 0; k < 3; ++k) nv[k] /= norm_nv;
      o_cfg_amph(&c, base, nv);
      a++;
    }
  } else if (s->mode == MDPD_CFG_SPHERE) {
    const double min_box = fmin(fmin(L[0], L[1]), L[2]);
    if (!(s->sph_rad >= 0.0 && s->sph_rad <= 0.5 * min_box)) return MDPD_ERR_CONFIG;
    const double up_the = s->upper_the * M_PI / 180.0;
    const double q_thick = 0.5 * c.all_unit * s->bond_leng;
    const double out_rad = s->sph_rad + q_thick, in_rad = s->sph_rad - q_thick;
    double d_the_out = s->lip_len / out_rad, d_the_in = s->lip_len / in_rad;
    const int32_t out_elem = o_cfg_elem(up_the / d_the_out), in_elem = o_cfg_elem(up_the / d_the_in);
    d_the_out = up_the / out_elem;
    d_the_in = up_the / in_elem;
    o_cfg_theta_pass(&c, out_elem, d_the_out, 0.0, out_rad, -1.0);
    o_cfg_theta_pass(&c, in_elem, d_the_in, 0.0, in_rad, 1.0);
    const uint32_t residue = c.end_idx - c.idx;
    if (residue > 0) {
      o_cfg_theta_pass(&c, in_elem, d_the_in, 0.5, in_rad, 1.0);
      if (c.end_idx != c.idx) return MDPD_ERR_CONFIG;
    }
    sol_dim = 3;
    sol_rad = s->sph_rad;
  } else if (s->mode == MDPD_CFG_CYLIND) {
    const int axis = 2;
    const double q_thick = 0.5 * c.all_unit * s->bond_leng;
    const double out_rad = s->cyl_r + q_thick, in_rad = s->cyl_r - q_thick;
    const double d_the_in = s->lip_len / in_rad;
    const int32_t in_elem = o_cfg_elem(M_PI / d_the_in);
    if (!(s->cyl_l >= 0.0 && s->cyl_l <= L[axis])) return MDPD_ERR_CONFIG;
    if (!(s->cyl_r >= 0.0 && (2.0 * s->cyl_r + c.all_unit * s->bond_leng) < L[axis])) return MDPD_ERR_CONFIG;
    const int32_t z_elem = o_cfg_elem(L[axis] / s->lip_len);
    o_cfg_axis_pass(&c, z_elem, out_rad, -1.0);
    o_cfg_axis_pass(&c, z_elem, in_rad, 1.0);
    const uint32_t residue = c.end_idx - c.idx;
    if (residue > 0) {
      o_cfg_axis_pass(&c, in_elem, out_rad, -1.0);
      if (c.end_idx != c.idx) return MDPD_ERR_CONFIG;
    }
    sol_dim = 2;
    sol_rad = s->cyl_r;
  } else {
    return MDPD_ERR_ARG;
  }
  /* solvent (:282-321) */
  if (s->with_solvent) {
    const double h_thick = 0.5 * c.all_unit * s->bond_leng;
    for (uint32_t k = 0; k < s->sol_num; ++k) {
      const uint32_t i = amp_ptcl + k;
      o_saru r = o_saru3(k, CFG_TAG_SOL, s->seed);
      double p[3];
      if (s->mode == MDPD_CFG_FLAT) {
        double up[3] = {L[0], L[1], L[2]};
        const double dw[3] = {0.0, 0.0, 0.0};
        up[1] *= 0.5;
        for (int a = 0; a < 3; ++a) p[a] = (up[a] - dw[a]) * o_cfg_uni(&r, 1) + dw[a];
      } else {
        uint32_t t = 0;
        for (;; ++t) {
          if (t == CFG_SOL_MAX_ATTEMPTS) return MDPD_ERR_CONFIG;
          double d[3];
          for (int a = 0; a < 3; ++a) p[a] = L[a] * o_cfg_uni(&r, 1);
          for (int a = 0; a < 3; ++a) d[a] = p[a] - 0.5 * L[a];
          double dist = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
          if (sol_dim == 2) dist -= d[2] * d[2];
          dist = sqrt(dist);
          if (dist < sol_rad - h_thick) break;
        }
      }
      memcpy(pos + 3 * (size_t)i, p, sizeof(p));
      id[i] = i;
      prop[i] = MDPD_SOLVENT;
      unit[i] = MDPD_UNIT_NONE;
      amp_id[i] = MDPD_UNIT_NONE;
    }
  }
  /* GenVeloc :60-72 */
  const double sd = sqrt(s->temperature);
  for (uint32_t i = 0; i < n; ++i) {
    o_saru r = o_saru3(i, CFG_TAG_VEL, s->seed);
    for (int a = 0; a < 3; ++a) {
      const double u1 = o_cfg_uni(&r, 1);
      const double u2 = o_cfg_uni(&r, 1);
      vel[3 * (size_t)i + a] = 0.0 + sd * sqrt(-2.0 * log(u1)) * cos(2.0 * M_PI * u2);
    }
  }
  oracle_remove_cm_drift(n, vel);
  /* CheckParticleConfigIsValid :420-452 */
  double kin = 0.0;
  for (uint32_t i = 0; i < n; ++i) {
    const double* v = vel + 3 * (size_t)i;
    kin += v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
    for (int a = 0; a < 3; ++a) {
      const double x = pos[3 * (size_t)i + a];
      if (!(x <= L[a] && x >= 0.0)) return MDPD_ERR_CONFIG;
    }
  }
  kin /= 3.0 * n;
  if (kin_temp) *kin_temp = kin;
  if (fabs(kin - s->temperature) >= 1.0e-2) return MDPD_ERR_CONFIG;
  return MDPD_OK;
}
