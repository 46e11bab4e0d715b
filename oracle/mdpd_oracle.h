/*
 * mdpd_oracle.h -- CPU oracle of the MDPD step path.  TEST INFRASTRUCTURE
 * ONLY: used by tests/, __graft_entry__.smoke() and the cpu_baseline leg of
 * bench.py to check the engine against; never linked into or called by the
 * engine.
 *
 * Plain C restatement, FP64, of the reference's algorithms (file:line of
 * kohnakagawa/implicit_dpd cited at each function in mdpd_oracle.c).
 * Per-bead gather loops without Newton's third law, as the reference does;
 * neighbours come from a brute-force minimum-image scan or a cell list, both
 * deciding r^2 < cutoff^2 with the same FP64 expression.
 *
 * Parity pins (which check pins which function):
 *   saru_*          bit-exact vs the reference header's own vectors
 *                   (tests/golden/saru_ref_vectors.txt, built from
 *                   include/saruprng.hpp by oracle/build_ref.sh)
 *   calc_inter_coef debug_dump_rank0.txt fixture of the reference
 *                   (test/mpi_bonded/result_mpi/31x31) + closed forms
 *   pairs           brute force == cell list (test/nearlist/near_list.cpp:73-122)
 *   density/force   closed-form 2- and 3-bead cases, Newton-3 antisymmetry,
 *                   Warren/PAIRWISE reduction identities
 *   bonded          LAMMPS-form closed forms (test/bondbend_test/bondbendtest.cpp:197-302)
 *                   and finite differences of the bond/bend energies
 *   drift/kick      exact algebra of driftkick.hpp:5-54
 *   observe         direct sums
 *   chemistry       constructed cones, brute-force candidate counts and
 *                   binomial statistics (ids, counts, geometry) with the
 *                   keyed Saru draws the engine uses; a reference-MT mode
 *                   consumes one MT19937 stream in the reference's order
 *                   (the generator pinned by its published 10000th output
 *                   and numpy's legacy MT19937; the mode by its draw count
 *                   and by matching the keyed mode wherever the draws do
 *                   not decide); on one process the reference seeds the
 *                   stream with time(nullptr) * rank = 0 (main.cpp:150),
 *                   which seed 0 replays; the engine's reference-MT mode
 *                   is checked against it.  The keyed draws (the engine's
 *                   default) are another stream by design, not a
 *                   restatement of the reference's MT sequence.
 *   make_config     closed-form geometry of each mode (bond vectors, leaflet
 *                   heights and radii, ring counts floor(2 pi r sin / lip),
 *                   per-100 leaflet split), Box-Muller over the pinned Saru
 *                   uniforms, zero drift, and the reference's error cases;
 *                   keyed draws in place of the wall-clock-seeded MT stream
 *                   ("parity unpinned" against that stream).
 */
#ifndef MDPD_ORACLE_H
#define MDPD_ORACLE_H

#include <stdint.h>
#include "../include/mdpd_capi.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---- Saru (saruprng.hpp) ---- */
void oracle_saru_batch(uint32_t n, const uint32_t* s1, const uint32_t* s2,
                       const uint32_t* s3, uint32_t* out_u32 /*[8n]*/,
                       double* out_uni /*[2n]*/, double* out_nrml /*[n]*/);
double oracle_pair_nrml(uint32_t id_i, uint32_t id_j, uint32_t seed3);

/* ---- coefficients (parameter.hpp:131-186) ---- */
void oracle_calc_inter_coef(double chi, double kappa, double rho_co,
                            double vbb, const double cf_r_upper[6],
                            double dt, double reo, uint32_t all_unit,
                            double arc, mdpd_params* out);

/* ---- neighbours ---- */
/* ordered pairs i != j with r^2 < cutoff^2, sorted by (i, j) */
uint64_t oracle_pairs(const mdpd_params* p, uint32_t n, const double* pos,
                      double cutoff, int32_t brute, uint32_t* out_i,
                      uint32_t* out_j, uint64_t cap);

/* ---- non-bonded passes ---- */
void oracle_density(const mdpd_params* p, uint32_t n, const double* pos,
                    const uint32_t* prop, double* dens /*[3n]*/,
                    double* cm /*[6n] or NULL*/, uint32_t* cm_cnt /*[2n]*/);
void oracle_force(const mdpd_params* p, uint32_t n, const double* pos,
                  const double* vel, const uint32_t* id, const uint32_t* prop,
                  const double* dens, uint32_t seed3, double* acc,
                  double* press);
/* sum_j |F_ij| per bead (normaliser of the force tolerance) */
void oracle_force_abs(const mdpd_params* p, uint32_t n, const double* pos,
                      const double* vel, const uint32_t* id,
                      const uint32_t* prop, const double* dens, uint32_t seed3,
                      double* fabs_sum);

/* ---- bonded, integrator, observers ---- */
void oracle_bonded(const mdpd_params* p, uint32_t n_amp, const uint32_t* topo,
                   const double* cf_bend, const double* pos, double* acc,
                   double vir[3]);
void oracle_drift(const mdpd_params* p, uint32_t n, double* pos, double* vel,
                  double* vel_buf, const double* acc, double* delta);
void oracle_kick(const mdpd_params* p, uint32_t n, double* vel,
                 const double* vel_buf, const double* acc);
void oracle_remove_cm_drift(uint32_t n, double* vel);
void oracle_observe(const mdpd_params* p, uint32_t n, uint32_t amp_num,
                    uint32_t sol_num, const double* vel, const double* press,
                    const double* delta, const double bonded_vir[3],
                    mdpd_observables* out);

/* ---- reactions ---- */
uint32_t oracle_head_conversion(const mdpd_params* p, uint32_t n,
                                const uint32_t* amp_id, const uint32_t* unit,
                                uint32_t* prop, uint32_t time);

/* ---- whole system, one step = main.cpp:167-196 ---- */
typedef struct oracle_sys oracle_sys;
oracle_sys* oracle_sys_new(const mdpd_params* p, uint32_t cap,
                           uint32_t amp_cap);
void oracle_sys_free(oracle_sys* s);
int oracle_sys_set(oracle_sys* s, uint32_t n, const double* pos,
                   const double* vel, const double* vel_buf,
                   const double* acc, const uint32_t* id,
                   const uint32_t* prop, const uint32_t* amp_id,
                   const uint32_t* unit);
int oracle_sys_set_topology(oracle_sys* s, uint32_t n_amp,
                            const uint32_t* topo, const uint32_t* core_amp_id,
                            uint32_t n_core);
void oracle_sys_set_clock(oracle_sys* s, uint32_t time);
uint32_t oracle_sys_get(const oracle_sys* s, double* pos, double* vel,
                        double* vel_buf, double* acc, double* press,
                        double* dens, double* delta, uint32_t* id,
                        uint32_t* prop, uint32_t* amp_id, uint32_t* unit);
uint32_t oracle_sys_num_amp(const oracle_sys* s);
void oracle_sys_get_topology(const oracle_sys* s, uint32_t* topo);
/* phases, for explicit sequences */
void oracle_sys_drift(oracle_sys* s);
void oracle_sys_density(oracle_sys* s);
void oracle_sys_force(oracle_sys* s);
void oracle_sys_bonded(oracle_sys* s, double vir[3]);
void oracle_sys_kick(oracle_sys* s);
void oracle_sys_remove_cm_drift(oracle_sys* s);
uint32_t oracle_sys_head_conversion(oracle_sys* s);
int oracle_sys_chem_event(oracle_sys* s, const mdpd_chem_params* cp,
                          int32_t* stop);
/* reference-MT mode of oracle_sys_chem_event (enable != 0): the draws come
   from one MT19937 stream seeded like ChemManager(seed) and consumed in the
   reference's order (cone test per core and amphiphile, then per target the
   tangent angle and the new beads' velocities); enable = 0 restores the
   keyed draws.  oracle_sys_mt_draws counts the stream's outputs. */
void oracle_sys_set_mt(oracle_sys* s, uint32_t seed, int32_t enable);
uint64_t oracle_sys_mt_draws(const oracle_sys* s);
/* the first n genrand_int32 outputs of MT19937 after init_genrand(seed) */
void oracle_mt_u32(uint32_t seed, uint32_t n, uint32_t* out);
void oracle_sys_observe(oracle_sys* s, uint32_t sol_num,
                        mdpd_observables* out);
/* nsteps of the main loop body; returns 0 or 1 if chemistry asked to stop */
int oracle_sys_run(oracle_sys* s, uint32_t nsteps);
int oracle_num_threads(void);
/* OpenMP threads of the following calls (the CPU baseline's 1-thread figure) */
void oracle_set_num_threads(int n);

/* ---- initial configurations (src/config_maker.cpp) ---- */
int oracle_config_count(const mdpd_config_spec* s, uint32_t* n);
/* pos/vel [n][3]; returns MDPD_OK or the status the engine gives */
int oracle_make_config(const mdpd_config_spec* s, uint32_t n, double* pos,
                       double* vel, uint32_t* id, uint32_t* prop,
                       uint32_t* amp_id, uint32_t* unit, double* kin_temp);

#ifdef __cplusplus
}
#endif
#endif
