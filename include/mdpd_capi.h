/*
 * mdpd_capi.h -- C ABI of the MI355X MDPD engine (implicit_dpd_amd).
 *
 * The engine replaces the step path of kohnakagawa/implicit_dpd: the FDPS
 * short-range functors CalcDensity / CalcForceEpEpDPD
 * (include/nbdf_calculator.hpp:11-145), the serial bonded force
 * ForceBonded::CalcListedForce (include/bdf_calculator.hpp:191-220), the
 * integrator drift_and_predict / kick / remove_cmdrift_global
 * (include/driftkick.hpp:5-54), the serial chemistry
 * ChemManager::RandomChemEvent (include/chemmanager.hpp:244-281) and the
 * global reductions of Observer (include/observer.hpp:82-178).  The FDPS
 * neighbour machinery they sit on (DomainInfo, exchangeParticle,
 * TreeForForceShort::Gather, called at src/main.cpp:111-132,169-174) becomes
 * a single-device cell list inside the engine.
 *
 * Conventions: plain pointers and sizes only; every call returns an MDPD_*
 * status (no exceptions cross the ABI); the context owns all device memory;
 * host buffers are borrowed for the duration of a call; one host thread per
 * context.  Calls enqueue on the context's own HIP stream; the calls that
 * hand data to the host (download, observe, bonded virial, chemistry) wait
 * for it.  mdpd_last_error() gives the message of the last failing call.
 *
 * Force law (per ordered pair i != j with r = |r_i - r_j| < 1, r^ = r_ij/r),
 * the generalisation of nbdf_calculator.hpp:108-132:
 *   F_ij = [ (a0 + c1*max(r-arc,0)) * (1-r)
 *          + sum_k mb[k] (rho_i^k + rho_j^k) * max(1 - r/rdm, 0)
 *          + sigma * wR(r) * xi_ij
 *          - gamma * wD(r) * (r^ . v_ij) ] r^
 *   rho_i^k = sum_{j != i, r_ij < rd, r_ij > 0} crho (1 - r_ij/rd)^2,
 *             k = type of j (n_dens == 3) or k = 0 (n_dens == 1)
 *   wR = sqrt(1-r), wD = 1-r  (envelope 0: nbdf_calculator.hpp:125-127)
 *   wR = 1-r,       wD = (1-r)^2 (envelope 1: Groot-Warren)
 *   xi_ij = Saru(min(id), max(id), pair_seed + time).nrml()
 *           (nbdf_calculator.hpp:99-105, saruprng.hpp:239-255,427-445)
 * The reference model is a0 = 0, c1 = cf_c, arc = 0.9, mb = cf_m, rd = rdm
 * = crho = 1, n_dens = 3 (parameter.hpp:138-186); Warren MDPD is a0 = A,
 * c1 = 0, mb = B, rdm = rd, crho = 15/(2 pi rd^3), n_dens = 1; the
 * PAIRWISE_DPD build is a0 = 25, c1 = mb = 0 (nbdf_calculator.hpp:114-117).
 */
#ifndef MDPD_CAPI_H
#define MDPD_CAPI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MDPD_MAX_SPECIES 3
/* sizeof(FPDPD) in include/ptcl_class.hpp:29-36 (u32 x4, F64vec x8, F64 x3) */
#define MDPD_FPDPD_STRIDE 232
#define MDPD_UNIT_NONE 0xffffffffu

/* status codes */
enum {
  MDPD_OK = 0,
  MDPD_ERR_ARG = 1,           /* bad argument (null, size, range) */
  MDPD_ERR_HIP = 2,           /* a HIP runtime call failed */
  MDPD_ERR_CAPACITY = 3,      /* more beads than the context holds */
  MDPD_ERR_NONFINITE = 4,     /* a non-finite position/velocity/force */
  MDPD_ERR_OUT_OF_BOX = 5,    /* a bead outside [0, L] (parameter.hpp:454-460) */
  MDPD_ERR_NLIST_OVERFLOW = 6,/* more neighbours than the pair-list width */
  MDPD_ERR_PARSE = 7,         /* run_param.txt / xyz parse error */
  MDPD_ERR_STATE = 8,         /* call out of order (e.g. force before density) */
  MDPD_ERR_NO_DEVICE = 9,     /* no HIP device */
  MDPD_ERR_UNSUPPORTED = 10,  /* option not available for this context */
  MDPD_ERR_CONFIG = 11        /* a generated configuration fails config_maker's checks */
};

/* bead types, parameter.hpp:211-217 */
enum { MDPD_HYPHIL = 0, MDPD_HYPHOB = 1, MDPD_SOLVENT = 2 };

/* Model and run constants of one system (one replica). */
typedef struct mdpd_params {
  double box[3];            /* box_leng (parameter.hpp:75) */
  double dt;                /* dt */
  int32_t n_species;        /* bead types present, 1..3 */
  int32_t n_dens;           /* density slots: 3 (per neighbour type) or 1 */
  int32_t envelope;         /* 0: reference sqrt envelope, 1: Groot-Warren */
  int32_t pad0;
  double rd;                /* density cutoff (1 in the reference) */
  double crho;              /* density weight normalisation */
  double rdm;               /* many-body envelope cutoff */
  double arc;               /* onset of the c1 term (Parameter::arc = 0.9) */
  double a0[3][3];          /* constant conservative amplitude */
  double c1[3][3];          /* cf_c */
  double mb[3][3][3];       /* cf_m */
  double sigma[3][3];       /* cf_r, already divided by sqrt(dt) */
  double gamma[3][3];       /* cf_g */
  /* bonded chains, bdf_calculator.hpp:95-150 */
  double cf_s;              /* spring: F = cf_s (1/r - 1/bond_leng) dr */
  double cf_b;              /* bending U = cf_b (1 - cos) */
  double cf_b_rigid;        /* bending of core amphiphiles */
  double bond_leng;         /* Parameter::bond_leng = 0.5 */
  /* random streams */
  uint32_t pair_seed;       /* CalcForceEpEpDPD::m_seed */
  uint32_t chem_seed;       /* keyed draws of the chemistry */
  /* head-conversion reaction: a unit-0 bead of type conv_from becomes
     conv_to with probability conv_p per step (no reference counterpart) */
  double conv_p;
  int32_t conv_from;
  int32_t conv_to;
  uint32_t cm_corr;         /* steps between CM-drift removals (100) */
  uint32_t pad1;
} mdpd_params;

/* Chemistry of the serial reference (chemmanager.hpp:116-173,244-281). */
typedef struct mdpd_chem_params {
  double p_thresld;         /* insertion probability per candidate */
  double eps;               /* tangential offset of the copy */
  double influ_rad, influ_hei, influ_dep, influ_grd; /* truncated cone */
  uint32_t max_amp_num;     /* stop once amp_num >= max_amp_num */
  uint32_t n_core;          /* number of core amphiphiles */
  const uint32_t* core_ptcl_id; /* bead ids of the core heads [n_core] */
} mdpd_chem_params;

/* Global reductions of Observer (observer.hpp:82-178), FP64. */
typedef struct mdpd_observables {
  double kin_temp[4];       /* <vx^2>, <vy^2>, <vz^2>, mean  (KineticTempera) */
  double pressure[3];       /* P_xx, P_yy, P_zz                (Pressure) */
  double diffusion[3];      /* MSD per bead, per amphiphile CM, per solvent */
  double bonded_vir[3];     /* twice the bonded virial (bdf_calculator.hpp:105-142) */
  double pair_vir[3];       /* sum_i press_i (twice the pair virial) */
  double n_pairs;           /* ordered pairs inside the cutoff, last pair list */
  uint32_t n_beads;
  uint32_t n_amp;
  uint32_t time;
  uint32_t pad;
} mdpd_observables;

/* Launch choices the autotuner picks.  Block sizes, nlist_width and
   use_graph do not change results; verlet_skin changes only the FP32
   rounding of pair terms (the pair sets are the same). */
typedef struct mdpd_tuning {
  int32_t block_scan;       /* threads per block, density/pair-list pass */
  int32_t block_force;      /* threads per block, force pass */
  int32_t cell_div;         /* cells per unit cutoff along each axis; must be 1 */
  int32_t nlist_width;      /* pair-list slots per bead, a multiple of 4 */
  int32_t reorder_every;    /* ignored: bead storage is put in cell order at
                               every cell-list rebuild */
  int32_t use_graph;        /* capture each step in a hipGraph */
  float verlet_skin;        /* 0: mdpd_run rebuilds the cell list every step.
                               In 0.05..1: cells are at least 1 + skin wide,
                               each build keeps every pair closer than
                               1 + skin, and mdpd_run rebuilds cells and list
                               only in steps where some bead has moved more
                               than skin / 2 since the last build (decided on
                               the device, inside the captured step).  Every
                               box edge must hold 3 such cells. */
  int32_t verlet_width;     /* skin-list slots per bead, multiple of 4 (8..2048) */
} mdpd_tuning;

typedef struct mdpd_ctx mdpd_ctx;

/* ---- lifecycle ---------------------------------------------------------- */
/* One system (FDPS ParticleSystem + DomainInfo + two trees, main.cpp:93-127). */
int mdpd_create(const mdpd_params* p, uint32_t capacity, mdpd_ctx** out);
/* n_rep independent replicas in one context; replica r holds n_beads[r]
   beads, uploaded back to back.  No reference counterpart: the reference
   runs one system per process (test/simple_dpdfluid/run_all.sh:5-13). */
int mdpd_create_ensemble(const mdpd_params* p, const uint32_t* n_beads,
                         int32_t n_rep, mdpd_ctx** out);
int mdpd_destroy(mdpd_ctx* ctx);
const char* mdpd_last_error(const mdpd_ctx* ctx);
/* message of the last failure that had no context (create) */
const char* mdpd_global_error(void);
int mdpd_device_name(char* buf, size_t len);

/* ---- state -------------------------------------------------------------- */
/* FPDPD records (ptcl_class.hpp:29-36), `stride` bytes apart (>= 232). */
int mdpd_upload_fpdpd(mdpd_ctx* ctx, const void* aos, uint32_t n, size_t stride);
int mdpd_download_fpdpd(mdpd_ctx* ctx, void* aos, uint32_t n, size_t stride);
/* Structure-of-arrays transfer; any pointer may be NULL on download.  vec3
   arrays are [n][3] doubles; dens is [n][3]. */
int mdpd_upload_soa(mdpd_ctx* ctx, uint32_t n, const double* pos,
                    const double* vel, const double* acc, const uint32_t* id,
                    const uint32_t* prop, const uint32_t* amp_id,
                    const uint32_t* unit);
int mdpd_download_soa(mdpd_ctx* ctx, uint32_t n, double* pos, double* vel,
                      double* vel_buf, double* acc, double* press,
                      double* dens, double* delta_sumr, uint32_t* id,
                      uint32_t* prop, uint32_t* amp_id, uint32_t* unit);
int mdpd_num_beads(const mdpd_ctx* ctx, uint32_t* n);
/* glob_topol_[amp*4+unit] = bead index (bdf_calculator.hpp:61-88); the
   bending constant of amphiphile a is cf_b_rigid if a is listed in
   core_amp_id, else cf_b.  n_amp may be 0 (no chains). */
int mdpd_set_topology(mdpd_ctx* ctx, const uint32_t* glob_topol,
                      uint32_t n_amp, const uint32_t* core_amp_id,
                      uint32_t n_core);
/* Parameter::time and CalcForceEpEpDPD::m_seed */
int mdpd_set_clock(mdpd_ctx* ctx, uint32_t time, uint32_t pair_seed);
int mdpd_get_clock(const mdpd_ctx* ctx, uint32_t* time);
int mdpd_set_tuning(mdpd_ctx* ctx, const mdpd_tuning* t);
int mdpd_get_tuning(const mdpd_ctx* ctx, mdpd_tuning* t);

/* ---- one step, in the order of src/main.cpp:167-196 --------------------- */
int mdpd_drift_predict(mdpd_ctx* ctx);          /* driftkick.hpp:5-30 */
int mdpd_rebuild_cells(mdpd_ctx* ctx);          /* exchangeParticle + tree */
int mdpd_density(mdpd_ctx* ctx);                /* CalcDensity */
int mdpd_force(mdpd_ctx* ctx);                  /* CalcForceEpEpDPD */
int mdpd_bonded(mdpd_ctx* ctx, double vir[3]);  /* ForceBonded::CalcListedForce */
int mdpd_kick(mdpd_ctx* ctx);                   /* driftkick.hpp:32-40 */
int mdpd_remove_cm_drift(mdpd_ctx* ctx);        /* driftkick.hpp:42-54 */
int mdpd_head_conversion(mdpd_ctx* ctx);        /* keyed type flip, no counterpart */
/* RandomChemEvent (chemmanager.hpp:244-281) with keyed draws (or the
   reference's MT19937 stream, mdpd_set_chem_mt) and the new
   amphiphiles appended to the bonded topology; *stop = 1 once
   amp_num >= max_amp_num.  The draws are keyed on the context clock.  The
   event runs on the device with the context's own scratch (no allocation,
   no copy of the topology); inserted beads do not invalidate the captured
   steps of mdpd_run.  More amphiphiles than the capacity holds: nothing is
   inserted and MDPD_ERR_CAPACITY. */
int mdpd_chem_event(mdpd_ctx* ctx, const mdpd_chem_params* cp, int* stop);
/* The chemistry branch of the main loop (main.cpp:186-196) inside
   mdpd_run: with enable != 0 every step whose time >= beg_chem ends with
   RandomChemEvent after the CM-drift removal, at that step's time, inside
   the captured step (the reference's beg_chem is a compile-time constant,
   parameter.hpp:254; here a runtime field).  When an event brings amp_num to
   max_amp_num the run ends after that step, as the reference's loop breaks
   (main.cpp:188): the clock stays at that step's time, mdpd_chem_status
   reports stopped = 1 and mdpd_run refuses to continue until
   mdpd_set_chemistry is called again.  mdpd_run then synchronises at its end
   (it reads back the bead count).  Single-system contexts with
   verlet_skin = 0; at most 256 core amphiphiles.  enable = 0 switches it off. */
int mdpd_set_chemistry(mdpd_ctx* ctx, const mdpd_chem_params* cp, uint32_t beg_chem,
                       int32_t enable);
/* Draws of the chemistry (mdpd_chem_event and the events inside mdpd_run):
   enable != 0 takes them from one MT19937 stream started by
   init_genrand(seed), consumed in the reference's order, as ChemManager's
   PS::MT is (chemmanager.hpp:176-179, seeded at main.cpp:151 with
   time(nullptr) * rank, 0 on one process): one genrand_real1 per chain head
   inside a core's cone (cores in order, chains in order; :140-161), then per
   target one genrand_real1 for the tangent angle (:167) and two
   genrand_real3 per velocity component of each of its four beads (:18-20,
   the logarithm's first).  Calling it again restarts the stream from seed.
   enable = 0 returns to the keyed draws (the default).  A checkpoint saved
   with the mode on holds the stream, and loading it switches the mode on at
   that position.  Single-system contexts. */
int mdpd_set_chem_mt(mdpd_ctx* ctx, uint32_t seed, int32_t enable);
/* outputs of the MT19937 stream consumed since mdpd_set_chem_mt (0 when
   the keyed draws are on).  No reference counterpart. */
int mdpd_chem_mt_draws(mdpd_ctx* ctx, uint64_t* draws);
/* stopped: an event inside the last mdpd_run reached max_amp_num; events and
   inserted (may be NULL): in-run events and amphiphiles added since
   mdpd_set_chemistry.  No reference counterpart. */
int mdpd_chem_status(mdpd_ctx* ctx, int32_t* stopped, uint32_t* events, uint32_t* inserted);
int mdpd_observe(mdpd_ctx* ctx, mdpd_observables* out);
/* observables of every replica of an ensemble context, out[n_rep] */
int mdpd_observe_replicas(mdpd_ctx* ctx, mdpd_observables* out);
/* Trajectory / configuration output in the reference's xyz text format
   (io_util::WriteXYZForm + FPDPD::writeAscii, io_util.hpp:25-31,
   ptcl_class.hpp:67-73; the writer of traject.xyz every step_mic and of
   fin_config.xyz, main.cpp:199-208): count line, "time T", one 17-column
   line per bead in upload order, reals as %.15g.  The state is copied into
   pinned host memory on the context's stream and the call returns; a host
   thread formats (in parallel) and writes the file while the caller goes on
   (append != 0: add a frame).  The next write, mdpd_io_wait and
   mdpd_destroy wait for it; mdpd_io_wait returns the status of the write
   (file errors as MDPD_ERR_ARG).  No reference counterpart for the
   asynchrony: the reference formats on the calling thread. */
int mdpd_write_xyz_async(mdpd_ctx* ctx, const char* path, uint32_t time, int32_t append);
int mdpd_io_wait(mdpd_ctx* ctx);
/* Reads a configuration in the reference's xyz text format into FPDPD
   records `stride` bytes apart (io_util::ReadXYZForm + FPDPD::readAscii,
   io_util.hpp:96-129, ptcl_class.hpp:60-65; the loader of init_config.xyz,
   Parameter::LoadParticleConfig, parameter.hpp:420-428).  Host only, no
   context: the file is read once and its lines parsed in parallel.  The
   header must be "<n>" then "time <t>"; more than `cap` records gives
   MDPD_ERR_CAPACITY, a line count other than <n> or a line without its 17
   columns MDPD_ERR_PARSE (the reference aborts in these cases); the message
   is in mdpd_global_error.  delta_sumr, press, nei_cm_pos and density are
   zero.  mdpd_read_xyz_header reads only <n> and <t>, to size the records. */
int mdpd_read_xyz_header(const char* path, uint32_t* n, uint32_t* time);
int mdpd_read_xyz(const char* path, void* aos, uint32_t cap, size_t stride, uint32_t* n,
                  uint32_t* time);

/* Binary checkpoint of the whole evolving state in the device's storage
   order (positions, velocities, vel_buf, delta_sumr, accelerations, virials,
   ids and types, chain map and topology, clock and pair seeds).  Loading it
   into a context created with the same parameters and capacity continues
   the run bit for bit (the every-step cell list; in skin mode the first
   step after a load rebuilds, so FP32 rounding may differ).  The
   reference restarts from ASCII (restart.py, fin_config.xyz); this is the
   exact-restart sidecar. */
int mdpd_checkpoint_save(mdpd_ctx* ctx, const char* path);
int mdpd_checkpoint_load(mdpd_ctx* ctx, const char* path);
/* Pair-distance histogram (radial distribution function; no reference
   counterpart): ordered pairs i != j of the current positions with r in
   [k dr, (k+1) dr), dr = rmax / nbins, decided on r^2 against the FP64
   squared edges (k dr)^2; out[n_rep][nbins].  0 < rmax <= 1, nbins <= 1024;
   needs the cell list of the current positions (after mdpd_run or
   mdpd_rebuild_cells). */
int mdpd_rdf(mdpd_ctx* ctx, double rmax, int32_t nbins, double* out);
/* mdpd_rdf without any device allocation: no counterpart in the reference.
   For callers that cannot allocate device memory (code under graph capture,
   tight serving loops).  Its scratch (bin edges, counters) is one fixed
   device array of 1,057,024 bytes per GPU, held until the process ends, so
   n_rep * nbins may be at most 65536 (more gives MDPD_ERR_CAPACITY, before
   any work).  Same result as mdpd_rdf, bit for bit.  All calls on one GPU
   share that array: calls that overlap on a GPU race, and the caller must
   serialise them. */
int mdpd_rdf_noalloc(mdpd_ctx* ctx, double rmax, int32_t nbins, double* out);
/* Chain geometry histograms, the quantities of
   analysis/bond_angle/h2e.cpp:64-104, out[n_rep][4][nbins]: row 0 the first
   bond |r1 - r0|, row 1 the other bonds (bins of bond_max / nbins, on the
   squared length), row 2 the cosine of the angle between consecutive bonds
   (bins of 2 / nbins over [-1, 1], decided without a square root), row 3
   the end-to-end |r3 - r0| (bins of 3 bond_max / nbins). */
int mdpd_chain_histograms(mdpd_ctx* ctx, int32_t nbins, double bond_max,
                          double* out);
/* 3-D count map of the beads whose type bit is set in type_mask
   (analysis/trjanalysis/density_map.hpp:13-87, without its rotation into
   the membrane's principal axes): voxels box / g per axis, index
   (int)(x / edge) clamped to g - 1, out[n_rep][gz][gy][gx]. */
int mdpd_count_map(mdpd_ctx* ctx, int32_t gx, int32_t gy, int32_t gz,
                   uint32_t type_mask, double* out);
/* bead-count profile along `axis` in nbins bins for every replica,
   out[n_rep][nbins] (counts per bin, FP64) */
int mdpd_density_profile(mdpd_ctx* ctx, int32_t axis, int32_t nbins,
                         double* out);
/* `nsteps` iterations of main.cpp:166-196 (drift .. time++); the pair
   virial of the last step is kept for mdpd_observe. */
int mdpd_run(mdpd_ctx* ctx, uint32_t nsteps);
/* mdpd_run with HIP events recorded on the context's stream around the
   steps; *gpu_ms is the device time between them.  Used by the benchmark. */
int mdpd_run_timed(mdpd_ctx* ctx, uint32_t nsteps, double* gpu_ms);
/* Records, without running anything, the step graphs that the next
   mdpd_run(ctx, nsteps) launches (one per parity of the master storage and
   kind of step: virial, CM-drift removal, key pass, chemistry).  mdpd_run
   and mdpd_run_timed do this themselves before their first launch (and
   mdpd_run_timed before its first event); a caller that times its own runs
   calls it first.  No reference counterpart (the reference has no graphs). */
int mdpd_prepare_run(mdpd_ctx* ctx, uint32_t nsteps);
/* Skin mode (mdpd_tuning.verlet_skin > 0): the number of steps so far that
   rebuilt cells and skin list, and (entries may be NULL) the skin-list
   entries of the last build summed over the beads.  No reference
   counterpart. */
int mdpd_skin_stats(mdpd_ctx* ctx, uint64_t* builds, uint64_t* entries);
/* Blocks until the context's stream is idle. */
int mdpd_sync(mdpd_ctx* ctx);

/* ---- checks and measurement -------------------------------------------- */
/* Ordered neighbour pairs (i != j, r^2 < cutoff^2, i and j bead indices)
   seen by the last rebuild, sorted by (i, j).  cutoff must be 1 or the
   density cutoff.  Mirrors test/nearlist_gpu/near_list.cu:70-138. */
int mdpd_neighbor_pairs(mdpd_ctx* ctx, double cutoff, uint32_t* out_i,
                        uint32_t* out_j, uint64_t cap, uint64_t* n_out);
/* Saru on the device: out_u32[8*k] = first 8 u32() of Saru(s1,s2,s3)[k],
   out_nrml[k] = FP32 nrml() of a fresh stream (saruprng.hpp:239-255,
   322-334,427-445). */
int mdpd_saru_device(const uint32_t* s1, const uint32_t* s2,
                     const uint32_t* s3, uint32_t n, uint32_t* out_u32,
                     float* out_nrml);
/* Per-phase average kernel time (ms) over the steps run since the last
   call with profiling on; names are "drift", "sort", ..., NUL separated. */
int mdpd_set_profiling(mdpd_ctx* ctx, int32_t on);
int mdpd_phase_times(mdpd_ctx* ctx, double* ms, int32_t cap, int32_t* n,
                     char* names, size_t names_len);
/* Device copy of `bytes` bytes (read + write), `reps` times, with the
   engine's own copy kernel; *gbps = 2 * bytes * reps / time.  Calibrates
   HBM counters and confirms the bandwidth of the device. */
int mdpd_bandwidth_probe(uint64_t bytes, int32_t reps, double* gbps);
/* Times candidate launch choices on the current state (nsteps each) and
   keeps the fastest; the state is restored after every candidate and the
   results are bit-identical across candidates.  out may be NULL.  No
   reference counterpart (project-only autotuner). */
int mdpd_autotune(mdpd_ctx* ctx, uint32_t nsteps, mdpd_tuning* out);
/* mdpd_autotune, and with MDPD_TUNE_SKIN also the neighbour-list bin: the
   cell list rebuilt every step, or skins 0.1, 0.2, 0.3 (cells of 1 + skin,
   where the boxes allow them).  A skin changes FP32 rounding only; the pair
   sets stay exact. */
#define MDPD_TUNE_SKIN 1u
int mdpd_autotune_ex(mdpd_ctx* ctx, uint32_t nsteps, uint32_t flags,
                     mdpd_tuning* out);

/* ---- host-side parameter input (parameter.hpp:74-186,352-387) ---------- */
/* Reads run_param.txt in `dir`; fills p with the reference coefficients of
   CalcInterCoef and cp/run with the run and chemistry constants.  Values
   are parsed with std::stof like the reference (parameter.hpp:332-340). */
typedef struct mdpd_run_input {
  uint32_t init_amp_num, sol_num;
  uint32_t all_time, step_mic, step_mac;
  uint32_t max_amp_num;
  double p_thresld, eps;
  double influ_rad, influ_hei, influ_dep, influ_grd, influ_vol, upside_rad;
  double chi, kappa, rho_co, vbb;
  double cf_r_in[6];
  uint32_t n_core;
  uint32_t core_amp_id[64];
} mdpd_run_input;
int mdpd_read_run_param(const char* dir, mdpd_params* p, mdpd_run_input* in);

/* ---- initial configurations (src/config_maker.cpp) ---------------------
   The input of the reference's ConfigMaker (config_maker.cpp:14-38, tags
   read by MatchingTagValues :386-418).  Modes (GenConfig, :342-384):
     FLAT       bilayer normal to y, leaflets alternating amphiphile by
                amphiphile (MakeFlatSheet :103-138); solvent in the lower
                half of y (MakeCuboidSolvent :282-293)
     ASYM_FLAT  as FLAT, but each block of 100 amphiphiles puts its first
                (int)(100 in_out_rat) in the upper leaflet (:113-137)
     SPHERE     vesicle of rings of spacing lip_len, outer leaflet then
                inner (MakeSphSheet/MakeSphLine :139-213); solvent inside
                (MakeSphericalSolvent :295-321)
     CYLIND     cylinder along z (MakeCylindSheet :215-262); solvent inside
     RANDOM     random head positions, directions in the positive octant
                (MakeRandomConfig :264-280)
   Velocities are N(0, Temperature) per component with the centre-of-mass
   drift removed, vel_buf = vel, acc = 0 (GenVeloc :60-72).
   Random draws: the reference seeds MT19937 with time(NULL) and draws in
   bead order; here every amphiphile, solvent bead and velocity has its own
   keyed Saru stream (bead or amphiphile index, stream tag, seed), so the
   generator runs one thread per bead and a given seed always gives the
   same configuration. */
enum { MDPD_CFG_FLAT = 0, MDPD_CFG_ASYM_FLAT = 1, MDPD_CFG_SPHERE = 2,
       MDPD_CFG_CYLIND = 3, MDPD_CFG_RANDOM = 4 };
typedef struct mdpd_config_spec {
  int32_t mode;             /* MDPD_CFG_* */
  int32_t with_solvent;     /* tag with_solvent present */
  double box[3];            /* box_leng */
  double temperature;       /* Temperature */
  uint32_t amp_num, head_unit, tail_unit, sol_num;
  double bond_leng, lip_len;
  double sph_rad, upper_the;  /* SPHERE; upper_the in degrees */
  double cyl_l, cyl_r;        /* CYLIND */
  double in_out_rat;          /* ASYM_FLAT */
  uint32_t seed;              /* key of every draw */
  uint32_t pad;
} mdpd_config_spec;
/* Reads config_param.txt in `dir` (ConfigMaker::LoadParam, :455-467):
   required tags Temperature, box_leng, amp_num, mode, head_unit, tail_unit,
   bond_leng, lip_len, plus sph_rad/upper_the, cyl_l/cyl_r, in_out_rat,
   sol_num by mode; values through std::stof / std::stoi like the
   reference.  seed is left 0. */
int mdpd_read_config_param(const char* dir, mdpd_config_spec* s);
/* beads the spec makes: amp_num (head_unit + tail_unit) + sol_num if
   with_solvent (GenParticles :469-477) */
int mdpd_config_count(const mdpd_config_spec* s, uint32_t* n_beads);
/* Generates the configuration on the device into host arrays of n =
   mdpd_config_count beads, in the reference's bead order (amphiphile a is
   beads a*all_unit .. a*all_unit + all_unit-1, solvent follows; id = bead
   index).  pos/vel [n][3].  Errors follow config_maker's exits: sphere
   radius too large, too many amphiphiles for the rings, a bead outside
   [0, L], |kinetic temperature - Temperature| >= 1e-2
   (CheckParticleConfigIsValid :420-452), with the message in
   mdpd_global_error(); kin_temp (may be NULL) gets sum v.v / (3 n). */
int mdpd_make_config_host(const mdpd_config_spec* s, uint32_t n, double* pos,
                          double* vel, uint32_t* id, uint32_t* prop,
                          uint32_t* amp_id, uint32_t* unit, double* kin_temp);
/* The same configuration generated straight into a single-system context
   (replaces config_maker + init_config.xyz + mdpd_upload_fpdpd): the beads
   become the context's state and, for 4-bead amphiphiles (head_unit +
   tail_unit == 4, the engine's chain model), the chains its topology with
   no core amphiphiles (mdpd_set_topology may follow).  Other chain lengths
   are refused with MDPD_ERR_UNSUPPORTED.  A configuration that fails the
   checks above leaves the context with no beads. */
int mdpd_make_config(mdpd_ctx* ctx, const mdpd_config_spec* s, double* kin_temp);

/* ---- equation-of-state surrogate (no reference counterpart) ------------
   A 4 -> 128 -> 128 -> 2 perceptron with SiLU hidden layers that maps
   (A, B, rd, kBT) (normalised by the caller) to (rho, P) of a plain run.
   Products on the matrix cores with BF16 operands and FP32 accumulation;
   FP32 master parameters, Adam.  Parameter layout (17,410 floats, weights
   row-major [out][in]): W1[128][4], b1[128], W2[128][128], b2[128],
   W3[2][128], b3[2].  Loss = sum over samples and outputs of
   (y - t)^2 / (2 n). */
typedef struct mdpd_mlp mdpd_mlp;
int mdpd_mlp_num_params(uint32_t* n);
/* max_batch: samples per device batch; parameters drawn from a seeded
   U(-1/sqrt(fan_in), 1/sqrt(fan_in)) */
int mdpd_mlp_create(uint32_t max_batch, uint32_t seed, mdpd_mlp** out);
int mdpd_mlp_destroy(mdpd_mlp* m);
const char* mdpd_mlp_last_error(const mdpd_mlp* m);
int mdpd_mlp_set_params(mdpd_mlp* m, const float* params, uint32_t n);
int mdpd_mlp_get_params(mdpd_mlp* m, float* params, uint32_t n);
/* y[n][2] = f(x[n][4]), host arrays, any n */
int mdpd_mlp_forward(mdpd_mlp* m, const float* x, uint32_t n, float* y);
/* loss and its gradient w.r.t. the parameters over n <= max_batch samples;
   grad may be NULL */
int mdpd_mlp_loss_grad(mdpd_mlp* m, const float* x, const float* t, uint32_t n,
                       double* loss, float* grad);
/* Adam (beta 0.9/0.999, eps 1e-8) over mini-batches of `batch` samples of a
   per-epoch shuffle (seeded); loss_per_epoch[epochs] may be NULL */
int mdpd_mlp_train(mdpd_mlp* m, const float* x, const float* t, uint32_t n,
                   uint32_t epochs, uint32_t batch, float lr, uint32_t seed,
                   float* loss_per_epoch);

#ifdef __cplusplus
}
#endif
#endif /* MDPD_CAPI_H */
