// Device-state view handed (by value) to the launch wrappers.
#pragma once
#include "common.hpp"

namespace mdpd {

// Per-bead master storage, SoA with stride cap.  Bead storage is kept in
// cell order: every cell-list rebuild permutes it into the new sorted order
// (a near-identity permutation, so the copies stream), and `orig` carries
// each bead's upload slot for the host.
struct Master {
  double* pos;      // [3][cap]
  double* vel;      // [3][cap]  velocity after the last kick
  double* vbuf;     // [3][cap]  predicted velocity = vel_buf (driftkick.hpp:25-27)
  double* delta;    // [3][cap]  delta_sumr
  float* acc;       // [3][cap]
  float* press;     // [3][cap]  per-bead pair virial (written on virial steps)
  uint32_t* id;
  uint32_t* prop;
  uint32_t* amp_id;
  uint32_t* unit;
  uint32_t* orig;   // upload slot of the bead in this storage slot
  uint32_t* chain;  // chain index << 2 | position in the chain, or ~0u
};

// The force pass's per-bead record, 32 bytes: a neighbour gather reads one
// 32-byte piece of a cache line instead of pieces of two lines.  (The
// density pass streams the positions alone, 16 bytes apart, so they are
// also kept in their own array.)
struct SBead {
  float4 pos;  // local coordinates in the bead's cell, .w = id << 2 | type
  float4 vel;  // predicted velocity; .w = density when n_dens == 1
};

struct DevState : Master {
  Master alt;       // the other buffer set, target of the next permutation
  uint32_t cap;
  uint32_t n;       // beads in use, as the host last knew them
  uint32_t nl;      // launch bound of per-bead kernels: n, or cap in grow mode
  // grow mode (set once chemistry is used): the live bead and chain counts
  // are rep_off[n_rep] / amp_off[n_rep] on the device, which chemistry
  // advances inside captured steps, and *run_ok = 0 makes every remaining
  // kernel of a run return at once (the chemistry stopped the run); null
  // otherwise
  uint32_t* run_ok;
  // cell list, rebuilt every step
  uint32_t* key;        // [cap] global cell of each bead
  uint32_t* cell_cnt;   // [ncell + 1]
  uint32_t* cell_start; // [ncell + 1]
  uint32_t* cursor;     // [ncell]
  uint32_t* bsum;       // scan block sums
  uint32_t* perm;       // [cap] new slot -> old slot of the last rebuild
  uint32_t* inv_perm;   // [cap] old slot -> new slot
  uint32_t ncell;
  // FP32 copies in cell order (same slots as the master storage)
  float4* spos;         // [cap] the positions alone, for the cell and list walks
  SBead* sb;            // [cap] position + velocity, for the force pass
  uint32_t* scell;      // global cell of each slot at the last rebuild
  float* sdens;         // [cap] density (one species), else float4 [cap] (rho_0, rho_1, rho_2, 0)
  uint32_t* nlist;      // wave-tiled pair list [cap/64][W/4][64][4]
  uint32_t* ncount;     // [cap]
  int32_t W;
  // skin mode: neighbour list of radius 1 + skin, rebuilt on demand
  uint32_t* vlist;      // [cap/64][Wv/4][64][4], slot | stencil code << 26
  uint32_t* vcount;     // [cap]
  double* dref;         // [3][cap] delta_sumr at the last build
  uint32_t* rebuild;    // [0]: 1 = rebuild cells and list in this step; [1]: builds so far
  int32_t Wv;
  float vcut2;          // squared list radius (1 + skin + margin)^2
  double disp2;         // (skin / 2)^2: a larger displacement forces a rebuild
  double cell_slack;    // how far a bead may sit outside its cell (0 or skin)
  // replicas
  Replica* reps;
  uint32_t* rep_off;    // [n_rep + 1] first bead of each replica
  uint32_t* rep_cell_off; // [n_rep + 1]
  int32_t n_rep;
  int32_t n_dens;
  int32_t rd_le_1;      // every replica's rd <= 1 (the density pass's cutoff test)
  // bonded
  uint32_t* topo;       // [4][amp_cap] chain bead slots, amp-major
  double4* chainpos;    // [4 amp_cap] FP64 positions of chain beads, chain-major
  double* cf_bend;      // [amp_cap]
  uint32_t* amp_rep;    // [amp_cap] replica of each chain
  uint32_t* amp_off;    // [n_rep + 1] first chain of each replica
  uint32_t n_amp;
  // scalars and scratch
  uint32_t* d_time;
  uint32_t* flags;
  double* red;          // reduction partials
  double* red_out;      // reduction results
};

// live bead count: the device's in grow mode, the by-value host count otherwise
__device__ __forceinline__ uint32_t dev_n(const DevState& d) {
  return d.run_ok ? d.rep_off[d.n_rep] : d.n;
}
__device__ __forceinline__ uint32_t dev_namp(const DevState& d) {
  return d.run_ok ? d.amp_off[d.n_rep] : d.n_amp;
}
// tiles of bs beads that hold live beads: the grid's size, or in grow mode
// (grid over the capacity) those of the device's count, so that the XCD
// mapping spreads the live tiles over all XCDs
__device__ __forceinline__ uint32_t live_tiles(const DevState& d, uint32_t bs) {
  return d.run_ok ? (dev_n(d) + bs - 1u) / bs : gridDim.x;
}
// a chemistry stop ended the run: the step's kernels do nothing
__device__ __forceinline__ bool halted(const DevState& d) {
  return d.run_ok && *d.run_ok == 0u;
}

// RandomChemEvent parameters on the device (chemmanager.hpp:244-281 and the
// Parameter fields it reads); slot 0 serves the chemistry inside mdpd_run,
// slot 1 the explicit mdpd_chem_event.
constexpr uint32_t kMaxCore = 256;
struct ChemDev {
  double p_thresld, eps, influ_rad, influ_hei, influ_dep, influ_grd, bond_leng, cf_b;
  uint32_t max_amp_num, n_core, in_run, pad;
  uint32_t core_id[kMaxCore];  // bead ids of the core heads (Parameter::core_ptcl_id)
};
// per-event state and results of a chemistry slot
struct ChemState {
  uint32_t nt;       // amphiphiles inserted by the last event
  uint32_t n0, namp0;  // counts before it
  uint32_t stop;     // n_amp >= max_amp_num after it
  uint32_t refused;  // it would have exceeded the capacity (nothing inserted)
  uint32_t events;   // events so far
  uint32_t total;    // amphiphiles inserted so far
  uint32_t n_in;     // reference-MT mode: heads inside a cone at the last event
  uint32_t core_slot[kMaxCore];   // storage slot of each core head (~0: absent)
  double core_geo[kMaxCore][6];   // corrected head position, membrane normal h2t
};

// The reference-MT mode of the chemistry: the draws come from one MT19937
// stream (the reference's PS::MT, seeded by ChemManager(seed),
// chemmanager.hpp:176-179, main.cpp:151) consumed in the reference's order: one
// genrand_real1 per head inside a core's cone, cores in order and chains in
// order within a core (:140-161), then per target one genrand_real1 for the
// tangent angle (:167) and two genrand_real3 per velocity component of its
// four beads (:18-20, the log's draw first).
constexpr uint32_t kMtN = 624;
constexpr uint32_t kMtDrawsPerTarget = 25;
struct MtState {
  uint32_t mt[kMtN];
  uint32_t mti;                // index of the next output in mt (kMtN: twist first)
  uint32_t pad;
  unsigned long long draws;    // outputs consumed so far
};
struct ChemMt {
  MtState* g;            // null: keyed draws
  uint32_t* cone_offs;   // [kMaxCore * chem_count_blocks] in-cone heads, then their offsets
  uint32_t* cone_bits;   // [cone_words] bit r: in-cone draw r is <= p_thresld
  uint32_t* tdraws;      // [kMtDrawsPerTarget * amp_cap] the targets' raw outputs
  uint32_t cone_words;   // kMaxCore * amp_cap / 32 + 2 (at most one draw per core and chain)
};

// launch wrappers (defined in the .hip files)
void launch_fill_u32(uint32_t* p, uint32_t n, uint32_t v, hipStream_t st);
// tick: advance the device clock (the step's first kernel)
void launch_drift(const DevState& d, bool keys, bool predict_only, hipStream_t st,
                  bool tick = false);
void launch_keys(const DevState& d, hipStream_t st);
// cell key + histogram of each bead's drifted position (the state itself
// is drifted by launch_step_permute)
void launch_step_keys(const DevState& d, hipStream_t st);
// the step's permutation: clock, drift + move into the other set in cell order
void launch_step_permute(const DevState& d, hipStream_t st);
void launch_scan_cells(const DevState& d, hipStream_t st);
void launch_scatter(const DevState& d, hipStream_t st);
void launch_cell_sort(const DevState& d, hipStream_t st);
// permutes the master storage into d.alt in cell order and writes the FP32
// copies; `full` also carries vel, acc and press (needed when they are live)
void launch_permute(const DevState& d, bool predicted, bool full, hipStream_t st);
void launch_topo_remap(const DevState& d, hipStream_t st);
// topo rebuilt from chain[] (the step keeps only chain[] current)
void launch_topo_from_chain(const DevState& d, hipStream_t st);
void launch_density(const DevState& d, int ns, int nd, int block, hipStream_t st);
// fuse: the force pass ends with the bead's bonded terms and the kick
// (virial steps: per-bead chain virials for launch_bead_vir_reduce)
// conv (with fuse): the head conversion of the bead ends the epilogue
// keys (with fuse): the epilogue also writes the cell key and histogram of
// the bead's position after the next step's drift (launch_step_keys' work)
void launch_force(const DevState& d, int ns, int nd, int env, bool virial, int block,
                  hipStream_t st, bool fuse = false, bool conv = false, bool keys = false);
double* bead_vir_buffer(const DevState& d);
void launch_bead_vir_reduce(const DevState& d, double* vir_out, hipStream_t st);
void launch_kick(const DevState& d, hipStream_t st);
// bonded forces of every bead's own chain added to its acceleration, then
// the kick, in one pass over the beads (chain virial on virial steps)
void launch_bonded_kick(const DevState& d, bool virial, bool kick, double* vir_out,
                        hipStream_t st);
// chainpos from the master positions (the permutation keeps it current)
void launch_chainpos_fill(const DevState& d, hipStream_t st);
void launch_cm_drift(const DevState& d, hipStream_t st);
void launch_head_conversion(const DevState& d, hipStream_t st);
// skin mode (launches of the rebuild kernels return at once unless
// *d.rebuild is set)
void launch_drift_skin(const DevState& d, hipStream_t st);
void launch_keys_gated(const DevState& d, hipStream_t st);
void launch_scan_cells_gated(const DevState& d, hipStream_t st);
void launch_scatter_gated(const DevState& d, hipStream_t st);
void launch_cell_sort_gated(const DevState& d, hipStream_t st);
void launch_permute_gated(const DevState& d, hipStream_t st);
void launch_copy_back_gated(const DevState& d, hipStream_t st);
void launch_skin_build(const DevState& d, hipStream_t st);
// also counts the build of this step (rebuild[1]) and clears rebuild[0]
void launch_density_skin(const DevState& d, int nd, int block, hipStream_t st);
// observables: out[n_rep][13] FP64; dl_load scratch FP64 [3][cap]
void launch_observe(const DevState& d, const uint32_t* amp_num /*[n_rep]*/, double* dl_load,
                    double* out, hipStream_t st);
void launch_density_profile(const DevState& d, int axis, int nbins, uint32_t* hist,
                            hipStream_t st);
// observables of k_obs.hip: e2 / eb2 / ee2 are squared bin edges [nb + 1]
void launch_rdf(const DevState& d, const double* e2, int nb, double rmax,
                unsigned long long* hist, hipStream_t st);
void launch_chain_hist(const DevState& d, const double* eb2, const double* ee2, int nb,
                       unsigned long long* hist, hipStream_t st);
void launch_count_map(const DevState& d, int g0, int g1, int g2, uint32_t type_mask,
                      unsigned int* map, hipStream_t st);
void launch_neighbor_cm(const DevState& d, const uint32_t* beads, uint32_t nb, double* cm_out,
                        hipStream_t st);
void launch_pairs_list(const DevState& d, double cutoff, uint32_t* out_i, uint32_t* out_j,
                       unsigned long long* counter, uint64_t cap, hipStream_t st);
void launch_saru_test(const uint32_t* s1, const uint32_t* s2, const uint32_t* s3, uint32_t n,
                      uint32_t* u32out, float* nrml_out, hipStream_t st);
void launch_copy_d(double* dst, const double* src, size_t n, hipStream_t st);
// one bead's columns of the xyz trajectory format, in upload order
struct XyzRec {
  double pos[3], vel[3], vbuf[3];
  float acc[3];
  uint32_t id, prop, amp_id, unit;
};
void launch_xyz_gather(const DevState& d, const double* vel, XyzRec* out, hipStream_t st);

}  // namespace mdpd
