// Engine context shared by the host translation units.
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../include/mdpd_capi.h"
#include "common.hpp"
#include "state.hpp"

namespace mdpd {
extern thread_local std::string g_err;
struct XyzWriter;  // asynchronous trajectory output (io.cpp)
void xyz_writer_free(XyzWriter* w);
constexpr int kPhases = 10;
extern const char* kPhaseNames[kPhases];
}  // namespace mdpd

struct mdpd_ctx {
  int dev = 0;
  hipStream_t st = nullptr;
  std::string err;
  std::vector<mdpd_params> P;
  std::vector<mdpd::Replica> R;
  std::vector<uint32_t> rep_off, cell_off, amp_off, amp_num_rep, sol_num_rep;
  int n_rep = 1;
  uint32_t cap = 0, n = 0, ncell = 0, amp_cap = 0, n_amp = 0;
  uint32_t ncell_alloc = 0;  // cells the cell arrays hold (the cutoff grid)
  int ns = 1, nd = 1, env = 0;
  bool any_conv = false;
  uint32_t cm_corr = 100;
  mdpd_tuning tune{};
  mdpd::DevState d{};
  double* vir_bonded = nullptr;  // [n_rep][3]
  double* obs_out = nullptr;     // [n_rep][16]
  uint32_t* d_amp_num = nullptr; // [n_rep]
  void* pinned = nullptr;
  size_t pinned_bytes = 0;
  uint32_t time = 0;
  bool drifted = false, have_cells = false, cells_fresh = false, have_density = false;
  // skin mode: the skin list, its cells and the displacement reference match
  // the bead storage (cleared by every call that moves or reorders beads)
  bool skin_valid = false;
  // captured steps by storage parity and variant (1: virial step, 2: CM-drift
  // step, 4: the step computes its own cell keys, 8: the step ends with the
  // chemistry event)
  hipGraphExec_t gexec[2][16] = {};
  // chemistry (single system).  Slot 0 of chem_dev / chem_state serves the
  // event inside mdpd_run, slot 1 the explicit mdpd_chem_event; the counts
  // and targets scratch is shared (the two never overlap on the stream).
  mdpd::ChemDev* chem_dev = nullptr;      // [2]
  mdpd::ChemState* chem_state = nullptr;  // [2]
  uint32_t* chem_counts = nullptr;        // [kMaxCore * chem_count_blocks(cap)]
  uint32_t* chem_targets = nullptr;       // [amp_cap]
  unsigned long long* chem_masks = nullptr;  // [chem_mask_words(cap)] flags per (core, wave)
  uint32_t* run_ok_word = nullptr;        // the device's run word of grow mode
  mdpd::ChemDev chem_host[2] = {};        // host copies (sources of the uploads)
  // reference-MT mode (mdpd_set_chem_mt): chem_mt is the buffers of
  // mt_store when on, all null (keyed draws) when off; mt_store is allocated
  // on the first switch-on and kept
  mdpd::ChemMt chem_mt = {};
  mdpd::ChemMt mt_store = {};
  bool grow = false;          // chemistry may append beads: see DevState::nl
  bool chem_on = false;       // mdpd_run ends steps with time >= beg_chem with an event
  bool chem_stopped = false;  // an in-run event reached max_amp_num
  uint32_t beg_chem = 0;
  int parity = 0;            // which buffer set is current (flips every rebuild)
  double* dl_load = nullptr;  // observer scratch: delta_sumr in load order, [3][cap]
  bool profiling = false;
  // profiling: kPhases + 1 events per profiled step, read back lazily
  std::vector<hipEvent_t> evpool;
  uint32_t ev_steps = 0;          // profiled steps recorded in evpool
  double ph_ms[mdpd::kPhases + 1] = {0};  // accumulated; last = whole step
  uint64_t ph_steps = 0;
  std::vector<void*> allocs;
  mdpd::XyzWriter* xyz = nullptr;
};

namespace mdpd {
int fail(mdpd_ctx* c, int code, const std::string& msg);
void invalidate_graphs(mdpd_ctx* c);
void swap_master(mdpd_ctx* c);
int run_steps(mdpd_ctx* c, uint32_t nsteps);
int prepare_graphs(mdpd_ctx* c, uint32_t nsteps);
int check_flags(mdpd_ctx* c);
// device address of the fixed scratch of mdpd_rdf_noalloc on device `dev`
// (looked up once per device, at context creation)
int noalloc_scratch_init(int dev);
// grow mode: per-bead launches cover the capacity and kernels read the live
// counts on the device, so chemistry never forces a graph re-capture
int enter_grow(mdpd_ctx* c);
inline void set_launch_bound(mdpd_ctx* c) { c->d.nl = (c->grow && c->d.n) ? c->cap : c->d.n; }
// the host's copies of the counts after device-side chemistry
int pull_counts(mdpd_ctx* c);
// reference-MT chemistry: switch the mode on from a stream state; read the
// state of the stream in use
int chem_mt_enable(mdpd_ctx* c, const MtState& h);
int chem_mt_state(mdpd_ctx* c, MtState* h);
uint32_t chem_count_blocks(uint32_t cap);
size_t chem_mask_words(uint32_t cap);
void launch_chem_event(const DevState& d, const ChemDev* cd, ChemState* cs, uint32_t n_core,
                       uint32_t* counts, unsigned long long* masks, uint32_t* targets,
                       uint32_t* amp_num, bool keys, const ChemMt& mt, hipStream_t st);
}  // namespace mdpd

#define HIPCHK(c, call)                                                                    \
  do {                                                                                     \
    hipError_t e_ = (call);                                                                \
    if (e_ != hipSuccess)                                                                  \
      return ::mdpd::fail((c), MDPD_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)
