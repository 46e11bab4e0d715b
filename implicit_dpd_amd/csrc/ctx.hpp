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
  // step, 4: the step computes its own cell keys)
  hipGraphExec_t gexec[2][8] = {};
  int parity = 0;            // which buffer set is current (flips every rebuild)
  uint32_t* inv_orig = nullptr;  // scratch [cap]
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
int check_flags(mdpd_ctx* c);
// device address of the fixed scratch of mdpd_rdf_noalloc on device `dev`
// (looked up once per device, at context creation)
int noalloc_scratch_init(int dev);
}  // namespace mdpd

#define HIPCHK(c, call)                                                                    \
  do {                                                                                     \
    hipError_t e_ = (call);                                                                \
    if (e_ != hipSuccess)                                                                  \
      return ::mdpd::fail((c), MDPD_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)
