// Allocation-free pair-distance histogram: the work of mdpd_rdf with every
// piece of its scratch (squared bin edges, the 64-bit counters and their
// FP64 copy) carved out of one fixed device array instead of per-call
// hipMalloc/hipFree.  For callers that cannot allocate device memory (code
// under graph capture, serving loops).  No counterpart in the reference.
//
// The array exists once per GPU for the life of the process.  Its device
// address is looked up per device with hipGetSymbolAddress when a context
// is created on that device (the array's name itself is only a host-side
// stand-in and never reaches a kernel).  Every call on one GPU shares the
// array: calls that overlap on a GPU (several host threads, or contexts on
// different streams of the same GPU) race and must be serialised by the
// caller.  Each call clears the counters it uses with a fill kernel (no
// memset node), so the contents a previous call left behind do not matter.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "ctx.hpp"

namespace {

constexpr int kNoallocMaxBins = 1024;            // per replica, as mdpd_rdf
constexpr size_t kNoallocMaxCounters = 65536;    // n_rep * nbins
constexpr size_t kEdgeBytes = 8448;              // (kNoallocMaxBins + 1) doubles, 256-aligned
constexpr size_t kHistOff = kEdgeBytes;
constexpr size_t kOutOff = kHistOff + 8 * kNoallocMaxCounters;
constexpr size_t kScratchBytes = kOutOff + 8 * kNoallocMaxCounters;  // 1,057,024 bytes
static_assert(8 * (kNoallocMaxBins + 1) <= kEdgeBytes, "edge region too small");
static_assert(kEdgeBytes % 256 == 0 && kOutOff % 256 == 0, "regions 256-byte aligned");

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
unsigned char* g_scratch_dev[kMaxDevices] = {};  // device address of `scratch`, per device

// e2[k] = (k * width)^2 for k <= nb (the host's sq_edges of mdpd_rdf, same
// FP64 operations), and hist[0..cnt) = 0
__global__ void k_noalloc_rdf_prep(double* __restrict__ e2, int nb, double width,
                                   unsigned long long* __restrict__ hist, uint32_t cnt) {
#pragma clang fp contract(off)
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i <= (uint32_t)nb) {
    const double x = (double)(int)i * width;
    e2[i] = x * x;
  }
  if (i < cnt) hist[i] = 0ull;
}

__global__ void k_noalloc_u64_to_f64(const unsigned long long* __restrict__ h,
                                     double* __restrict__ out, uint32_t cnt) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < cnt) out[i] = (double)h[i];
}

}  // namespace

namespace mdpd {

// Looks up this device's copy of `scratch` once (the device current at
// context creation); later calls on the same device return at once.
int noalloc_scratch_init(int dev) {
  if (dev < 0 || dev >= kMaxDevices) return MDPD_ERR_UNSUPPORTED;
  std::lock_guard<std::mutex> lk(g_mu);
  if (g_scratch_dev[dev]) return MDPD_OK;
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p) {
    g_err = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
    return MDPD_ERR_HIP;
  }
  g_scratch_dev[dev] = (unsigned char*)p;
  return MDPD_OK;
}

}  // namespace mdpd

extern "C" int mdpd_rdf_noalloc(mdpd_ctx* c, double rmax, int32_t nbins, double* out) {
  using namespace mdpd;
  if (!c || !out || !(rmax > 0.0 && rmax <= 1.0) || nbins < 1 || nbins > kNoallocMaxBins)
    return fail(c, MDPD_ERR_ARG, "rdf: 0 < rmax <= 1 and 1 <= nbins <= 1024");
  const size_t cnt = (size_t)c->n_rep * (size_t)nbins;
  if (cnt > kNoallocMaxCounters)
    return fail(c, MDPD_ERR_CAPACITY,
                "rdf_noalloc: n_rep * nbins = " + std::to_string(cnt) +
                    " exceeds the fixed scratch of 65536 counters");
  if (!c->have_cells || !c->cells_fresh) return fail(c, MDPD_ERR_STATE, "no current cell list");
  unsigned char* base = nullptr;
  {
    std::lock_guard<std::mutex> lk(g_mu);
    if (c->dev >= 0 && c->dev < kMaxDevices) base = g_scratch_dev[c->dev];
  }
  if (!base) return fail(c, MDPD_ERR_STATE, "rdf_noalloc: scratch of this device not set up");
  double* e2 = (double*)base;
  unsigned long long* hist = (unsigned long long*)(base + kHistOff);
  double* res = (double*)(base + kOutOff);
  const uint32_t m = (uint32_t)cnt > (uint32_t)nbins + 1 ? (uint32_t)cnt : (uint32_t)nbins + 1;
  hipLaunchKernelGGL(k_noalloc_rdf_prep, dim3(div_up(m, 256)), dim3(256), 0, c->st, e2, nbins,
                     rmax / nbins, hist, (uint32_t)cnt);
  launch_rdf(c->d, e2, nbins, rmax, hist, c->st);
  hipLaunchKernelGGL(k_noalloc_u64_to_f64, dim3(div_up((uint32_t)cnt, 256)), dim3(256), 0, c->st,
                     hist, res, (uint32_t)cnt);
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipMemcpyAsync(out, res, 8 * cnt, hipMemcpyDeviceToHost, c->st));
  HIPCHK(c, hipStreamSynchronize(c->st));
  return MDPD_OK;
}
