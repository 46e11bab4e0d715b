// Host side of the equation-of-state surrogate: parameter initialisation,
// forward evaluation, loss/gradient and Adam training over mini-batches,
// all through the matrix-core kernels of k_mlp.hip.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <random>
#include <string>
#include <vector>

#include "../../include/mdpd_capi.h"
#include "common.hpp"
#include "mlp.hpp"

using namespace mdpd;

struct mdpd_mlp {
  hipStream_t st = nullptr;
  MlpDev m{};
  uint32_t cap = 0;       // samples per device batch (multiple of 32)
  float* dx = nullptr;    // [cap][4]
  float* dt = nullptr;    // [cap][2]
  float* dy = nullptr;    // [cap][2]
  float* loss = nullptr;  // [max minibatches per epoch][cap / 32]
  uint32_t loss_cap = 0;
  uint32_t step = 0;      // Adam steps taken
  std::vector<void*> allocs;
  std::string err;
};

namespace {

int mfail(mdpd_mlp* m, int code, const std::string& msg) {
  if (m) m->err = msg;
  return code;
}

#define MCHK(m, x)                                                              \
  do {                                                                          \
    hipError_t e_ = (x);                                                        \
    if (e_ != hipSuccess) return mfail(m, MDPD_ERR_HIP, std::string(#x) + ": " + \
                                              hipGetErrorString(e_));           \
  } while (0)

template <class T>
int malloc_dev(mdpd_mlp* m, T** p, size_t count) {
  MCHK(m, hipMalloc((void**)p, sizeof(T) * std::max<size_t>(count, 1)));
  MCHK(m, hipMemset(*p, 0, sizeof(T) * std::max<size_t>(count, 1)));
  m->allocs.push_back(*p);
  return MDPD_OK;
}

// PyTorch's default Linear initialisation: U(-1/sqrt(fan_in), 1/sqrt(fan_in))
// for weights and biases, from a seeded MT19937.
void init_params(std::vector<float>& p, uint32_t seed) {
  std::mt19937 g(seed);
  auto fill = [&](uint32_t off, uint32_t cnt, int fan_in) {
    std::uniform_real_distribution<float> u(-1.0f / std::sqrt((float)fan_in),
                                            1.0f / std::sqrt((float)fan_in));
    for (uint32_t i = 0; i < cnt; ++i) p[off + i] = u(g);
  };
  fill(kOffW1, kH * kIn, kIn);
  fill(kOffB1, kH, kIn);
  fill(kOffW2, kH * kH, kH);
  fill(kOffB2, kH, kH);
  fill(kOffW3, kOut * kH, kH);
  fill(kOffB3, kOut, kH);
}

int upload_params(mdpd_mlp* m, const float* p) {
  MCHK(m, hipMemcpyAsync(m->m.param, p, sizeof(float) * kParams, hipMemcpyHostToDevice, m->st));
  MCHK(m, hipMemsetAsync(m->m.adam_m, 0, sizeof(float) * kParams, m->st));
  MCHK(m, hipMemsetAsync(m->m.adam_v, 0, sizeof(float) * kParams, m->st));
  m->step = 0;
  launch_mlp_adam(m->m, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, false, m->st);  // BF16 copies only
  MCHK(m, hipGetLastError());
  MCHK(m, hipStreamSynchronize(m->st));
  return MDPD_OK;
}

uint32_t build_tiles(std::vector<MlpTile>& t, const MlpDev& d) {
  t.clear();
  for (uint32_t m0 = 0; m0 < (uint32_t)kH; m0 += 32)  // dW1 [128][4]
    t.push_back({d.DZ1T, d.XT, d.grad + kOffW1, m0, 0, (uint32_t)kH, (uint32_t)kIn, (uint32_t)kIn});
  for (uint32_t m0 = 0; m0 < (uint32_t)kH; m0 += 32)  // dW2 [128][128]
    for (uint32_t k0 = 0; k0 < (uint32_t)kH; k0 += 32)
      t.push_back({d.DZ2T, d.H1T, d.grad + kOffW2, m0, k0, (uint32_t)kH, (uint32_t)kH, (uint32_t)kH});
  for (uint32_t k0 = 0; k0 < (uint32_t)kH; k0 += 32)  // dW3 [2][128]
    t.push_back({d.DYT, d.H2T, d.grad + kOffW3, 0, k0, (uint32_t)kOut, (uint32_t)kH, (uint32_t)kH});
  return (uint32_t)t.size();
}

int ensure_loss(mdpd_mlp* m, uint32_t nparts) {
  if (nparts <= m->loss_cap) return MDPD_OK;
  if (m->loss) {
    auto it = std::find(m->allocs.begin(), m->allocs.end(), (void*)m->loss);
    if (it != m->allocs.end()) m->allocs.erase(it);
    MCHK(m, hipFree(m->loss));
    m->loss = nullptr;
  }
  int rc = malloc_dev(m, &m->loss, nparts);
  if (rc) return rc;
  m->loss_cap = nparts;
  return MDPD_OK;
}

// One forward/backward pass over n <= cap device-resident samples; leaves
// the gradient in m.grad and the per-wave loss partials in loss[part0..].
int pass(mdpd_mlp* m, const float* x, const float* t, uint32_t n, uint32_t part0) {
  launch_mlp_rows(m->m, x, t, n, true, nullptr, m->loss + part0, m->st);
  launch_mlp_grads(m->m, n, 24, m->st);
  MCHK(m, hipGetLastError());
  return MDPD_OK;
}

}  // namespace

extern "C" {

int mdpd_mlp_num_params(uint32_t* n) {
  if (!n) return MDPD_ERR_ARG;
  *n = kParams;
  return MDPD_OK;
}

int mdpd_mlp_create(uint32_t max_batch, uint32_t seed, mdpd_mlp** out) {
  if (!out || max_batch == 0 || max_batch > (1u << 24)) return MDPD_ERR_ARG;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return MDPD_ERR_NO_DEVICE;
  mdpd_mlp* m = new mdpd_mlp();
  m->cap = div_up(max_batch, 32) * 32;
  int rc = MDPD_OK;
  auto bail = [&](int code) {
    mdpd_mlp_destroy(m);
    return code;
  };
  if (hipStreamCreateWithFlags(&m->st, hipStreamNonBlocking) != hipSuccess) return bail(MDPD_ERR_HIP);
  MlpDev& d = m->m;
  d.npad = m->cap;
  const size_t np = d.npad;
  if ((rc = malloc_dev(m, &d.param, kParams)) || (rc = malloc_dev(m, &d.grad, kParams)) ||
      (rc = malloc_dev(m, &d.adam_m, kParams)) || (rc = malloc_dev(m, &d.adam_v, kParams)) ||
      (rc = malloc_dev(m, &d.W1b, (size_t)kH * kIn)) || (rc = malloc_dev(m, &d.W2b, (size_t)kH * kH)) ||
      (rc = malloc_dev(m, &d.W2Tb, (size_t)kH * kH)) || (rc = malloc_dev(m, &d.W3b, (size_t)kOut * kH)) ||
      (rc = malloc_dev(m, &d.XT, kIn * np)) || (rc = malloc_dev(m, &d.H1T, kH * np)) ||
      (rc = malloc_dev(m, &d.H2T, kH * np)) || (rc = malloc_dev(m, &d.DZ1T, kH * np)) ||
      (rc = malloc_dev(m, &d.DZ2T, kH * np)) || (rc = malloc_dev(m, &d.DYT, kOut * np)) ||
      (rc = malloc_dev(m, &m->dx, 4 * np)) || (rc = malloc_dev(m, &m->dt, 2 * np)) ||
      (rc = malloc_dev(m, &m->dy, 2 * np)) || (rc = malloc_dev(m, &d.tiles, 24)))
    return bail(rc);
  d.b1 = d.param + kOffB1;
  d.b2 = d.param + kOffB2;
  d.b3 = d.param + kOffB3;
  std::vector<MlpTile> tiles;
  build_tiles(tiles, d);
  if (hipMemcpy(d.tiles, tiles.data(), sizeof(MlpTile) * tiles.size(), hipMemcpyHostToDevice) != hipSuccess)
    return bail(MDPD_ERR_HIP);
  if ((rc = ensure_loss(m, m->cap / 32))) return bail(rc);
  std::vector<float> p(kParams);
  init_params(p, seed);
  if ((rc = upload_params(m, p.data()))) return bail(rc);
  *out = m;
  return MDPD_OK;
}

int mdpd_mlp_destroy(mdpd_mlp* m) {
  if (!m) return MDPD_ERR_ARG;
  if (m->st) hipStreamSynchronize(m->st);
  for (void* p : m->allocs) hipFree(p);
  if (m->st) hipStreamDestroy(m->st);
  delete m;
  return MDPD_OK;
}

const char* mdpd_mlp_last_error(const mdpd_mlp* m) { return m ? m->err.c_str() : "null surrogate"; }

int mdpd_mlp_set_params(mdpd_mlp* m, const float* p, uint32_t n) {
  if (!m || !p) return MDPD_ERR_ARG;
  if (n != kParams) return mfail(m, MDPD_ERR_ARG, "expected " + std::to_string(kParams) + " parameters");
  return upload_params(m, p);
}

int mdpd_mlp_get_params(mdpd_mlp* m, float* p, uint32_t n) {
  if (!m || !p) return MDPD_ERR_ARG;
  if (n != kParams) return mfail(m, MDPD_ERR_ARG, "expected " + std::to_string(kParams) + " parameters");
  MCHK(m, hipMemcpyAsync(p, m->m.param, sizeof(float) * kParams, hipMemcpyDeviceToHost, m->st));
  MCHK(m, hipStreamSynchronize(m->st));
  return MDPD_OK;
}

int mdpd_mlp_forward(mdpd_mlp* m, const float* x, uint32_t n, float* y) {
  if (!m || (n && (!x || !y))) return MDPD_ERR_ARG;
  for (uint32_t b = 0; b < n; b += m->cap) {
    const uint32_t k = std::min(m->cap, n - b);
    MCHK(m, hipMemcpyAsync(m->dx, x + 4 * (size_t)b, 16 * (size_t)k, hipMemcpyHostToDevice, m->st));
    launch_mlp_rows(m->m, m->dx, nullptr, k, false, m->dy, nullptr, m->st);
    MCHK(m, hipGetLastError());
    MCHK(m, hipMemcpyAsync(y + 2 * (size_t)b, m->dy, 8 * (size_t)k, hipMemcpyDeviceToHost, m->st));
  }
  MCHK(m, hipStreamSynchronize(m->st));
  return MDPD_OK;
}

int mdpd_mlp_loss_grad(mdpd_mlp* m, const float* x, const float* t, uint32_t n, double* loss,
                       float* grad) {
  if (!m || !x || !t || n == 0) return MDPD_ERR_ARG;
  if (n > m->cap) return mfail(m, MDPD_ERR_CAPACITY, "batch larger than max_batch");
  MCHK(m, hipMemcpyAsync(m->dx, x, 16 * (size_t)n, hipMemcpyHostToDevice, m->st));
  MCHK(m, hipMemcpyAsync(m->dt, t, 8 * (size_t)n, hipMemcpyHostToDevice, m->st));
  int rc = pass(m, m->dx, m->dt, n, 0);
  if (rc) return rc;
  const uint32_t nw = div_up(n, 32);
  std::vector<float> lp(nw);
  MCHK(m, hipMemcpyAsync(lp.data(), m->loss, 4 * (size_t)nw, hipMemcpyDeviceToHost, m->st));
  if (grad) MCHK(m, hipMemcpyAsync(grad, m->m.grad, 4 * (size_t)kParams, hipMemcpyDeviceToHost, m->st));
  MCHK(m, hipStreamSynchronize(m->st));
  if (loss) {
    double s = 0.0;
    for (float v : lp) s += v;
    *loss = s / (2.0 * n);
  }
  return MDPD_OK;
}

int mdpd_mlp_train(mdpd_mlp* m, const float* x, const float* t, uint32_t n, uint32_t epochs,
                   uint32_t batch, float lr, uint32_t seed, float* loss_per_epoch) {
  if (!m || !x || !t || n == 0 || batch == 0 || !(lr > 0.0f)) return MDPD_ERR_ARG;
  if (batch > m->cap) return mfail(m, MDPD_ERR_CAPACITY, "batch larger than max_batch");
  const uint32_t nmb = div_up(n, batch);
  const uint32_t wpb = div_up(batch, 32);
  int rc = ensure_loss(m, nmb * wpb);
  if (rc) return rc;
  // the whole data set, reshuffled on the host each epoch (fixed seed)
  float* ex = nullptr;
  float* et = nullptr;
  MCHK(m, hipMalloc((void**)&ex, 16 * (size_t)n));
  MCHK(m, hipMalloc((void**)&et, 8 * (size_t)n));
  std::vector<uint32_t> perm(n);
  std::iota(perm.begin(), perm.end(), 0u);
  std::mt19937 g(seed);
  std::vector<float> hx(4 * (size_t)n), ht(2 * (size_t)n), lp((size_t)nmb * wpb);
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  for (uint32_t ep = 0; ep < epochs && rc == MDPD_OK; ++ep) {
    std::shuffle(perm.begin(), perm.end(), g);
    for (uint32_t i = 0; i < n; ++i) {
      std::memcpy(&hx[4 * (size_t)i], x + 4 * (size_t)perm[i], 16);
      std::memcpy(&ht[2 * (size_t)i], t + 2 * (size_t)perm[i], 8);
    }
    if (hipMemcpyAsync(ex, hx.data(), 16 * (size_t)n, hipMemcpyHostToDevice, m->st) != hipSuccess ||
        hipMemcpyAsync(et, ht.data(), 8 * (size_t)n, hipMemcpyHostToDevice, m->st) != hipSuccess) {
      rc = mfail(m, MDPD_ERR_HIP, "upload of the training set failed");
      break;
    }
    for (uint32_t b = 0; b < nmb && rc == MDPD_OK; ++b) {
      const uint32_t off = b * batch, k = std::min(batch, n - off);
      rc = pass(m, ex + 4 * (size_t)off, et + 2 * (size_t)off, k, b * wpb);
      ++m->step;
      const float c1 = 1.0f / (1.0f - std::pow(b1, (float)m->step));
      const float c2 = 1.0f / (1.0f - std::pow(b2, (float)m->step));
      launch_mlp_adam(m->m, lr, b1, b2, eps, c1, c2, true, m->st);
    }
    if (loss_per_epoch && rc == MDPD_OK) {
      if (hipMemcpyAsync(lp.data(), m->loss, 4 * lp.size(), hipMemcpyDeviceToHost, m->st) != hipSuccess ||
          hipStreamSynchronize(m->st) != hipSuccess) {
        rc = mfail(m, MDPD_ERR_HIP, "loss readback failed");
        break;
      }
      // every partial is a sum of squared errors over its samples
      double s = 0.0;
      for (float v : lp) s += v;
      loss_per_epoch[ep] = (float)(s / (2.0 * n));
    }
  }
  hipStreamSynchronize(m->st);
  hipFree(ex);
  hipFree(et);
  if (rc == MDPD_OK && hipGetLastError() != hipSuccess) rc = mfail(m, MDPD_ERR_HIP, "training kernels failed");
  return rc;
}

}  // extern "C"
