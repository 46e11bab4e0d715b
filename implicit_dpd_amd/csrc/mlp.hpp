// Device view of the equation-of-state surrogate (4 -> 128 -> 128 -> 2).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mdpd {

constexpr int kIn = 4, kH = 128, kOut = 2;
// flat FP32 parameter layout: W1 [128][4], b1 [128], W2 [128][128], b2 [128],
// W3 [2][128], b3 [2] (weights row-major [out][in])
constexpr uint32_t kOffW1 = 0;
constexpr uint32_t kOffB1 = kOffW1 + kH * kIn;
constexpr uint32_t kOffW2 = kOffB1 + kH;
constexpr uint32_t kOffB2 = kOffW2 + kH * kH;
constexpr uint32_t kOffW3 = kOffB2 + kH;
constexpr uint32_t kOffB3 = kOffW3 + kOut * kH;
constexpr uint32_t kParams = kOffB3 + kOut;  // 17,410

// One 32 x 32 tile of a weight gradient dW[m][k] = sum_s G[m][s] H[k][s].
struct MlpTile {
  const uint16_t* G;   // [mvalid][npad] BF16
  const uint16_t* H;   // [kvalid][npad] BF16
  float* out;          // dW, row stride ld
  uint32_t m0, k0, mvalid, kvalid, ld;
};

struct MlpDev {
  float* param;        // [kParams] FP32 master copy
  float* grad;         // [kParams]
  float* adam_m;
  float* adam_v;
  const float* b1;     // views into param
  const float* b2;
  const float* b3;
  uint16_t* W1b;       // BF16 copies read by the products
  uint16_t* W2b;
  uint16_t* W2Tb;      // W2 transposed [in][out]
  uint16_t* W3b;
  // BF16 transposes [feature][sample] of the product inputs of one batch
  uint16_t* XT;        // [4][npad]
  uint16_t* H1T;       // [128][npad]
  uint16_t* H2T;       // [128][npad]
  uint16_t* DZ1T;      // [128][npad]
  uint16_t* DZ2T;      // [128][npad]
  uint16_t* DYT;       // [2][npad]
  uint32_t npad;       // batch capacity, a multiple of 32
  MlpTile* tiles;      // the 24 weight-gradient tiles
};

void launch_mlp_rows(const MlpDev& m, const float* x, const float* t, uint32_t n, bool train,
                     float* y, float* loss_part, hipStream_t st);
void launch_mlp_grads(const MlpDev& m, uint32_t n, uint32_t ntiles, hipStream_t st);
void launch_mlp_adam(const MlpDev& m, float lr, float b1, float b2, float eps, float c1, float c2,
                     bool apply, hipStream_t st);

}  // namespace mdpd
