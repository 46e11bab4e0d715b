// Equation-of-state surrogate: a 4 -> 128 -> 128 -> 2 perceptron with SiLU
// hidden activations (no counterpart in the reference: it proposes MDPD
// parameters (A, B, rd, kBT) for a target liquid density and pressure).
//
// Every matrix product runs on the matrix cores with BF16 operands and
// FP32 accumulation: gfx950's v_mfma_f32_32x32x16_bf16 (K = 16 per
// instruction) for the products over the 128 hidden features and over the
// batch, v_mfma_f32_32x32x8_bf16 for the two whose K is 4 or 2 (layer 1 and
// dY W3).  Operand layout for lane l (r = l % 32, h = l / 32):
//   A (32 x 16): A[r][8h + 0..7]      (x8 form: A[r][4h + 0..3])
//   B (16 x 32): B[8h + 0..7][r]      (x8 form: B[4h + 0..3][r])
//   D (32 x 32): D[8(q/4) + 4h + q%4][r] in accumulator q = 0..15 (both)
// Activations are rounded to BF16 (round to nearest even) where they enter
// a product; biases, pre-activations, SiLU and the optimizer stay FP32.
//
// Kernels:
//   k_mlp_rows   one wavefront per 32 samples: forward pass and, when
//                training, the backward pass down to the pre-activation
//                gradients; writes BF16 transposes of every product input
//                the weight gradients need
//   k_mlp_wgrad  weight gradients dW = G^T H summed over the batch, one
//                wavefront per 32 x 32 tile, fixed order (deterministic);
//                bias gradients as FP32 column sums
//   k_mlp_adam   Adam step on the FP32 master parameters, refreshes the
//                BF16 copies the products read
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "common.hpp"
#include "mlp.hpp"

namespace mdpd {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx16 __attribute__((ext_vector_type(16)));

__device__ __forceinline__ uint16_t to_bf16(float f) {
  const uint32_t u = __float_as_uint(f);
  return (uint16_t)((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);
}
__device__ __forceinline__ float from_bf16(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }

__device__ __forceinline__ floatx16 mfma(bf16x4 a, bf16x4 b, floatx16 c) {
  return __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ floatx16 mfma16(bf16x8 a, bf16x8 b, floatx16 c) {
  return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
}
// eight consecutive BF16 values (16-byte aligned)
__device__ __forceinline__ bf16x8 load8(const uint16_t* p) {
  return __builtin_bit_cast(bf16x8, *reinterpret_cast<const uint4*>(p));
}
__device__ __forceinline__ bf16x8 zero8() { return __builtin_bit_cast(bf16x8, make_uint4(0u, 0u, 0u, 0u)); }

__device__ __forceinline__ bf16x4 pack4(uint16_t a, uint16_t b, uint16_t c, uint16_t d) {
  bf16x4 v;
  v[0] = (short)a;
  v[1] = (short)b;
  v[2] = (short)c;
  v[3] = (short)d;
  return v;
}
__device__ __forceinline__ bf16x4 load4(const uint16_t* p) {
  const uint2 u = *reinterpret_cast<const uint2*>(p);
  bf16x4 v;
  v[0] = (short)(u.x & 0xFFFFu);
  v[1] = (short)(u.x >> 16);
  v[2] = (short)(u.y & 0xFFFFu);
  v[3] = (short)(u.y >> 16);
  return v;
}

__device__ __forceinline__ float silu(float z) { return z / (1.0f + expf(-z)); }
__device__ __forceinline__ float dsilu(float z) {
  const float s = 1.0f / (1.0f + expf(-z));
  return s * (1.0f + z * (1.0f - s));
}

// Row index of accumulator q of lane half h inside a 32-row tile.
__device__ __forceinline__ int drow(int q, int h) { return 8 * (q >> 2) + 4 * h + (q & 3); }

// Stores a 32 x 32 accumulator tile (columns col0..col0+31) as BF16 into
// the wave's LDS tile [32][ld] and into the transposed global buffer
// T[col][sample] (4 consecutive samples per store).
__device__ __forceinline__ void store_tile_bf16(const floatx16& v, uint16_t* lds, int ld, int col0,
                                                uint16_t* T, uint32_t npad, uint32_t row0, int r,
                                                int h) {
#pragma unroll
  for (int g = 0; g < 4; ++g) {
    uint16_t b[4];
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      b[e] = to_bf16(v[4 * g + e]);
      if (lds) lds[drow(4 * g + e, h) * ld + col0 + r] = b[e];
    }
    if (T) {
      const uint32_t s = row0 + 8 * g + 4 * h;
      uint2 u;
      u.x = (uint32_t)b[0] | ((uint32_t)b[1] << 16);
      u.y = (uint32_t)b[2] | ((uint32_t)b[3] << 16);
      *reinterpret_cast<uint2*>(T + (size_t)(col0 + r) * npad + s) = u;
    }
  }
}

template <bool TRAIN>
__global__ void __launch_bounds__(64) k_mlp_rows(MlpDev m, const float* __restrict__ x,
                                                 const float* __restrict__ t, uint32_t n,
                                                 float inv_n, float* __restrict__ y,
                                                 float* __restrict__ loss_part) {
  __shared__ __align__(16) uint16_t act[32 * kH];  // 32 samples x 128 features, BF16
  __shared__ __align__(16) float dyl[32 * 2];
  const int l = threadIdx.x, r = l & 31, h = l >> 5;
  const uint32_t row0 = blockIdx.x * 32u;
  const uint32_t npad = m.npad;

  // ---- layer 1: Z1 = X W1^T + b1 (K = 4 padded to 8)
  bf16x4 ax = pack4(0, 0, 0, 0);
  {
    const uint32_t s = row0 + r;
    if (h == 0 && s < n) {
      const float4 xv = *reinterpret_cast<const float4*>(x + 4 * (size_t)s);
      ax = pack4(to_bf16(xv.x), to_bf16(xv.y), to_bf16(xv.z), to_bf16(xv.w));
    }
    if (TRAIN) {
      // X^T for the layer-1 weight gradient: XT[k][s], k < 4
      if (h == 0) {
        const uint32_t s2 = row0 + r;
        for (int k = 0; k < 4; ++k) m.XT[(size_t)k * npad + s2] = (uint16_t)ax[k];
      }
    }
  }
  floatx16 z1[4];
#pragma unroll
  for (int nt = 0; nt < 4; ++nt) {
    const bf16x4 bw = h == 0 ? load4(m.W1b + (size_t)(32 * nt + r) * 4) : pack4(0, 0, 0, 0);
    floatx16 c = {};
    c = mfma(ax, bw, c);
    const float bias = m.b1[32 * nt + r];
    floatx16 hv;
#pragma unroll
    for (int q = 0; q < 16; ++q) {
      c[q] += bias;
      hv[q] = silu(c[q]);
    }
    z1[nt] = c;
    store_tile_bf16(hv, act, kH, 32 * nt, TRAIN ? m.H1T : nullptr, npad, row0, r, h);
  }
  __syncthreads();

  // ---- layer 2: Z2 = H1 W2^T + b2
  floatx16 z2[4];
#pragma unroll
  for (int nt = 0; nt < 4; ++nt) z2[nt] = floatx16{};
  for (int ks = 0; ks < kH / 16; ++ks) {
    const bf16x8 a = load8(act + r * kH + 16 * ks + 8 * h);
#pragma unroll
    for (int nt = 0; nt < 4; ++nt)
      z2[nt] = mfma16(a, load8(m.W2b + (size_t)(32 * nt + r) * kH + 16 * ks + 8 * h), z2[nt]);
  }
  __syncthreads();
#pragma unroll
  for (int nt = 0; nt < 4; ++nt) {
    const float bias = m.b2[32 * nt + r];
    floatx16 hv;
#pragma unroll
    for (int q = 0; q < 16; ++q) {
      z2[nt][q] += bias;
      hv[q] = silu(z2[nt][q]);
    }
    store_tile_bf16(hv, act, kH, 32 * nt, TRAIN ? m.H2T : nullptr, npad, row0, r, h);
  }
  __syncthreads();

  // ---- layer 3: Y = H2 W3^T + b3 (N = 2 padded to 32)
  floatx16 yo = {};
  for (int ks = 0; ks < kH / 16; ++ks) {
    const bf16x8 a = load8(act + r * kH + 16 * ks + 8 * h);
    const bf16x8 b = r < kOut ? load8(m.W3b + (size_t)r * kH + 16 * ks + 8 * h) : zero8();
    yo = mfma16(a, b, yo);
  }
  float lsum = 0.f;
  if (r < kOut) {
    const float bias = m.b3[r];
#pragma unroll
    for (int q = 0; q < 16; ++q) {
      const uint32_t s = row0 + drow(q, h);
      const float v = yo[q] + bias;
      if (s < n) {
        if (!TRAIN) y[2 * (size_t)s + r] = v;
        if (TRAIN) {
          const float dlt = v - t[2 * (size_t)s + r];
          lsum += dlt * dlt;
          dyl[drow(q, h) * 2 + r] = dlt * inv_n;  // d(loss)/dY, loss = sum (y-t)^2 / (2n)
        }
      } else if (TRAIN) {
        dyl[drow(q, h) * 2 + r] = 0.f;
      }
    }
  }
  if (!TRAIN) return;
  // per-wave loss partial (lanes in a fixed order)
  for (int off = 32; off > 0; off >>= 1) lsum += __shfl_down(lsum, off);
  if (l == 0) loss_part[blockIdx.x] = lsum;
  __syncthreads();

  // ---- backward: dY^T for dW3, dH2 = dY W3 (K = 2 padded to 8)
  uint16_t dyb[2] = {to_bf16(dyl[r * 2]), to_bf16(dyl[r * 2 + 1])};
  if (h == 0) {
    m.DYT[(size_t)0 * npad + row0 + r] = dyb[0];
    m.DYT[(size_t)1 * npad + row0 + r] = dyb[1];
  }
  const bf16x4 ady = h == 0 ? pack4(dyb[0], dyb[1], 0, 0) : pack4(0, 0, 0, 0);
#pragma unroll
  for (int nt = 0; nt < 4; ++nt) {
    const bf16x4 b = h == 0 ? pack4(m.W3b[32 * nt + r], m.W3b[kH + 32 * nt + r], 0, 0)
                            : pack4(0, 0, 0, 0);
    floatx16 c = {};
    c = mfma(ady, b, c);
#pragma unroll
    for (int q = 0; q < 16; ++q) c[q] *= dsilu(z2[nt][q]);  // dZ2
    store_tile_bf16(c, act, kH, 32 * nt, m.DZ2T, npad, row0, r, h);
  }
  __syncthreads();

  // ---- dH1 = dZ2 W2, dZ1 = dH1 * silu'(Z1)
  floatx16 d1[4];
#pragma unroll
  for (int nt = 0; nt < 4; ++nt) d1[nt] = floatx16{};
  for (int ks = 0; ks < kH / 16; ++ks) {
    const bf16x8 a = load8(act + r * kH + 16 * ks + 8 * h);
#pragma unroll
    for (int nt = 0; nt < 4; ++nt)
      d1[nt] = mfma16(a, load8(m.W2Tb + (size_t)(32 * nt + r) * kH + 16 * ks + 8 * h), d1[nt]);
  }
#pragma unroll
  for (int nt = 0; nt < 4; ++nt) {
#pragma unroll
    for (int q = 0; q < 16; ++q) d1[nt][q] *= dsilu(z1[nt][q]);
    store_tile_bf16(d1[nt], nullptr, kH, 32 * nt, m.DZ1T, npad, row0, r, h);
  }
}

// dW[mrow][kcol] = sum_s G[mrow][s] * H[kcol][s] over the padded batch, for
// one 32 x 32 tile; rows >= mvalid / cols >= kvalid read as zero.
__global__ void __launch_bounds__(64) k_mlp_wgrad(MlpDev m, uint32_t nsteps16) {
  const int l = threadIdx.x, r = l & 31, h = l >> 5;
  const MlpTile tl = m.tiles[blockIdx.x];
  const uint32_t npad = m.npad;
  const bool arow = tl.m0 + r < tl.mvalid, bcol = tl.k0 + r < tl.kvalid;
  const uint16_t* ga = tl.G + (size_t)(tl.m0 + r) * npad + 8 * h;
  const uint16_t* hb = tl.H + (size_t)(tl.k0 + r) * npad + 8 * h;
  floatx16 c = {};
  for (uint32_t st = 0; st < nsteps16; ++st) {
    const bf16x8 a = arow ? load8(ga + 16 * st) : zero8();
    const bf16x8 b = bcol ? load8(hb + 16 * st) : zero8();
    c = mfma16(a, b, c);
  }
#pragma unroll
  for (int q = 0; q < 16; ++q) {
    const uint32_t mr = tl.m0 + drow(q, h), kc = tl.k0 + r;
    if (mr < tl.mvalid && kc < tl.kvalid) tl.out[(size_t)mr * tl.ld + kc] = c[q];
  }
}

// Bias gradients: column sums of the BF16 pre-activation gradients, FP32,
// sequential over the batch (one thread per output unit).
__global__ void __launch_bounds__(256) k_mlp_bgrad(MlpDev m, uint32_t n) {
  const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  const uint16_t* src;
  float* dst;
  if (j < kH) { src = m.DZ1T + (size_t)j * m.npad; dst = m.grad + kOffB1 + j; }
  else if (j < 2 * kH) { src = m.DZ2T + (size_t)(j - kH) * m.npad; dst = m.grad + kOffB2 + (j - kH); }
  else if (j < 2 * kH + kOut) { src = m.DYT + (size_t)(j - 2 * kH) * m.npad; dst = m.grad + kOffB3 + (j - 2 * kH); }
  else return;
  float s = 0.f;
  for (uint32_t i = 0; i < n; ++i) s += from_bf16(src[i]);
  *dst = s;
}

__global__ void __launch_bounds__(256) k_mlp_adam(MlpDev m, float lr, float b1, float b2, float eps,
                                                  float c1, float c2, int apply) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= kParams) return;
  float p = m.param[i];
  if (apply) {
    const float g = m.grad[i];
    const float mm = b1 * m.adam_m[i] + (1.0f - b1) * g;
    const float vv = b2 * m.adam_v[i] + (1.0f - b2) * g * g;
    m.adam_m[i] = mm;
    m.adam_v[i] = vv;
    p -= lr * (mm * c1) / (sqrtf(vv * c2) + eps);
    m.param[i] = p;
  }
  // BF16 copies: W1 [128][4], W2 [128][128] and its transpose, W3 [2][128]
  const uint16_t b = to_bf16(p);
  if (i < kOffB1) m.W1b[i] = b;
  else if (i >= kOffW2 && i < kOffB2) {
    const uint32_t k = i - kOffW2, o = k / kH, in = k % kH;
    m.W2b[k] = b;
    m.W2Tb[(size_t)in * kH + o] = b;
  } else if (i >= kOffW3 && i < kOffB3) {
    m.W3b[i - kOffW3] = b;
  }
}

void launch_mlp_rows(const MlpDev& m, const float* x, const float* t, uint32_t n, bool train,
                     float* y, float* loss_part, hipStream_t st) {
  const uint32_t nb = div_up(n, 32);
  if (!nb) return;
  if (train)
    hipLaunchKernelGGL(k_mlp_rows<true>, dim3(nb), dim3(64), 0, st, m, x, t, n, 1.0f / (float)n, y,
                       loss_part);
  else
    hipLaunchKernelGGL(k_mlp_rows<false>, dim3(nb), dim3(64), 0, st, m, x, t, n, 0.0f, y,
                       loss_part);
}

void launch_mlp_grads(const MlpDev& m, uint32_t n, uint32_t ntiles, hipStream_t st) {
  const uint32_t npad32 = div_up(n, 32) * 32;
  hipLaunchKernelGGL(k_mlp_wgrad, dim3(ntiles), dim3(64), 0, st, m, npad32 / 16);
  hipLaunchKernelGGL(k_mlp_bgrad, dim3(2), dim3(256), 0, st, m, npad32);
}

void launch_mlp_adam(const MlpDev& m, float lr, float b1, float b2, float eps, float c1, float c2,
                     bool apply, hipStream_t st) {
  hipLaunchKernelGGL(k_mlp_adam, dim3(div_up(kParams, 256)), dim3(256), 0, st, m, lr, b1, b2, eps,
                     c1, c2, apply ? 1 : 0);
}

}  // namespace mdpd
