// Saru counter-based PRNG for the device and the host.
//
// Same integer algorithm as the reference's CPU generator
// (include/saruprng.hpp:208-255 three-seed constructor, :199-202 and
// :322-334 one-step u32, :427-435 uni_open, :442-445 nrml with the first
// draw feeding log and the second cos), written here for HIP.  The u32
// stream is bit-exact; nrml_f32 is an FP32 evaluation whose error against
// the FP64 nrml is bounded in the tests.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mdpd {

struct Saru {
  uint32_t state, wstate;
};

__host__ __device__ __forceinline__ uint32_t asr(uint32_t x, int s) {
  return (uint32_t)(((int32_t)x) >> s);
}

__host__ __device__ __forceinline__ Saru saru_seed(uint32_t seed1, uint32_t seed2,
                                                   uint32_t seed3) {
  seed3 ^= (seed1 << 7) ^ (seed2 >> 6);
  seed2 += (seed1 >> 4) ^ (seed3 >> 15);
  seed1 ^= (seed2 << 9) + (seed3 << 8);
  seed3 ^= 0xA5366B4Du * ((seed2 >> 11) ^ (seed1 << 1));
  seed2 += 0x72BE1579u * ((seed1 << 4) ^ (seed3 >> 16));
  seed1 ^= 0x3F38A6EDu * ((seed3 >> 5) ^ asr(seed2, 22));
  seed2 += seed1 * seed3;
  seed1 += seed3 ^ (seed2 >> 2);
  seed2 ^= asr(seed2, 17);
  Saru s;
  s.state = 0x79dedea3u * (seed1 ^ asr(seed1, 14));
  s.wstate = (s.state + seed2) ^ asr(s.state, 8);
  s.state = s.state + (s.wstate * (s.wstate ^ 0xdddf97f5u));
  s.wstate = 0xABCB96F7u + (s.wstate >> 1);
  return s;
}

__host__ __device__ __forceinline__ uint32_t saru_u32(Saru& s) {
  s.state = 0x4beb5d59u * s.state + 0x2600e1f7u;
  s.wstate = s.wstate + 0x8009d14bu + (asr(s.wstate, 31) & 0xda879addu);
  const uint32_t v = (s.state ^ (s.state >> 26)) + s.wstate;
  return (v ^ (v >> 20)) * 0x6957f5a7u;
}

// (0,1] uniform in FP64, exact: ((int)(u >> 1) + 0.5) / 2^31
__host__ __device__ __forceinline__ double saru_uni_f64(Saru& s) {
  return ((double)(int32_t)(saru_u32(s) >> 1) + 0.5) * (1.0 / 2147483648.0);
}

// -ln(u) for u = (k + 0.5) / 2^31, k = x >> 1, in FP32.  Near u = 1 the FP32
// rounding of u would dominate, so there t = 1 - u is formed exactly from
// the integer and -ln(1 - t) = t (1 + t (1/2 + t/3)) (|t| <= 2^-8, relative
// truncation error below 2^-26).
__device__ __forceinline__ float neg_log_uni(uint32_t x) {
  const uint32_t k = x >> 1;
  const float u = ((float)(int32_t)k + 0.5f) * 4.656612873077392578125e-10f;
  const float g = -0.69314718055994530942f * __builtin_amdgcn_logf(u);
  const uint32_t m = 0x80000000u - k;  // 1 <= m <= 2^31
  const float t = ((float)m - 0.5f) * 4.656612873077392578125e-10f;
  const float near1 = t * (1.0f + t * (0.5f + t * 0.333333333f));
  return (m <= (1u << 23)) ? near1 : g;
}

// FP32 Box-Muller normal of the pair noise: sqrt(-2 ln u1) cos(2 pi u2).
// The hardware cosine takes its argument in turns, so cos(2 pi u2) is one
// v_cos_f32 of u2 in (0, 1] (absolute error below 3e-7 there).
__device__ __forceinline__ float saru_nrml_f32(Saru& s) {
  const uint32_t x1 = saru_u32(s);
  const uint32_t x2 = saru_u32(s);
  const float r = __builtin_amdgcn_sqrtf(2.0f * neg_log_uni(x1));
  const float u2 = ((float)(int32_t)(x2 >> 1) + 0.5f) * 4.656612873077392578125e-10f;
  return r * __builtin_amdgcn_cosf(u2);
}

}  // namespace mdpd
