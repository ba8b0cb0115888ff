// Shared device helpers and host error plumbing for libssmhip.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>

#include "ssmhip.h"

namespace ssmhip {

// ---- host error handling ----------------------------------------------------
void set_error(const std::string& msg);
int32_t fail(int32_t code, const std::string& msg);
int32_t hip_fail(hipError_t e, const char* what);

#define SSM_HIP(call)                                   \
  do {                                                  \
    hipError_t _e = (call);                             \
    if (_e != hipSuccess) return ::ssmhip::hip_fail(_e, #call); \
  } while (0)

int device_cu_count();

// ---- surrogate MLP training launch arguments (mlp_kernel.hip) ------------------------
// mode 0: dy = grad_y[row]; mode 1: Huber(y - target, delta) times `scale`.
// acts / deltas: feature-major bf16 [rows][ld] (NULL = not stored); dx [n][8];
// y [n]; loss [n] (all optional).
struct TrainArgs {
  const float* x;
  const float* target_or_grad;
  const __bf16* frags;
  const float* bias;
  const float* xnorm;  // optional shift[7], scale[7]: the network sees (x - shift) * scale
  __bf16* acts;
  __bf16* deltas;
  float* dx;
  float* y;
  float* loss;
  int64_t n;
  int64_t ld;
  int32_t mode;
  float delta;
  float scale;
};

// ---- Philox4x32-10 -------------------------------------------------------------
// Counter-based generator of Salmon, Moraes, Dror and Shaw, "Parallel random
// numbers: as easy as 1, 2, 3" (SC'11): 10 rounds of two 32x32->64 multiplies
// with the published multipliers and Weyl key increments.
struct u32x4 {
  uint32_t x, y, z, w;
};

__device__ __forceinline__ u32x4 philox4x32_10(u32x4 c, uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    const uint64_t p0 = (uint64_t)0xD2511F53u * (uint64_t)c.x;
    const uint64_t p1 = (uint64_t)0xCD9E8D57u * (uint64_t)c.z;
    const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    // three-input XORs as one gfx950 v_bitop3_b32 each (truth table 0x96 = a ^ b ^ c)
    c.x = __builtin_amdgcn_bitop3_b32(hi1, c.y, k0, 0x96);
    c.y = lo1;
    c.z = __builtin_amdgcn_bitop3_b32(hi0, c.w, k1, 0x96);
    c.w = lo0;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  return c;
}

// uint32 -> float in (0, 1]: (u + 0.5) * 2^-32 rounded to float32; never 0.
__device__ __forceinline__ float u32_to_open01(uint32_t u) {
  return __builtin_fmaf((float)u, 2.3283064365386963e-10f, 1.1641532182693481e-10f);
}
// uint32 -> float in [0, 1): top 24 bits * 2^-24 (exact, never 1).
__device__ __forceinline__ float u32_to_unit(uint32_t u) {
  return (float)(u >> 8) * 5.9604644775390625e-08f;
}

// uint32 -> angle in revolutions, 1 + (top 23 bits) * 2^-23 in [1, 2): the bits placed under
// the exponent of 1.0f by one v_alignbit_b32 (0x7F:u >> 9). v_sin_f32 / v_cos_f32 take their
// argument as x * 2 pi and reduce it to its fraction, which is exact here (x - 1), so the
// angle is 2 pi (top 23 bits) 2^-23.
__device__ __forceinline__ float u32_to_rev(uint32_t u) {
  return __builtin_bit_cast(float, __builtin_amdgcn_alignbit(0x7Fu, u, 9u));
}

// -2 ln 2: -2 ln(u) = kBoxMullerLog2 * log2(u)
constexpr float kBoxMullerLog2 = -1.3862943611198906f;

// Box–Muller (Box & Muller 1958): two uniforms -> two independent N(0, c / (-2 ln 2)).
// r = sqrt(c log2 u1) with the hardware log2; c = -2 ln 2 gives N(0, 1), and
// c = -2 ln 2 s^2 gives normals already scaled by s (|s|, the same law), so a walk's step
// needs no multiply for its noise scale.
__device__ __forceinline__ void box_muller_c(uint32_t a, uint32_t b, float c, float& z0, float& z1) {
  const float u1 = u32_to_open01(a);
  const float rev = u32_to_rev(b);
  const float r = __builtin_amdgcn_sqrtf(c * __builtin_amdgcn_logf(u1));
  z0 = r * __builtin_amdgcn_cosf(rev);
  z1 = r * __builtin_amdgcn_sinf(rev);
}
__device__ __forceinline__ void box_muller(uint32_t a, uint32_t b, float& z0, float& z1) {
  box_muller_c(a, b, kBoxMullerLog2, z0, z1);
}

// four normals from one Philox4x32-10 block, scaled by s where c = -2 ln 2 s^2
__device__ __forceinline__ void philox_normals4_c(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
                                                  uint32_t k0, uint32_t k1, float c, float out[4]) {
  u32x4 r = philox4x32_10(u32x4{c0, c1, c2, c3}, k0, k1);
  box_muller_c(r.x, r.y, c, out[0], out[1]);
  box_muller_c(r.z, r.w, c, out[2], out[3]);
}
__device__ __forceinline__ void philox_normals4(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
                                                uint32_t k0, uint32_t k1, float out[4]) {
  philox_normals4_c(c0, c1, c2, c3, k0, k1, kBoxMullerLog2, out);
}

// ---- wavefront helpers (wave64) ----------------------------------------------
__device__ __forceinline__ int lane_id() { return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0)); }

__device__ __forceinline__ int lanes_below(uint64_t mask) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0));
}

__device__ __forceinline__ uint64_t readlane_u64(uint64_t v, int lane) {
  const uint32_t lo = __builtin_amdgcn_readlane((int)(uint32_t)v, lane);
  const uint32_t hi = __builtin_amdgcn_readlane((int)(uint32_t)(v >> 32), lane);
  return ((uint64_t)hi << 32) | lo;
}

}  // namespace ssmhip
