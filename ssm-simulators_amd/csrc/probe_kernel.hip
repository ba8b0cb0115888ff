// VALU throughput probes: the compute roofs the EM and KDE kernels are
// measured against. Each probe runs a branch-free loop at full occupancy:
//   kind 0  EM block: Philox4x32-10 + 2 Box–Muller pairs + 4 angle-DDM steps
//           (the simulator's inner loop without walk termination or refill)
//   kind 1  KDE pair: sub, fma, exp2, add (the KDE evaluation inner loop)
//   kind 2..8  single instructions, 8 independent chains per lane:
//           2 v_mad_u64_u32, 3 v_xor_b32, 4 v_fma_f32, 5 v_log_f32,
//           6 v_exp_f32, 7 v_sin_f32, 8 v_sqrt_f32
//   kind 9  the surrogate's hidden unit u = 1 / (2^z + 1): exp2, add, rcp (+ one fma),
//           8 chains per lane
//   kind 10 LCA step, 4 accumulators, constant bound: bound test, Philox4x32-10 + 2
//           Box–Muller pairs, the leaky, inhibited update with its float32 sum
//   kind 11 race step, 4 accumulators, Weibull bound: the same with the race update
// Every lane writes one float so nothing is optimised away.
#include <hip/hip_runtime.h>

#include "ssmhip_internal.h"

namespace ssmhip {

__global__ void __launch_bounds__(256) probe_em_block(int64_t iters, float* out) {
  // the simulator's step as its iterations far from the end of t_s and from the deadline run
  // it, in the same form (float32 rounding of every operation, no contraction; exec-mask
  // branches): the bound test at b = a + (fx dt) slope with the step counted in float32,
  // then the position and counter updates (such iterations take t from the step count once
  // they end, not per step). The bounds are wide enough that no walk ends, so every lane
  // steps every time (the simulator's walks end and are refilled).
#pragma clang fp contract(off)
  const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  float y = 0.1f * (float)(gid & 7), fx = 0.f;
  const float dt = 0.001f, vdt = 0.0005f, sst = 0.0316f, a = 3.0e4f, slope = -0.2f;
  // native walks draw normals scaled by sqrt_st (Box–Muller radius sqrt(-2 ln u1 sqrt_st^2))
  const float bmc = kBoxMullerLog2 * (sst * sst);
  bool alive = true;
  for (int64_t it = 0; it < iters; ++it) {
    if (alive) {
      float nz[4];
      philox_normals4_c((uint32_t)gid, 7u, (uint32_t)it, 0u, 0x1234u, 0x5678u, bmc, nz);
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        if (alive) {
          const float b = a + (fx * dt) * slope;
          if (fabsf(y) <= b) {
            asm volatile("");  // a branch, as in the simulator (not speculated into selects)
            y += vdt + nz[i];
            fx += 1.0f;
          } else {
            alive = false;
          }
        }
      }
    }
  }
  out[gid] = y + fx;
}

template <bool LCA>
__global__ void __launch_bounds__(256) probe_acc_step(int64_t iters, float* out) {
  // one accumulator-model step per iteration, as the simulator's native accumulator loop does
  // it (bound test on every accumulator, four normals from one Philox block, the update; LCA:
  // in its untested iterations, with the update's products fused), without walk termination
  // or refill: a finished walk restarts in place
  const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const float dt = 0.001f, a = 2.0f, leak_dt = 0.2f * dt, inh_dt = 0.1f * dt, alpha = 2.0f, binv = 1.0f / 3.0f;
  float x[4], vraw[4], vdt[4], sst[4];
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    x[j] = 0.05f * (float)((gid + j) & 7);
    vraw[j] = 0.4f + 0.1f * (float)j;
    vdt[j] = vraw[j] * dt;
    sst[j] = 0.0316f;
  }
  float t = 0.f;
  int ix = 0;
  for (int64_t it = 0; it < iters; ++it) {
    float b = a;
    if constexpr (!LCA) {
      const float u = ((float)ix * dt) * binv;
      const float p = __builtin_amdgcn_exp2f(alpha * __builtin_amdgcn_logf(u));
      b = a * __builtin_amdgcn_exp2f(-1.4426950408889634f * p);
    }
    bool go = LCA || (ix < 20001 && t <= 20.f);
#pragma unroll
    for (int j = 0; j < 4; ++j) go = go && !(x[j] > b);
    float nz[4];
    philox_normals4((uint32_t)gid, 7u, (uint32_t)ix, 0u, 0x1234u, 0x5678u, nz);
    float xn[4];
    if constexpr (LCA) {
      float sum = 0.f;
#pragma unroll
      for (int j = 0; j < 4; ++j) sum += x[j];
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const float red = (-x[j]) + sum;
        const float dr = __builtin_fmaf(-leak_dt, x[j], __builtin_fmaf(-inh_dt, red, vdt[j]));
        xn[j] = fmaxf(0.0f, x[j] + __builtin_fmaf(sst[j], nz[j], dr));
      }
    } else {
#pragma unroll
      for (int j = 0; j < 4; ++j) xn[j] = fmaxf(0.0f, x[j] + (vdt[j] + sst[j] * nz[j]));
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) x[j] = go ? xn[j] : 0.5f * x[j];
    t = go ? t + dt : 0.f;
    ix = go ? ix + 1 : 0;
  }
  out[gid] = x[0] + x[1] + x[2] + x[3] + t;
}

__global__ void __launch_bounds__(256) probe_kde_pair(int64_t iters, float* out) {
  // the evaluation kernel's inner step: 4 points x 4 samples per iteration, each pair
  // d = x - s, 2^(fma(-d, d, 64)), add; the 4 sample updates stand in for its LDS read
  const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  float x[4], acc[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    x[q] = 0.01f * (float)((gid + q) & 255);
    acc[q] = 0.f;
  }
  float s0 = 0.25f, s1 = 0.35f, s2 = 0.45f, s3 = 0.55f;
  for (int64_t it = 0; it < iters; ++it) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      float d;
      d = x[q] - s0; acc[q] += __builtin_amdgcn_exp2f(__builtin_fmaf(-d, d, 64.f));
      d = x[q] - s1; acc[q] += __builtin_amdgcn_exp2f(__builtin_fmaf(-d, d, 64.f));
      d = x[q] - s2; acc[q] += __builtin_amdgcn_exp2f(__builtin_fmaf(-d, d, 64.f));
      d = x[q] - s3; acc[q] += __builtin_amdgcn_exp2f(__builtin_fmaf(-d, d, 64.f));
    }
    s0 += 1e-6f; s1 += 1e-6f; s2 += 1e-6f; s3 += 1e-6f;
  }
  out[gid] = acc[0] + acc[1] + acc[2] + acc[3];
}

template <int KIND>
__global__ void __launch_bounds__(256) probe_single(int64_t iters, float* out) {
  const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  float f[8];
  uint32_t u[8];
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    f[j] = 1.0f + 0.01f * (float)((gid + j) & 15);
    u[j] = (uint32_t)(gid * 8 + j) | 1u;
  }
  for (int64_t it = 0; it < iters; ++it) {
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      if constexpr (KIND == 2) {
        const uint64_t p = (uint64_t)0xD2511F53u * (uint64_t)u[j];
        u[j] = (uint32_t)(p >> 32) ^ (uint32_t)p;
      } else if constexpr (KIND == 3) {
        u[j] = (u[j] ^ 0x9E3779B9u) ^ (u[j] << 1);
      } else if constexpr (KIND == 4) {
        f[j] = __builtin_fmaf(f[j], 0.999f, 0.001f);
      } else if constexpr (KIND == 5) {
        f[j] = __builtin_amdgcn_logf(f[j]) + 2.0f;
      } else if constexpr (KIND == 6) {
        f[j] = __builtin_amdgcn_exp2f(f[j]) * 0.25f;
      } else if constexpr (KIND == 7) {
        f[j] = __builtin_amdgcn_sinf(f[j]) + 0.5f;
      } else if constexpr (KIND == 9) {
        // the surrogate's hidden unit on a pre-scaled argument, u = 1 / (2^z + 1) =
        // (1 - tanh) / 2, with one fma standing in for the bf16 repack of the result
        const float z = __builtin_fmaf(f[j], 0.5f, 0.25f);
        f[j] = __builtin_amdgcn_rcpf(__builtin_amdgcn_exp2f(z) + 1.0f);
      } else {
        f[j] = __builtin_amdgcn_sqrtf(f[j]) + 0.5f;
      }
    }
  }
  float acc = 0.f;
#pragma unroll
  for (int j = 0; j < 8; ++j) acc += f[j] + (float)u[j];
  out[gid] = acc;
}

int32_t probe_enqueue(int32_t kind, int64_t iters, int64_t n_threads, float* out, hipStream_t stream) {
  if (iters < 1 || n_threads < 256 || n_threads % 256) return fail(SSM_EINVAL, "probe needs iters >= 1, threads % 256 == 0");
  if (!out) return fail(SSM_EINVAL, "probe output is NULL");
  const dim3 grid((unsigned)(n_threads / 256)), block(256);
  switch (kind) {
    case 0: hipLaunchKernelGGL(probe_em_block, grid, block, 0, stream, iters, out); break;
    case 1: hipLaunchKernelGGL(probe_kde_pair, grid, block, 0, stream, iters, out); break;
    case 2: hipLaunchKernelGGL(probe_single<2>, grid, block, 0, stream, iters, out); break;
    case 3: hipLaunchKernelGGL(probe_single<3>, grid, block, 0, stream, iters, out); break;
    case 4: hipLaunchKernelGGL(probe_single<4>, grid, block, 0, stream, iters, out); break;
    case 5: hipLaunchKernelGGL(probe_single<5>, grid, block, 0, stream, iters, out); break;
    case 6: hipLaunchKernelGGL(probe_single<6>, grid, block, 0, stream, iters, out); break;
    case 7: hipLaunchKernelGGL(probe_single<7>, grid, block, 0, stream, iters, out); break;
    case 8: hipLaunchKernelGGL(probe_single<8>, grid, block, 0, stream, iters, out); break;
    case 9: hipLaunchKernelGGL(probe_single<9>, grid, block, 0, stream, iters, out); break;
    case 10: hipLaunchKernelGGL(probe_acc_step<true>, grid, block, 0, stream, iters, out); break;
    case 11: hipLaunchKernelGGL(probe_acc_step<false>, grid, block, 0, stream, iters, out); break;
    default: return fail(SSM_EINVAL, "unknown probe kind");
  }
  SSM_HIP(hipGetLastError());
  return SSM_OK;
}

}  // namespace ssmhip
