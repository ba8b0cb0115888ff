// Closed-form LBA sampler (no time stepping), one sample per thread.
// Follows lba_vanilla / lba_angle, src/cssm/lba_models.pyx:101-115 and :212-224:
//   zs_j ~ U(0, z), vs_j = |N(v_j, sd_j)|, x_j = (a - zs_j) / vs_j
//   (angle: / (vs_j + tan(theta))), choice = argmin_j x_j, rt = min_j x_j + t,
//   rt >= deadline -> -999.
#include <hip/hip_runtime.h>

#include "ssmhip_internal.h"

namespace ssmhip {

struct LbaArgs {
  const float* params;
  const int64_t* offsets;
  const float* replay;
  float* rts;
  void* choices;
  int64_t total;
  int32_t n_sets;
  int32_t nact;
  int32_t out_layout;
  int32_t choice_bytes;
  uint32_t k0, k1;
};

__device__ __forceinline__ int lba_find_set(const int64_t* __restrict__ off, int n_sets, int64_t g) {
  int lo = 0, hi = n_sets - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (off[mid] <= g) lo = mid; else hi = mid - 1;
  }
  return lo;
}

__global__ void __launch_bounds__(256) lba_kernel(LbaArgs A) {
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= A.total) return;
  const int k = lba_find_set(A.offsets, A.n_sets, g);
  const int n = (int)(g - A.offsets[k]);
  const float* sp = A.params + (int64_t)k * SSM_SET_STRIDE;
  const float a = sp[SSM_P_A], zr = sp[SSM_P_Z], t0 = sp[SSM_P_T0], deadline = sp[SSM_P_DEADLINE];
  const float tan_th = sp[SSM_P_B1];
  float best = 0.f;
  int choice = 0;
  for (int q = 0; q * 4 < A.nact; ++q) {
    float u[4], nz[4];
    if (A.replay) {
      const float* r = A.replay + g * (2 * A.nact);
      for (int i = 0; i < 4; ++i) {
        const int j = 4 * q + i;
        u[i] = j < A.nact ? r[j] : 0.f;
        nz[i] = j < A.nact ? r[A.nact + j] : 0.f;
      }
    } else {
      const u32x4 ru = philox4x32_10(u32x4{(uint32_t)n, (uint32_t)k, (uint32_t)q, 3u}, A.k0, A.k1);
      u[0] = u32_to_unit(ru.x); u[1] = u32_to_unit(ru.y); u[2] = u32_to_unit(ru.z); u[3] = u32_to_unit(ru.w);
      philox_normals4((uint32_t)n, (uint32_t)k, (uint32_t)q, 5u, A.k0, A.k1, nz);
    }
    for (int i = 0; i < 4; ++i) {
      const int j = 4 * q + i;
      if (j >= A.nact) break;
      const float zs = u[i] * zr;
      const float vs = fabsf(sp[SSM_P_VJ + j] + sp[SSM_P_SSTJ + j] * nz[i]);
      const float xj = (a - zs) / (vs + tan_th);
      if (j == 0 || xj < best) { best = xj; choice = j; }
    }
  }
  float rt = best + t0;
  if (rt >= deadline) rt = -999.f;
  const int64_t o = A.out_layout == 0 ? g : (int64_t)n * A.n_sets + k;
  if (o >= A.total) return;
  A.rts[o] = rt;
  if (A.choice_bytes == 1) ((int8_t*)A.choices)[o] = (int8_t)choice;
  else ((int32_t*)A.choices)[o] = choice;
}

int32_t lba_enqueue(int32_t nact, int32_t n_sets, int64_t n_total, int32_t out_layout, int32_t choice_bytes,
                    uint64_t seed, const float* params, const int64_t* offsets, const float* replay,
                    float* rts, void* choices, hipStream_t stream) {
  if (nact < 1 || nact > SSM_MAX_PARTICLES) return fail(SSM_EINVAL, "nact must be in [1, 16]");
  if (n_sets < 0 || n_total < 0) return fail(SSM_EINVAL, "negative sizes");
  if (n_total > 0 && n_sets == 0) return fail(SSM_EINVAL, "n_total > 0 with no sets");
  if (out_layout != 0 && out_layout != 1) return fail(SSM_EINVAL, "out_layout must be 0 or 1");
  if (choice_bytes != 1 && choice_bytes != 4) return fail(SSM_EINVAL, "choice_bytes must be 1 or 4");
  if (n_total == 0) return SSM_OK;
  if (!params || !offsets || !rts || !choices) return fail(SSM_EINVAL, "NULL buffer");
  LbaArgs A{params, offsets, replay, rts, choices, n_total, n_sets, nact, out_layout, choice_bytes,
            (uint32_t)seed, (uint32_t)(seed >> 32)};
  const int64_t grid = (n_total + 255) / 256;
  hipLaunchKernelGGL(lba_kernel, dim3((unsigned)grid), dim3(256), 0, stream, A);
  SSM_HIP(hipGetLastError());
  return SSM_OK;
}

// ---- RNG diagnostics ------------------------------------------------------------
__global__ void philox_fill_kernel(uint32_t k0, uint32_t k1, uint32_t c0, uint32_t c1, uint32_t c2,
                                   uint32_t c3, int64_t n, uint32_t* raw, float* normals) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const u32x4 r = philox4x32_10(u32x4{c0 + (uint32_t)i, c1, c2, c3}, k0, k1);
  if (raw) {
    raw[4 * i + 0] = r.x; raw[4 * i + 1] = r.y; raw[4 * i + 2] = r.z; raw[4 * i + 3] = r.w;
  }
  if (normals) {
    float z[4];
    box_muller(r.x, r.y, z[0], z[1]);
    box_muller(r.z, r.w, z[2], z[3]);
    for (int j = 0; j < 4; ++j) normals[4 * i + j] = z[j];
  }
}

int32_t philox_enqueue(uint64_t seed, uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, int64_t n,
                       uint32_t* raw, float* normals, hipStream_t stream) {
  if (n < 0) return fail(SSM_EINVAL, "n must be >= 0");
  if (n == 0) return SSM_OK;
  const int64_t grid = (n + 255) / 256;
  hipLaunchKernelGGL(philox_fill_kernel, dim3((unsigned)grid), dim3(256), 0, stream, (uint32_t)seed,
                     (uint32_t)(seed >> 32), c0, c1, c2, c3, n, raw, normals);
  SSM_HIP(hipGetLastError());
  return SSM_OK;
}

}  // namespace ssmhip
