// Per-set auxiliary labels of the LAN data generator and the uniform parameter sampler.
//
// aux_count_kernel restates SimulationPipeline._compute_auxiliary_labels
// (ssms/dataset_generators/pipelines/simulation_pipeline.py:258-351) as integer
// counts: draws per choice, the no-go count ((choice != max choice) or rt == -999),
// and per-choice RT histograms over np.linspace(0, max_t, B + 1) for B = 128 and 256
// with np.histogram's edge rules (bins closed on the left, the last bin closed on
// both sides, values outside [0, max_t] dropped; omitted rts excluded). The host
// turns counts into the float64 probabilities and float32 arrays the reference
// returns, so every label is exact.
//
// theta_uniform_kernel draws theta[k][j] = float32(lo_j + (hi_j - lo_j) u) with a
// 53-bit uniform u built from two Philox words (the construction of MT19937's
// genrand_res53), as UniformParameterSampler draws rng.uniform(lo, hi) in float64
// and casts to float32 (parameter_samplers/uniform_sampler.py:24-51).

#include <hip/hip_runtime.h>

#include <string>

#include "ssmhip_internal.h"

namespace ssmhip {

constexpr int kAuxThreads = 256;

struct AuxArgs {
  const float* rts;
  const void* choices;
  const int64_t* offsets;
  const int32_t* choice_values;
  int32_t* counts;   // [n_sets][C]
  int32_t* nogo;     // [n_sets]
  int32_t* hist128;  // [n_sets][128][C]
  int32_t* hist256;  // [n_sets][256][C]
  double max_t;
  int32_t n_choices;
  int32_t choice_bytes;
  int32_t max_choice;
};

__device__ __forceinline__ int aux_choice(const void* ch, int bytes, int64_t i) {
  return bytes == 1 ? (int)((const int8_t*)ch)[i] : ((const int32_t*)ch)[i];
}

// np.histogram bin of x for edges e_i = i * fl(max_t / B) (i < B), e_B = max_t, or -1
__device__ __forceinline__ int hist_bin(double x, double max_t, double step, int B) {
  if (!(x >= 0.0) || x > max_t) return -1;
  int i = (int)(x / step);
  if (i > B - 1) i = B - 1;
  if (i < 0) i = 0;
  // settle on the exact edges (as numpy's decrement / increment fix-up)
  if (x < (double)i * step) --i;
  if (i < B - 1 && x >= ((i + 1 == B) ? max_t : (double)(i + 1) * step)) ++i;
  return i;
}

__global__ void __launch_bounds__(kAuxThreads) aux_count_kernel(AuxArgs A) {
  extern __shared__ int lds[];  // counts[C], nogo, hist128[128*C], hist256[256*C]
  const int C = A.n_choices;
  int* cnt = lds;
  int* nogo = lds + C;
  int* h128 = nogo + 1;
  int* h256 = h128 + 128 * C;
  const int total = C + 1 + 384 * C;
  for (int i = threadIdx.x; i < total; i += kAuxThreads) lds[i] = 0;
  __syncthreads();
  const int set = blockIdx.x;
  const int64_t lo = A.offsets[set], hi = A.offsets[set + 1];
  const double step128 = A.max_t / 128.0, step256 = A.max_t / 256.0;
  for (int64_t i = lo + threadIdx.x; i < hi; i += kAuxThreads) {
    const int ch = aux_choice(A.choices, A.choice_bytes, i);
    const float r = A.rts[i];
    if (ch != A.max_choice || r == -999.f) atomicAdd(nogo, 1);
    int k = -1;
    for (int c = 0; c < C; ++c)
      if (k < 0 && ch == A.choice_values[c]) k = c;
    if (k < 0) continue;
    atomicAdd(&cnt[k], 1);
    if (r == -999.f) continue;
    const double x = (double)r;
    const int b1 = hist_bin(x, A.max_t, step128, 128);
    if (b1 >= 0) atomicAdd(&h128[b1 * C + k], 1);
    const int b2 = hist_bin(x, A.max_t, step256, 256);
    if (b2 >= 0) atomicAdd(&h256[b2 * C + k], 1);
  }
  __syncthreads();
  for (int i = threadIdx.x; i < C; i += kAuxThreads) A.counts[(int64_t)set * C + i] = cnt[i];
  if (threadIdx.x == 0) A.nogo[set] = nogo[0];
  for (int i = threadIdx.x; i < 128 * C; i += kAuxThreads) A.hist128[(int64_t)set * 128 * C + i] = h128[i];
  for (int i = threadIdx.x; i < 256 * C; i += kAuxThreads) A.hist256[(int64_t)set * 256 * C + i] = h256[i];
}

__global__ void theta_uniform_kernel(int32_t n_sets, int32_t n_params, const double* lo, const double* hi,
                                     uint32_t k0, uint32_t k1, float* theta) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (int64_t)n_sets * n_params) return;
  const int set = (int)(e / n_params), j = (int)(e % n_params);
  const u32x4 r = philox4x32_10(u32x4{(uint32_t)set, (uint32_t)j, 0u, 21u}, k0, k1);
  const double u = ((double)(r.x >> 5) * 67108864.0 + (double)(r.y >> 6)) * (1.0 / 9007199254740992.0);
  theta[e] = (float)(lo[j] + (hi[j] - lo[j]) * u);
}

int32_t aux_labels_enqueue(int32_t n_sets, int32_t n_choices, const int32_t* choice_values_dev,
                           int32_t max_choice, const float* rts, const void* choices, int32_t choice_bytes,
                           const int64_t* offsets, double max_t, int32_t* counts, int32_t* nogo, int32_t* hist128,
                           int32_t* hist256, hipStream_t s) {
  if (n_sets < 0) return fail(SSM_EINVAL, "n_sets must be >= 0");
  if (n_choices < 1 || n_choices > SSM_MAX_PARTICLES) return fail(SSM_EINVAL, "n_choices must be in [1, 16]");
  if (choice_bytes != 1 && choice_bytes != 4) return fail(SSM_EINVAL, "choice_bytes must be 1 or 4");
  if (!(max_t > 0.0)) return fail(SSM_EINVAL, "max_t must be positive");
  if (n_sets == 0) return SSM_OK;
  if (!choice_values_dev || !rts || !choices || !offsets || !counts || !nogo || !hist128 || !hist256)
    return fail(SSM_EINVAL, "NULL buffer");
  AuxArgs A{rts, choices, offsets, choice_values_dev, counts, nogo, hist128, hist256, max_t, n_choices, choice_bytes,
            max_choice};
  const size_t lds = (size_t)(n_choices + 1 + 384 * n_choices) * sizeof(int);
  hipLaunchKernelGGL(aux_count_kernel, dim3((unsigned)n_sets), dim3(kAuxThreads), lds, s, A);
  SSM_HIP(hipGetLastError());
  return SSM_OK;
}

int32_t theta_uniform_enqueue(int32_t n_sets, int32_t n_params, const double* lo, const double* hi, uint64_t seed,
                              float* theta, hipStream_t s) {
  if (n_sets < 0 || n_params < 1 || n_params > 64) return fail(SSM_EINVAL, "need n_sets >= 0, 1 <= n_params <= 64");
  if (n_sets == 0) return SSM_OK;
  if (!lo || !hi || !theta) return fail(SSM_EINVAL, "NULL buffer");
  const int64_t n = (int64_t)n_sets * n_params;
  hipLaunchKernelGGL(theta_uniform_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n_sets, n_params, lo,
                     hi, (uint32_t)seed, (uint32_t)(seed >> 32), theta);
  SSM_HIP(hipGetLastError());
  return SSM_OK;
}

// ---- row copies between set-major buffers (the simulation filter's replacements) -----
// dst row dst_idx[r] (or r) <- src row src_idx[r] (or r), rows of row_bytes bytes; one block
// per row, 4-byte words when every row starts 4-byte aligned.
__global__ void copy_rows_kernel(int64_t n_rows, int64_t row_bytes, const uint8_t* __restrict__ src,
                                 const int64_t* __restrict__ src_idx, uint8_t* __restrict__ dst,
                                 const int64_t* __restrict__ dst_idx, int words) {
  for (int64_t r = blockIdx.x; r < n_rows; r += gridDim.x) {
    const int64_t s = src_idx ? src_idx[r] : r;
    const int64_t d = dst_idx ? dst_idx[r] : r;
    const uint8_t* a = src + s * row_bytes;
    uint8_t* b = dst + d * row_bytes;
    if (words) {
      const int64_t nw = row_bytes >> 2;
      for (int64_t i = threadIdx.x; i < nw; i += blockDim.x)
        reinterpret_cast<uint32_t*>(b)[i] = reinterpret_cast<const uint32_t*>(a)[i];
    } else {
      for (int64_t i = threadIdx.x; i < row_bytes; i += blockDim.x) b[i] = a[i];
    }
  }
}

int32_t copy_rows_enqueue(int64_t n_rows, int64_t row_bytes, const void* src, const int64_t* src_idx, void* dst,
                          const int64_t* dst_idx, hipStream_t stream) {
  if (n_rows < 0 || row_bytes < 0) return fail(SSM_EINVAL, "negative sizes");
  if (n_rows == 0 || row_bytes == 0) return SSM_OK;
  if (!src || !dst) return fail(SSM_EINVAL, "NULL buffer");
  const int words = (row_bytes % 4 == 0) && ((uintptr_t)src % 4 == 0) && ((uintptr_t)dst % 4 == 0);
  const int64_t grid = n_rows < 65536 ? n_rows : 65536;
  hipLaunchKernelGGL(copy_rows_kernel, dim3((unsigned)grid), dim3(256), 0, stream, n_rows, row_bytes,
                     (const uint8_t*)src, src_idx, (uint8_t*)dst, dst_idx, words);
  SSM_HIP(hipGetLastError());
  return SSM_OK;
}

// ---- per-set records laid out on the device from per-slot columns ----------------------
// record k (stride floats): slot j = columns[col_of[j] * n_sets + k] when col_of[j] >= 0, else 0;
// one thread per record float, so the record writes are coalesced.
struct SlotColumns {
  int8_t col_of[SSM_SET_STRIDE];
};
__global__ void em_records_kernel(int64_t n_sets, int stride, SlotColumns map, const float* __restrict__ columns,
                                  float* __restrict__ records) {
  const int64_t n = n_sets * stride;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t k = e / stride;
    const int j = (int)(e - k * stride);
    const int c = map.col_of[j];
    records[e] = c >= 0 ? columns[(int64_t)c * n_sets + k] : 0.f;
  }
}

int32_t em_records_enqueue(int32_t n_sets, int32_t set_stride, int32_t n_slots, const int32_t* slots,
                           const float* columns, float* records, hipStream_t stream) {
  if (n_sets < 0) return fail(SSM_EINVAL, "n_sets must be >= 0");
  if (set_stride < 1 || set_stride > SSM_SET_STRIDE)
    return fail(SSM_EINVAL, "set_stride must be in [1, " + std::to_string(SSM_SET_STRIDE) + "]");
  if (n_slots < 0 || n_slots > set_stride) return fail(SSM_EINVAL, "n_slots must be in [0, set_stride]");
  if (n_sets == 0) return SSM_OK;
  if (!records || (n_slots > 0 && (!slots || !columns))) return fail(SSM_EINVAL, "NULL buffer");
  SlotColumns map;
  for (int j = 0; j < SSM_SET_STRIDE; ++j) map.col_of[j] = -1;
  for (int c = 0; c < n_slots; ++c) {
    const int j = slots[c];
    if (j < 0 || j >= set_stride) return fail(SSM_EINVAL, "slot index outside the record");
    if (map.col_of[j] >= 0) return fail(SSM_EINVAL, "slot index given twice");
    map.col_of[j] = (int8_t)c;
  }
  const int64_t n = (int64_t)n_sets * set_stride;
  const int64_t blocks = (n + 255) / 256 < 65536 ? (n + 255) / 256 : 65536;
  hipLaunchKernelGGL(em_records_kernel, dim3((unsigned)blocks), dim3(256), 0, stream, (int64_t)n_sets, set_stride,
                     map, columns, records);
  SSM_HIP(hipGetLastError());
  return SSM_OK;
}

// offsets[k] = k * n_samples, k <= n_sets
__global__ void uniform_offsets_kernel(int64_t n_sets, int64_t n_samples, int64_t* __restrict__ offsets) {
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k <= n_sets; k += (int64_t)gridDim.x * blockDim.x)
    offsets[k] = k * n_samples;
}

int32_t uniform_offsets_enqueue(int32_t n_sets, int64_t n_samples, int64_t* offsets, hipStream_t stream) {
  if (n_sets < 0 || n_samples < 0) return fail(SSM_EINVAL, "n_sets and n_samples must be >= 0");
  if (!offsets) return fail(SSM_EINVAL, "NULL buffer");
  const int64_t n = (int64_t)n_sets + 1;
  const int64_t blocks = (n + 255) / 256 < 65536 ? (n + 255) / 256 : 65536;
  hipLaunchKernelGGL(uniform_offsets_kernel, dim3((unsigned)blocks), dim3(256), 0, stream, (int64_t)n_sets, n_samples,
                     offsets);
  SSM_HIP(hipGetLastError());
  return SSM_OK;
}

}  // namespace ssmhip
