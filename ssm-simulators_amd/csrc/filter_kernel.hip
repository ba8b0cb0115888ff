// Per-set simulation statistics behind the LAN data generator's simulation filter.
//
// SimulationPipeline._is_valid_simulation (ssms/dataset_generators/pipelines/
// simulation_pipeline.py:199-256) rejects a parameter set from, for every possible
// choice, the rts of that choice without omissions (rt != -999): their count, mean,
// standard deviation, mode (scipy.stats.mode: the smallest of the most frequent
// values) and the mode's share of the count (0 below 5 values). Here each set's
// samples are sorted once by (choice index, rt) with rocPRIM's segmented radix sort
// over 64-bit keys, and one block per set reads every statistic from the sorted run:
// the choice ranges by binary search, sums in double, the mode as the longest run of
// equal keys (the first such run on ties, i.e. the smallest value). Count and mode
// are exact; mean and std are formed in double (NumPy reduces float32 arrays in
// float32), which can only matter for a threshold met to the last float bit.

#include <hip/hip_runtime.h>
#include <rocprim/device/device_segmented_radix_sort.hpp>

#include "ssmhip_internal.h"

namespace ssmhip {

constexpr int kFilterThreads = 256;

__device__ __forceinline__ uint32_t float_order_bits(float f) {
  const uint32_t u = __float_as_uint(f);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);  // unsigned order == float order
}
__device__ __forceinline__ float float_from_order_bits(uint32_t u) {
  return __uint_as_float((u & 0x80000000u) ? (u & 0x7fffffffu) : ~u);
}

__global__ void filter_keys_kernel(const float* __restrict__ rts, const void* __restrict__ choices, int choice_bytes,
                                   const int32_t* __restrict__ choice_values, int C, int64_t n_total,
                                   uint64_t* __restrict__ keys) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_total) return;
  const int ch = choice_bytes == 1 ? (int)((const int8_t*)choices)[i] : ((const int32_t*)choices)[i];
  const float r = rts[i];
  int c = C;  // not counted: omitted rt or a value outside the possible choices
  if (r != -999.f)
    for (int j = 0; j < C; ++j)
      if (c == C && ch == choice_values[j]) c = j;
  keys[i] = ((uint64_t)c << 32) | (c < C ? float_order_bits(r) : 0u);
}

// first index in [lo, hi) whose key's choice field is >= c
__device__ __forceinline__ int64_t choice_lower_bound(const uint64_t* __restrict__ k, int64_t lo, int64_t hi, int c) {
  while (lo < hi) {
    const int64_t mid = (lo + hi) >> 1;
    if ((int)(k[mid] >> 32) < c) lo = mid + 1; else hi = mid;
  }
  return lo;
}

template <typename T, typename Op>
__device__ T block_reduce(T v, T* scratch, Op op) {
  for (int off = 32; off > 0; off >>= 1) v = op(v, __shfl_xor(v, off));
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  __syncthreads();
  if (lane == 0) scratch[wave] = v;
  __syncthreads();
  T r = scratch[0];
  for (int w = 1; w < kFilterThreads / 64; ++w) r = op(r, scratch[w]);
  return r;
}

__global__ void __launch_bounds__(kFilterThreads) filter_stats_kernel(const uint64_t* __restrict__ keys,
                                                                      const int64_t* __restrict__ offsets, int C,
                                                                      double* __restrict__ stats) {
  __shared__ double dscr[kFilterThreads / 64];
  __shared__ unsigned long long uscr[kFilterThreads / 64];
  const int set = blockIdx.x;
  const int64_t lo = offsets[set], hi = offsets[set + 1];
  for (int c = 0; c < C; ++c) {
    const int64_t b = choice_lower_bound(keys, lo, hi, c);
    const int64_t e = choice_lower_bound(keys, b, hi, c + 1);
    const int64_t n = e - b;
    double* out = stats + ((int64_t)set * C + c) * 5;
    if (n == 0) {  // simulation_pipeline.py:231-236
      if (threadIdx.x == 0) { out[0] = -1.0; out[1] = -1.0; out[2] = 1.0; out[3] = 0.0; out[4] = 0.0; }
      continue;
    }
    double s = 0.0;
    for (int64_t i = b + threadIdx.x; i < e; i += kFilterThreads)
      s += (double)float_from_order_bits((uint32_t)keys[i]);
    const double mean = block_reduce(s, dscr, [](double x, double y) { return x + y; }) / (double)n;
    double q = 0.0;
    unsigned long long best = 0;  // (run length << 32) | (0xffffffff - run start offset): max = longest, then first
    for (int64_t i = b + threadIdx.x; i < e; i += kFilterThreads) {
      const uint32_t u = (uint32_t)keys[i];
      const double d = (double)float_from_order_bits(u) - mean;
      q += d * d;
      if (i == b || (uint32_t)keys[i - 1] != u) {  // a run of equal values starts here
        int64_t j = i + 1;
        while (j < e && (uint32_t)keys[j] == u) ++j;
        const unsigned long long cand = ((unsigned long long)(j - i) << 32) | (0xffffffffull - (uint64_t)(i - b));
        if (cand > best) best = cand;
      }
    }
    const double var = block_reduce(q, dscr, [](double x, double y) { return x + y; }) / (double)n;
    best = block_reduce(best, uscr, [](unsigned long long x, unsigned long long y) { return x > y ? x : y; });
    if (threadIdx.x == 0) {
      const int64_t start = b + (int64_t)(0xffffffffull - (best & 0xffffffffull));
      const double cnt = (double)(best >> 32);
      out[0] = (double)float_from_order_bits((uint32_t)keys[start]);
      out[1] = mean;
      out[2] = sqrt(var);
      out[3] = n < 5 ? 0.0 : cnt / (double)n;
      out[4] = (double)n;
    }
    __syncthreads();
  }
}

static size_t sort_temp_bytes(int32_t n_sets, int64_t n_total, const int64_t* offsets) {
  size_t bytes = 0;
  rocprim::segmented_radix_sort_keys((void*)nullptr, bytes, (const uint64_t*)nullptr, (uint64_t*)nullptr,
                                     (unsigned int)n_total, (unsigned int)n_sets, offsets, offsets + 1, 0, 37);
  return bytes;
}

int32_t filter_workspace_bytes(int32_t n_sets, int64_t n_total, size_t* bytes) {
  if (!bytes) return fail(SSM_EINVAL, "bytes is NULL");
  if (n_sets < 0 || n_total < 0) return fail(SSM_EINVAL, "negative sizes");
  if (n_total > 0xffffffffll) return fail(SSM_EINVAL, "n_total must be < 2^32");
  const size_t keys = ((size_t)n_total * 8 + 255) & ~(size_t)255;
  *bytes = 2 * keys + sort_temp_bytes(n_sets, n_total, nullptr) + 256;
  return SSM_OK;
}

int32_t filter_stats_enqueue(int32_t n_sets, int32_t n_choices, const int32_t* choice_values, const float* rts,
                             const void* choices, int32_t choice_bytes, const int64_t* offsets, int64_t n_total,
                             void* ws, size_t ws_bytes, double* stats, hipStream_t stream) {
  if (n_sets < 0 || n_total < 0) return fail(SSM_EINVAL, "negative sizes");
  if (n_choices < 1 || n_choices > 16) return fail(SSM_EINVAL, "n_choices must be in [1, 16]");
  if (choice_bytes != 1 && choice_bytes != 4) return fail(SSM_EINVAL, "choice_bytes must be 1 or 4");
  if (n_total > 0xffffffffll) return fail(SSM_EINVAL, "n_total must be < 2^32");
  if (n_sets == 0) return SSM_OK;
  if (!choice_values || !offsets || !stats || (n_total > 0 && (!rts || !choices || !ws)))
    return fail(SSM_EINVAL, "NULL buffer");
  size_t need = 0;
  int32_t rc = filter_workspace_bytes(n_sets, n_total, &need);
  if (rc != SSM_OK) return rc;
  if (ws_bytes < need) return fail(SSM_EINVAL, "workspace too small (see ssm_filter_workspace_bytes)");
  const size_t kb = ((size_t)n_total * 8 + 255) & ~(size_t)255;
  uint64_t* keys_in = (uint64_t*)ws;
  uint64_t* keys_out = (uint64_t*)((char*)ws + kb);
  void* temp = (char*)ws + 2 * kb;
  size_t temp_bytes = need - 2 * kb;
  if (n_total > 0) {
    const int threads = 256;
    const int64_t blocks = (n_total + threads - 1) / threads;
    hipLaunchKernelGGL(filter_keys_kernel, dim3((unsigned)blocks), dim3(threads), 0, stream, rts, choices, choice_bytes,
                       choice_values, n_choices, n_total, keys_in);
    SSM_HIP(hipGetLastError());
    // choice index in bits 32..36 (C <= 16 plus the excluded class), rt order in bits 0..31
    SSM_HIP(rocprim::segmented_radix_sort_keys(temp, temp_bytes, (const uint64_t*)keys_in, keys_out,
                                               (unsigned int)n_total, (unsigned int)n_sets, offsets, offsets + 1, 0,
                                               37, stream));
  }
  hipLaunchKernelGGL(filter_stats_kernel, dim3(n_sets), dim3(kFilterThreads), 0, stream, (const uint64_t*)keys_out,
                     offsets, n_choices, stats);
  SSM_HIP(hipGetLastError());
  return SSM_OK;
}

}  // namespace ssmhip
