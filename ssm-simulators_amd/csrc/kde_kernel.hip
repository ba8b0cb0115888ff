// KDE likelihood labelling on gfx950: per-(set, choice) Gaussian KDE on log RT.
//
// Restates ssms/support_utils/kde_class.py:
//   fit     _attach_data_from_simulator :438-504, compute_bandwidths :89-126,
//           bandwidth_silverman :508-553 (h = (4/(3n))^(1/5) * max(std, 1e-3),
//           n == 1 -> std = 10, n == 0 -> no base data)
//   sample  kde_sample :331-434 + sklearn KernelDensity.sample
//           (i = floor(u * n), x = N(data[i], h), rt = exp(x))
//   eval    __kde_eval_ :251-329 + KernelDensity.score_samples:
//           l = log p_c + log sum_i exp(-(x - x_i)^2 / 2h^2) - log n_c - log h
//               - log(2 pi)/2 - x,  x = log(max(rt, eps)); l <= lb or rt <= 0 -> lb;
//           a choice without base data gets -log(N * max_t).
// and the LAN row mixture of
//   ssms/dataset_generators/strategies/resample_mixture_strategy.py:64-214.
//
// Evaluation is exp-throughput bound: each (point, sample) pair costs one
// sub, mul, fma, exp2 and add. Samples are staged through LDS in tiles and
// read back as wave-uniform broadcasts; every thread evaluates four points so
// one LDS read feeds four exp2. Sums carry a fixed 2^64 scale instead of a
// running maximum: terms that underflow (> 131 nats below the scale) cannot
// lift a log-likelihood above the floor lb = -66.774.
#include <hip/hip_runtime.h>
#include <math.h>

#include <mutex>
#include <vector>

#include "ssmhip_internal.h"

namespace ssmhip {

constexpr int kFitThreads = 1024;
constexpr float kOmission = -999.f;
constexpr float kScaleLog2 = 64.f;  // sums are scaled by 2^64

// ---- block-wide helpers (blockDim.x == NT, wave64) ------------------------------
template <int NT>
__device__ __forceinline__ float block_sum(float v, float* red) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  __syncthreads();
  if (l == 0) red[w] = v;
  __syncthreads();
  float s = 0.f;
  if (threadIdx.x == 0) {
    for (int i = 0; i < NT / 64; ++i) s += red[i];  // fixed order: deterministic
    red[NT / 64] = s;
  }
  __syncthreads();
  return red[NT / 64];
}

// exclusive prefix of integer v over the block; returns prefix, total in *tot
template <int NT>
__device__ __forceinline__ int block_scan_excl(int v, int* red, int* tot) {
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  int x = v;
  for (int o = 1; o < 64; o <<= 1) {
    const int y = __shfl_up(x, o);
    if (l >= o) x += y;
  }
  __syncthreads();
  if (l == 63) red[w] = x;
  __syncthreads();
  if (threadIdx.x == 0) {
    int acc = 0;
    for (int i = 0; i < NT / 64; ++i) {
      const int c = red[i];
      red[i] = acc;
      acc += c;
    }
    red[NT / 64] = acc;
  }
  __syncthreads();
  const int res = red[w] + x - v;
  *tot = red[NT / 64];
  return res;
}

__device__ __forceinline__ int load_choice(const void* ch, int bytes, int64_t i) {
  return bytes == 1 ? (int)((const int8_t*)ch)[i] : ((const int32_t*)ch)[i];
}

// ---- fit: split by choice, compact log RT, Silverman bandwidth ------------------
struct FitArgs {
  const float* rts;
  const void* choices;
  const int64_t* offsets;
  const int32_t* choice_values;
  float* samples;        // [total] compacted log rt, per set grouped by choice
  int32_t* kde_count;    // [n_sets*C] samples kept (rt != -999)
  int32_t* choice_total; // [n_sets*C] all draws with that choice (for p_c)
  float* bandwidth;      // [n_sets*C] h, 0 when no base data
  int64_t* kde_start;    // [n_sets*C] absolute index into samples
  int32_t n_choices;
  int32_t choice_bytes;
};

// One block per set, two passes over the set's draws, every load and store coalesced:
//  1. per choice: the draws with that choice, those not omitted, and the sum of their
//     log(rt), which gives the mean;
//  2. compact log(rt) per choice in sample order, 4096 draws per tile: slot j of the tile
//     holds draws j*1024 + thread, so the tile's kept draws are placed by wave ballots and a
//     prefix over (slot, wave), which one wave scans, behind a running cursor per choice
//     (two barriers per tile; the counts alternate between two buffers).
//     The same pass sums (x - mean)^2 per choice: the population variance of the two-pass
//     algorithm without reading the compacted samples back.
// Then the Silverman bandwidth. Sums are per thread, then over the block in a fixed order
// (deterministic). MC >= n_choices bounds the per-choice registers.
// draws per thread per tile (and per pass-1 iteration); kFitSlots x (kFitThreads / 64) = 64
// counts per choice and tile, one per lane of the wave that scans them
constexpr int kFitSlots = 4;
static_assert(kFitSlots * (kFitThreads / 64) == 64, "one count per lane of the scanning wave");

template <int MC>
__device__ __forceinline__ int fit_choice_index(int ch, const int (&cv)[MC]) {
  int k = -1;
#pragma unroll
  for (int c = MC - 1; c >= 0; --c)
    if (ch == cv[c]) k = c;  // first match (unused slots hold a value no choice takes)
  return k;
}

template <int MC>
__global__ void __launch_bounds__(kFitThreads) kde_fit_kernel(FitArgs A) {
  constexpr int NW = kFitThreads / 64;
  __shared__ float redf[NW + 1];
  __shared__ int redi[NW + 1];
  constexpr int S = kFitSlots;
  __shared__ int wave_cnt[2][S][NW][MC];
  __shared__ int run[MC];
  const int set = blockIdx.x;
  const int64_t lo = A.offsets[set], hi = A.offsets[set + 1];
  const int C = A.n_choices;
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  int cv[MC];
#pragma unroll
  for (int c = 0; c < MC; ++c) cv[c] = c < C ? A.choice_values[c] : 0x7fffffff;
  // pass 1
  int kept[MC], all[MC];
  float sx[MC];
#pragma unroll
  for (int c = 0; c < MC; ++c) {
    kept[c] = all[c] = 0;
    sx[c] = 0.f;
  }
  for (int64_t i0 = lo + threadIdx.x; i0 < hi; i0 += S * kFitThreads) {
    float r[S];
    int k[S];
#pragma unroll
    for (int j = 0; j < S; ++j) {
      const int64_t i = i0 + (int64_t)j * kFitThreads;
      r[j] = kOmission;
      k[j] = -1;
      if (i < hi) {
        r[j] = A.rts[i];
        k[j] = fit_choice_index<MC>(load_choice(A.choices, A.choice_bytes, i), cv);
      }
    }
#pragma unroll
    for (int j = 0; j < S; ++j) {
      const bool keep = r[j] != kOmission;
      const float x = keep ? logf(r[j]) : 0.f;
#pragma unroll
      for (int c = 0; c < MC; ++c) {
        all[c] += (k[j] == c) ? 1 : 0;
        kept[c] += (k[j] == c && keep) ? 1 : 0;
        sx[c] += (k[j] == c) ? x : 0.f;
      }
    }
  }
  int tot[MC], tot_all[MC], base[MC];
  float mean[MC];
  int acc = 0;
#pragma unroll
  for (int c = 0; c < MC; ++c) {
    tot[c] = tot_all[c] = 0;
    mean[c] = 0.f;
    if (c < C) {
      block_scan_excl<kFitThreads>(kept[c], redi, &tot[c]);
      block_scan_excl<kFitThreads>(all[c], redi, &tot_all[c]);
      mean[c] = block_sum<kFitThreads>(sx[c], redf) / (float)(tot[c] > 0 ? tot[c] : 1);
    }
    base[c] = acc;
    acc += tot[c];
  }
  // pass 2
  if (threadIdx.x < MC) run[threadIdx.x] = 0;
  __syncthreads();
  float q2[MC];
#pragma unroll
  for (int c = 0; c < MC; ++c) q2[c] = 0.f;
  const uint64_t below = (l == 0) ? 0ull : (~0ull >> (64 - l));
  constexpr int kTile = S * kFitThreads;
  const int64_t ntiles = (hi - lo + kTile - 1) / kTile;
  // up to 4 choices, the next tile's draws are loaded while the current one is placed (with
  // 16 the per-choice registers leave no room for them)
  constexpr bool kAhead = MC <= 4;
  float rn[S];
  int cn[S];
  auto load_tile = [&](int64_t tt) {
#pragma unroll
    for (int j = 0; j < S; ++j) {
      const int64_t i = lo + tt * kTile + (int64_t)j * kFitThreads + threadIdx.x;
      rn[j] = i < hi ? A.rts[i] : kOmission;
      cn[j] = i < hi ? load_choice(A.choices, A.choice_bytes, i) : 0;
    }
  };
  if (kAhead) load_tile(0);
  for (int64_t t = 0; t < ntiles; ++t) {
    const int par = (int)(t & 1);
    if (!kAhead) load_tile(t);
    float rc[S];
    int cc[S];
#pragma unroll
    for (int j = 0; j < S; ++j) {
      rc[j] = rn[j];
      cc[j] = cn[j];
    }
    if (kAhead && t + 1 < ntiles) load_tile(t + 1);
    float x[S];
    int k[S], inwave[S];
#pragma unroll
    for (int j = 0; j < S; ++j) {
      k[j] = -1;
      x[j] = 0.f;
      if (rc[j] != kOmission) {  // also the padding past the set's end
        k[j] = fit_choice_index<MC>(cc[j], cv);
        x[j] = logf(rc[j]);
      }
      inwave[j] = 0;
#pragma unroll
      for (int c = 0; c < MC; ++c) {
        if (c < C) {
          const uint64_t m = __ballot(k[j] == c);
          if (k[j] == c) inwave[j] = __popcll(m & below);
          if (l == 0) wave_cnt[par][j][w][c] = __popcll(m);
        }
      }
    }
    __syncthreads();
    // wave 0 turns each choice's 64 (slot, wave) counts into the output offsets of those
    // groups: the running cursor plus an exclusive prefix in (slot, wave) order
    if (w == 0) {
      for (int c = 0; c < C; ++c) {
        int* cnt = &wave_cnt[par][l / NW][l % NW][c];
        const int v = *cnt;
        int incl = v;
        for (int o = 1; o < 64; o <<= 1) {
          const int y = __shfl_up(incl, o);
          if (l >= o) incl += y;
        }
        const int r0 = run[c];
        *cnt = r0 + incl - v;
        if (l == 63) run[c] = r0 + incl;
      }
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < S; ++j) {
      if (k[j] >= 0) {
        const int pos = wave_cnt[par][j][w][k[j]] + inwave[j];
        int b = 0;
        float mu = 0.f;
#pragma unroll
        for (int c = 0; c < MC; ++c) {
          b = (k[j] == c) ? base[c] : b;
          mu = (k[j] == c) ? mean[c] : mu;
        }
        A.samples[lo + b + pos] = x[j];
        const float d = x[j] - mu;
#pragma unroll
        for (int c = 0; c < MC; ++c) q2[c] += (k[j] == c) ? d * d : 0.f;
      }
    }
  }
  // bandwidths
#pragma unroll
  for (int c = 0; c < MC; ++c) {
    if (c < C) {
      const float ss = block_sum<kFitThreads>(q2[c], redf);
      if (threadIdx.x == 0) {
        const int tc = tot[c];
        const int g = set * C + c;
        float h = 0.f;
        if (tc > 1) {
          float sd = sqrtf(ss / (float)tc);  // np.std: population
          if (sd < 1e-3f) sd = 1e-3f;
          h = powf(4.0f / (3.0f * (float)tc), 0.2f) * sd;
        } else if (tc == 1) {
          h = powf(4.0f / 3.0f, 0.2f) * 10.0f;
        }
        A.kde_count[g] = tc;
        A.choice_total[g] = tot_all[c];
        A.bandwidth[g] = h;
        A.kde_start[g] = lo + base[c];
      }
    }
  }
}

// ---- sample: KDE resamples + uniform points, grouped by choice ------------------
struct SampleArgs {
  const float* samples;
  const int32_t* kde_count;
  const int32_t* choice_total;
  const float* bandwidth;
  const int64_t* kde_start;
  const int64_t* offsets;
  float* pts_rt;        // [n_sets*m]
  int8_t* pts_choice;   // [n_sets*m] choice index
  int64_t* grp_begin;   // [n_sets*C]
  int32_t* grp_count;   // [n_sets*C]
  const int32_t* n_by_override;  // [n_sets*C] per-choice KDE sample counts, or NULL
  int32_t n_choices;
  int32_t n_kde;
  int32_t n_unif;
  float max_t;
  float unif_lo;
  uint32_t k0, k1;
};

constexpr int kSampleThreads = 1024;

__global__ void __launch_bounds__(kSampleThreads) kde_sample_kernel(SampleArgs A) {
  __shared__ int n_by[SSM_MAX_PARTICLES];
  __shared__ int redi[kSampleThreads / 64 + 1];
  const int set = blockIdx.x;
  const int C = A.n_choices;
  const int m = A.n_kde + A.n_unif;
  const int64_t N = A.offsets[set + 1] - A.offsets[set];
  if (threadIdx.x == 0 && A.n_by_override) {
    for (int c = 0; c < C; ++c) n_by[c] = A.n_by_override[set * C + c];
  } else if (threadIdx.x == 0) {
    // n_c = round(n_kde * p_c) as kde_class.py:387-392 forms it: p_c = count / N is a Python
    // float (kde_class.py:496), the product a float64, and round() goes half to even on that
    // float64 value. Two scalar float64 operations per choice reproduce its rounding exactly
    // (an exact-rational rule differs on ties where the float64 product is not exactly .5).
    int sum = 0, arg = 0;
    for (int c = 0; c < C; ++c) {
      const double p = N > 0 ? (double)A.choice_total[set * C + c] / (double)N : 0.0;
      const int64_t f = (int64_t)rint((double)A.n_kde * p);
      n_by[c] = (int)f;
      sum += (int)f;
      if (n_by[c] > n_by[arg]) arg = c;  // np.argmax: first maximum
    }
    if (sum > A.n_kde) n_by[arg] -= 1;
    else if (sum < A.n_kde) n_by[arg] += 1;
  }
  __syncthreads();
  // one point per thread (m <= 1024): draw it, then a stable counting sort by choice
  const int j = threadIdx.x;
  int c = -1;
  float rt = 0.f;
  if (j < A.n_kde) {
    int acc = 0;
    for (int q = 0; q < C; ++q) {
      if (c < 0 && j < acc + n_by[q]) c = q;
      acc += n_by[q];
    }
    if (c < 0) c = C - 1;
    const int g = set * C + c;
    const u32x4 r = philox4x32_10(u32x4{(uint32_t)j, (uint32_t)set, 0u, 11u}, A.k0, A.k1);
    const int nc = A.kde_count[g];
    if (nc > 0 && A.bandwidth[g] > 0.f) {
      // KernelDensity.sample: i = floor(u * n), x = N(data[i], h); rt = exp(x)
      int i = (int)(u32_to_unit(r.x) * (float)nc);
      if (i >= nc) i = nc - 1;
      float z0, z1;
      box_muller(r.y, r.z, z0, z1);
      rt = expf(A.samples[A.kde_start[g] + i] + A.bandwidth[g] * z0);
    } else {
      rt = u32_to_unit(r.x) * A.max_t;  // no base data: U(0, max_t) (kde_class.py:417-420)
    }
  } else if (j < m) {
    // uniform rows: rt ~ U(unif_lo, max_t), choice uniform over the possible choices
    const u32x4 r = philox4x32_10(u32x4{(uint32_t)j, (uint32_t)set, 0u, 12u}, A.k0, A.k1);
    c = (int)(u32_to_unit(r.x) * (float)C);
    if (c >= C) c = C - 1;
    rt = A.unif_lo + u32_to_unit(r.y) * (A.max_t - A.unif_lo);
  }
  int running = 0;
  for (int q = 0; q < C; ++q) {
    int tot = 0;
    const int pos = block_scan_excl<kSampleThreads>(c == q ? 1 : 0, redi, &tot);
    if (c == q) {
      const int64_t o = (int64_t)set * m + running + pos;
      A.pts_rt[o] = rt;
      A.pts_choice[o] = (int8_t)q;
    }
    if (threadIdx.x == 0) {
      A.grp_begin[set * C + q] = (int64_t)set * m + running;
      A.grp_count[set * C + q] = tot;
    }
    running += tot;
  }
}

// ---- eval ----------------------------------------------------------------------------
struct EvalArgs {
  const float* pts_rt;
  const int64_t* grp_begin;
  const int32_t* grp_count;
  const float* samples;
  const int32_t* kde_count;
  const int32_t* choice_total;
  const float* bandwidth;
  const int64_t* kde_start;
  const int64_t* offsets;
  float* partial;  // [n_points][n_chunks] scaled partial sums
  float* out_ll;
  int32_t n_choices;
  int32_t n_chunks;
  float max_t;
  float lb;
  float eps;
};

constexpr int kEvalThreads = 256;
constexpr int kPtsPerThread = 4;
constexpr int kTile = 2048;

// Pass 1: block (group g, chunk c) sums the kernel over samples
// [c * SSM_KDE_CHUNK, (c + 1) * SSM_KDE_CHUNK) of group g for every point of g.
// Splitting groups along their samples balances the very unequal
// (points x samples) work of the groups across the chip. Inside a block each of the four
// waves takes every point of the group (up to kRows rows of 64 per pass) against its own
// quarter of each sample tile, so all waves carry the same number of point rows and none
// waits at the tile barriers for another; only the last row of a group has idle lanes.
// The four waves' sums are added in LDS in a fixed order.
constexpr int kWaves = kEvalThreads / 64;
constexpr int kRows = 2 * kPtsPerThread;  // point rows per wave per pass (512 points)

// NQ of a group of kPtsPerThread point rows against the samples [lo, hi) of the LDS tile,
// four samples per step in tile order (the per-point summation order does not depend on NQ)
template <int NQ>
__device__ __forceinline__ void kde_pairs(const float* tile, int lo, int hi, const float (&x)[kPtsPerThread],
                                          float (&acc)[kPtsPerThread]) {
  for (int i = lo; i < hi; i += 4) {
    const float4 s4 = *reinterpret_cast<const float4*>(&tile[i]);
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      float d;
      d = x[q] - s4.x; acc[q] += __builtin_amdgcn_exp2f(__builtin_fmaf(-d, d, kScaleLog2));
      d = x[q] - s4.y; acc[q] += __builtin_amdgcn_exp2f(__builtin_fmaf(-d, d, kScaleLog2));
      d = x[q] - s4.z; acc[q] += __builtin_amdgcn_exp2f(__builtin_fmaf(-d, d, kScaleLog2));
      d = x[q] - s4.w; acc[q] += __builtin_amdgcn_exp2f(__builtin_fmaf(-d, d, kScaleLog2));
    }
  }
}

__global__ void __launch_bounds__(kEvalThreads) kde_eval_partial_kernel(EvalArgs A) {
  __shared__ __attribute__((aligned(16))) float tile[kTile];
  __shared__ float red[(kWaves - 1) * kRows * 64];
  const int g = blockIdx.x;
  const int chunk = blockIdx.y;
  const int cnt = A.grp_count[g];
  const int nc = A.kde_count[g];
  const float h = A.bandwidth[g];
  const int s0 = chunk * SSM_KDE_CHUNK;
  if (cnt <= 0 || nc == 0 || !(h > 0.f) || s0 >= nc) return;
  const int s1 = (s0 + SSM_KDE_CHUNK < nc) ? s0 + SSM_KDE_CHUNK : nc;
  const int64_t pb = A.grp_begin[g];
  const float* smp = A.samples + A.kde_start[g];
  // Points and samples enter centred on the group's first sample and scaled by
  // sqrt(log2(e) / 2) / h, so a pair costs d = x - s, e = 64 - d*d (one fma), 2^e, add:
  // 2^(64 - log2(e) (x - s)^2 / (2 h^2)) = 2^64 exp(-(x - s)^2 / (2 h^2)).
  const float sc = 0.84932180028801904f / h;
  const float ctr = smp[0];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  constexpr int seg = kTile / kWaves;  // samples of each tile per wave (a multiple of 4)
  for (int pbase = 0; pbase < cnt; pbase += kRows * 64) {
    // rows of 64 points this pass: the same for every wave of the block
    const int rows = min(kRows, (cnt - pbase + 63) >> 6);
    float x[kRows / kPtsPerThread][kPtsPerThread], acc[kRows / kPtsPerThread][kPtsPerThread];
#pragma unroll
    for (int r4 = 0; r4 < kRows / kPtsPerThread; ++r4)
#pragma unroll
      for (int q = 0; q < kPtsPerThread; ++q) {
        const int pi = pbase + (r4 * kPtsPerThread + q) * 64 + lane;
        const float rt = pi < cnt ? A.pts_rt[pb + pi] : 1.f;
        x[r4][q] = (logf(fmaxf(rt, A.eps)) - ctr) * sc;
        acc[r4][q] = 0.f;
      }
    for (int t0 = s0; t0 < s1; t0 += kTile) {
      const int tn = (s1 - t0) < kTile ? (s1 - t0) : kTile;
      __syncthreads();
      for (int i = threadIdx.x; i < kTile; i += kEvalThreads) tile[i] = i < tn ? (smp[t0 + i] - ctr) * sc : 1e30f;
      __syncthreads();
      const int tn4 = (tn + 3) & ~3;
      const int lo = wave * seg;
      const int hi = lo + seg < tn4 ? lo + seg : tn4;
#pragma unroll
      for (int r4 = 0; r4 < kRows / kPtsPerThread; ++r4) {
        const int nq = min(kPtsPerThread, rows - r4 * kPtsPerThread);  // wave-uniform
        static_assert(kPtsPerThread == 4, "kde_pairs dispatch below covers 1..4 rows");
        switch (nq) {
          case 4: kde_pairs<4>(tile, lo, hi, x[r4], acc[r4]); break;
          case 3: kde_pairs<3>(tile, lo, hi, x[r4], acc[r4]); break;
          case 2: kde_pairs<2>(tile, lo, hi, x[r4], acc[r4]); break;
          case 1: kde_pairs<1>(tile, lo, hi, x[r4], acc[r4]); break;
          default: break;
        }
      }
    }
    // waves 1..3 hand their sums to wave 0, which adds them in wave order
    __syncthreads();
    if (wave > 0) {
#pragma unroll
      for (int r4 = 0; r4 < kRows / kPtsPerThread; ++r4)
#pragma unroll
        for (int q = 0; q < kPtsPerThread; ++q)
          red[((wave - 1) * kRows + r4 * kPtsPerThread + q) * 64 + lane] = acc[r4][q];
    }
    __syncthreads();
    if (wave == 0) {
#pragma unroll
      for (int r4 = 0; r4 < kRows / kPtsPerThread; ++r4)
#pragma unroll
        for (int q = 0; q < kPtsPerThread; ++q) {
          float v = acc[r4][q];
          for (int m = 1; m < kWaves; ++m) v += red[((m - 1) * kRows + r4 * kPtsPerThread + q) * 64 + lane];
          const int pi = pbase + (r4 * kPtsPerThread + q) * 64 + lane;
          if (pi < cnt) A.partial[(pb + pi) * A.n_chunks + chunk] = v;
        }
    }
  }
}

// Pass 2: per group, sum the chunks in a fixed order (deterministic) and finish l.
__global__ void __launch_bounds__(kEvalThreads) kde_eval_finish_kernel(EvalArgs A) {
  const int g = blockIdx.x;
  const int set = g / A.n_choices;
  const int cnt = A.grp_count[g];
  if (cnt <= 0) return;
  const int64_t pb = A.grp_begin[g];
  const int64_t N = A.offsets[set + 1] - A.offsets[set];
  const int nc = A.kde_count[g];
  const float h = A.bandwidth[g];
  if (nc == 0 || !(h > 0.f)) {
    // choice without base data: -log(N * max_t) (kde_class.py:291-296), no floor
    const float v = -logf((float)N * A.max_t);
    for (int i = threadIdx.x; i < cnt; i += kEvalThreads) A.out_ll[pb + i] = v;
    return;
  }
  int nch = (nc + SSM_KDE_CHUNK - 1) / SSM_KDE_CHUNK;
  if (nch > A.n_chunks) nch = A.n_chunks;  // stay inside the workspace even if max_samples was too small
  const float p = (float)A.choice_total[g] / (float)N;
  // log p - log n - log h - log(2 pi)/2 - 64 ln 2
  const float cst = logf(p) - logf((float)nc) - logf(h) - 0.91893853320467274f - kScaleLog2 * 0.69314718055994531f;
  for (int i = threadIdx.x; i < cnt; i += kEvalThreads) {
    const float rt = A.pts_rt[pb + i];
    const float x = logf(fmaxf(rt, A.eps));
    float sum = 0.f;
    for (int c = 0; c < nch; ++c) sum += A.partial[(pb + i) * A.n_chunks + c];
    float ll = cst + logf(sum) - x;  // sum == 0 -> -inf -> floored
    if (!(ll > A.lb) || !(rt > 0.f)) ll = A.lb;
    A.out_ll[pb + i] = ll;
  }
}

// ---- LAN rows: [theta..., rt, choice value, ll], shuffled per set -------------------
// Per set, m evaluated points plus n_neg negative-RT rows (resample_mixture_strategy.py:
// 202-209: rt ~ U(-1, 1e-4), a uniform random choice, label neg_ll), all m + n_neg rows of a
// set shuffled together (:211-212) by sorting random keys in LDS.
struct AssembleArgs {
  const float* theta;        // [n_sets][n_theta]
  const float* pts_rt;       // [n_sets*m]
  const int8_t* pts_choice;  // [n_sets*m] choice index
  const float* ll;           // [n_sets*m]
  const int32_t* choice_values;
  float* rows;               // [n_sets*(m + n_neg)][n_theta + 3]
  int32_t n_theta;
  int32_t m;
  int32_t n_neg;
  int32_t n_choices;
  int32_t shuffle;
  float neg_ll;
  uint32_t k0, k1;
};

constexpr int kSortN = 1024;

__global__ void __launch_bounds__(kSortN) lan_assemble_kernel(AssembleArgs A) {
  __shared__ uint32_t key[kSortN];
  __shared__ int idx[kSortN];
  const int set = blockIdx.x;
  const int t = threadIdx.x;
  const int m = A.m;
  const int M = A.m + A.n_neg;  // rows of this set
  const int W = A.n_theta + 3;
  // random keys (Philox) for a uniform random permutation; padding sorts last
  if (t < M && A.shuffle) {
    const u32x4 r = philox4x32_10(u32x4{(uint32_t)t, (uint32_t)set, 0u, 13u}, A.k0, A.k1);
    key[t] = r.x & 0xFFFFFC00u | (uint32_t)t;  // distinct keys
  } else {
    key[t] = t < M ? (uint32_t)t : 0xFFFFFFFFu;
  }
  idx[t] = t;
  __syncthreads();
  for (int k = 2; k <= kSortN; k <<= 1) {
    for (int j = k >> 1; j > 0; j >>= 1) {
      const int p = t ^ j;
      if (p > t) {
        const bool up = (t & k) == 0;
        const uint32_t a = key[t], b = key[p];
        if ((a > b) == up) {
          key[t] = b; key[p] = a;
          const int ti = idx[t]; idx[t] = idx[p]; idx[p] = ti;
        }
      }
      __syncthreads();
    }
  }
  if (t < M) {
    const int src = idx[t];
    float* row = A.rows + ((int64_t)set * M + t) * W;
    for (int q = 0; q < A.n_theta; ++q) row[q] = A.theta[(int64_t)set * A.n_theta + q];
    if (src < m) {
      const int64_t si = (int64_t)set * m + src;
      row[A.n_theta] = A.pts_rt[si];
      row[A.n_theta + 1] = (float)A.choice_values[A.pts_choice[si]];
      row[A.n_theta + 2] = A.ll[si];
    } else {
      // negative-RT row src - m of this set: rt = -1 + (1e-4 + 1) u, choice index floor(u' C)
      const u32x4 r = philox4x32_10(u32x4{(uint32_t)(src - m), (uint32_t)set, 0u, 14u}, A.k0, A.k1);
      int c = (int)(u32_to_unit(r.y) * (float)A.n_choices);
      if (c >= A.n_choices) c = A.n_choices - 1;
      row[A.n_theta] = __builtin_fmaf(1.0001f, u32_to_unit(r.x), -1.0f);
      row[A.n_theta + 1] = (float)A.choice_values[c];
      row[A.n_theta + 2] = A.neg_ll;
    }
  }
}

// ---- host launchers --------------------------------------------------------------------
int32_t kde_fit_enqueue(int32_t n_sets, int32_t n_choices, const int32_t* choice_values, const float* rts,
                        const void* choices, int32_t choice_bytes, const int64_t* offsets, float* samples,
                        int32_t* kde_count, int32_t* choice_total, float* bandwidth, int64_t* kde_start,
                        hipStream_t stream) {
  if (n_sets < 0) return fail(SSM_EINVAL, "n_sets must be >= 0");
  if (n_choices < 1 || n_choices > SSM_MAX_PARTICLES) return fail(SSM_EINVAL, "n_choices must be in [1, 16]");
  if (choice_bytes != 1 && choice_bytes != 4) return fail(SSM_EINVAL, "choice_bytes must be 1 or 4");
  if (n_sets == 0) return SSM_OK;
  if (!choice_values || !rts || !choices || !offsets || !samples || !kde_count || !choice_total || !bandwidth ||
      !kde_start)
    return fail(SSM_EINVAL, "NULL buffer");
  FitArgs A{rts, choices, offsets, choice_values, samples, kde_count, choice_total, bandwidth, kde_start,
            n_choices, choice_bytes};
  if (n_choices <= 2)
    hipLaunchKernelGGL(kde_fit_kernel<2>, dim3(n_sets), dim3(kFitThreads), 0, stream, A);
  else if (n_choices <= 4)
    hipLaunchKernelGGL(kde_fit_kernel<4>, dim3(n_sets), dim3(kFitThreads), 0, stream, A);
  else
    hipLaunchKernelGGL(kde_fit_kernel<SSM_MAX_PARTICLES>, dim3(n_sets), dim3(kFitThreads), 0, stream, A);
  SSM_HIP(hipGetLastError());
  return SSM_OK;
}

int32_t kde_sample_enqueue(int32_t n_sets, int32_t n_choices, int32_t n_kde, int32_t n_unif, float max_t,
                           float unif_lo, uint64_t seed, const float* samples, const int32_t* kde_count,
                           const int32_t* choice_total, const float* bandwidth, const int64_t* kde_start,
                           const int64_t* offsets, const int32_t* n_by_override, float* pts_rt,
                           int8_t* pts_choice, int64_t* grp_begin, int32_t* grp_count, hipStream_t stream) {
  if (n_sets < 0 || n_kde < 0 || n_unif < 0) return fail(SSM_EINVAL, "negative sizes");
  if (n_choices < 1 || n_choices > SSM_MAX_PARTICLES) return fail(SSM_EINVAL, "n_choices must be in [1, 16]");
  if (n_kde + n_unif > kSampleThreads) return fail(SSM_EINVAL, "n_kde + n_unif must be <= 1024 points per set");
  if (n_sets == 0 || n_kde + n_unif == 0) return SSM_OK;
  if (!samples || !kde_count || !choice_total || !bandwidth || !kde_start || !offsets || !pts_rt || !pts_choice ||
      !grp_begin || !grp_count)
    return fail(SSM_EINVAL, "NULL buffer");
  SampleArgs A{samples,   kde_count, choice_total,  bandwidth, kde_start, offsets, pts_rt,  pts_choice,
               grp_begin, grp_count, n_by_override, n_choices, n_kde,     n_unif,  max_t,   unif_lo,
               (uint32_t)seed, (uint32_t)(seed >> 32)};
  hipLaunchKernelGGL(kde_sample_kernel, dim3(n_sets), dim3(kSampleThreads), 0, stream, A);
  SSM_HIP(hipGetLastError());
  return SSM_OK;
}

int32_t kde_eval_enqueue(int32_t n_sets, int32_t n_choices, const int64_t* grp_begin, const int32_t* grp_count,
                         const float* pts_rt, const float* samples, const int32_t* kde_count,
                         const int32_t* choice_total, const float* bandwidth, const int64_t* kde_start,
                         const int64_t* offsets, float max_t, float lb, float eps, int64_t max_samples,
                         float* workspace, float* out_ll, hipStream_t stream) {
  if (n_sets < 0 || max_samples < 0) return fail(SSM_EINVAL, "negative sizes");
  if (n_choices < 1 || n_choices > SSM_MAX_PARTICLES) return fail(SSM_EINVAL, "n_choices must be in [1, 16]");
  if (n_sets == 0) return SSM_OK;
  if (!grp_begin || !grp_count || !pts_rt || !samples || !kde_count || !choice_total || !bandwidth || !kde_start ||
      !offsets || !out_ll || !workspace)
    return fail(SSM_EINVAL, "NULL buffer");
  const int64_t nch = (max_samples + SSM_KDE_CHUNK - 1) / SSM_KDE_CHUNK;
  if (nch > 65535) return fail(SSM_EINVAL, "max_samples too large");
  EvalArgs A{pts_rt,   grp_begin, grp_count, samples,  kde_count,         choice_total, bandwidth, kde_start,
             offsets,  workspace, out_ll,    n_choices, (int32_t)(nch > 0 ? nch : 1), max_t, lb, eps};
  if (nch > 0) {
    hipLaunchKernelGGL(kde_eval_partial_kernel, dim3(n_sets * n_choices, (unsigned)nch), dim3(kEvalThreads), 0,
                       stream, A);
    SSM_HIP(hipGetLastError());
  }
  hipLaunchKernelGGL(kde_eval_finish_kernel, dim3(n_sets * n_choices), dim3(kEvalThreads), 0, stream, A);
  SSM_HIP(hipGetLastError());
  return SSM_OK;
}

// ---- evaluation without allocation ----------------------------------------------------
// Partial sums of ssm_kde_eval_noalloc: a fixed array in device memory (one copy per GPU,
// held until the process ends). Host code never passes its name to a kernel: the name is a
// host-side stand-in, so the device address is looked up once per device with
// hipGetSymbolAddress (with that device current) and kept.
static __device__ __align__(256) float scratch[SSM_KDE_EVAL_SCRATCH_FLOATS];

static int32_t scratch_for_current_device(float** out) {
  static std::mutex mu;
  static std::vector<float*> per_device;
  int dev = 0;
  SSM_HIP(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lock(mu);
  if ((int)per_device.size() <= dev) per_device.resize(dev + 1, nullptr);
  if (!per_device[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return hip_fail(e, "hipGetSymbolAddress(scratch)");
    if (!p) return fail(SSM_EHIP, "hipGetSymbolAddress(scratch) returned NULL");
    per_device[dev] = (float*)p;
  }
  *out = per_device[dev];
  return SSM_OK;
}

int32_t kde_eval_noalloc_enqueue(int32_t n_sets, int32_t n_choices, const int64_t* grp_begin,
                                 const int32_t* grp_count, const float* pts_rt, const float* samples,
                                 const int32_t* kde_count, const int32_t* choice_total, const float* bandwidth,
                                 const int64_t* kde_start, const int64_t* offsets, float max_t, float lb, float eps,
                                 int64_t max_samples, int64_t n_points, float* out_ll, hipStream_t stream) {
  if (n_sets < 0 || max_samples < 0 || n_points < 0) return fail(SSM_EINVAL, "negative sizes");
  const int64_t nch = max_samples > 0 ? (max_samples + SSM_KDE_CHUNK - 1) / SSM_KDE_CHUNK : 1;
  if (nch > 65535) return fail(SSM_EINVAL, "max_samples too large");
  if (n_points > (int64_t)SSM_KDE_EVAL_SCRATCH_FLOATS / nch)
    return fail(SSM_EINVAL, "n_points * ceil(max_samples / 8192) = " + std::to_string(n_points * nch) +
                                " partial sums exceed the no-allocation scratch of " +
                                std::to_string((int64_t)SSM_KDE_EVAL_SCRATCH_FLOATS) +
                                " floats; evaluate fewer points per call or use ssm_kde_eval with a workspace");
  if (n_sets == 0) return SSM_OK;
  float* ws = nullptr;
  const int32_t rc = scratch_for_current_device(&ws);
  if (rc) return rc;
  return kde_eval_enqueue(n_sets, n_choices, grp_begin, grp_count, pts_rt, samples, kde_count, choice_total,
                          bandwidth, kde_start, offsets, max_t, lb, eps, max_samples, ws, out_ll, stream);
}

int32_t lan_assemble_enqueue(int32_t n_sets, int32_t n_theta, int32_t m, int32_t n_neg, int32_t n_choices,
                             float neg_ll, int32_t shuffle, uint64_t seed, const float* theta, const float* pts_rt,
                             const int8_t* pts_choice, const float* ll, const int32_t* choice_values, float* rows,
                             hipStream_t stream) {
  if (n_sets < 0 || n_theta < 0 || m < 0 || n_neg < 0) return fail(SSM_EINVAL, "negative sizes");
  if (m + n_neg > kSortN) return fail(SSM_EINVAL, "at most 1024 rows per set");
  if (n_choices < 1 || n_choices > SSM_MAX_PARTICLES) return fail(SSM_EINVAL, "n_choices must be in [1, 16]");
  if (n_sets == 0 || m + n_neg == 0) return SSM_OK;
  if (!theta || !choice_values || !rows || (m > 0 && (!pts_rt || !pts_choice || !ll)))
    return fail(SSM_EINVAL, "NULL buffer");
  AssembleArgs A{theta, pts_rt, pts_choice, ll, choice_values, rows, n_theta, m, n_neg, n_choices, shuffle, neg_ll,
                 (uint32_t)seed, (uint32_t)(seed >> 32)};
  hipLaunchKernelGGL(lan_assemble_kernel, dim3(n_sets), dim3(kSortN), 0, stream, A);
  SSM_HIP(hipGetLastError());
  return SSM_OK;
}

}  // namespace ssmhip
