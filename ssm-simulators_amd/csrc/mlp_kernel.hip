// Likelihood-surrogate MLP on gfx950: 7 -> 100 -> 100 -> 120 -> 1, tanh hidden
// layers, BF16 MFMA (v_mfma_f32_16x16x32_bf16) with FP32 accumulation and
// FP32 master weights. The reference has no counterpart (LAN training lives
// outside ssm-simulators); the inputs are its lan_data columns
// (5 angle-DDM theta, rt, choice; resample_mixture_strategy.py:136-139).
//
// Data flow inside a wavefront (16 rows per MFMA row tile):
//   every layer computes Z^T = W * H^T (M = output features, N = rows,
//   K = input features), so the accumulator of lane l holds row l%16 and
//   output features 16a + 4(l/16) + i (tile a, i < 4). The next layer takes
//   K block b from output tiles 2b and 2b+1: lane l's 8 K slots are exactly
//   the 8 features it already holds, so activations never leave registers.
//   The weight fragments are packed in that permuted K order (pack kernel).
// Hidden widths are padded to 128 (8 tiles), the output to one 16-wide tile;
// padded weights and biases are zero, so padded units stay at tanh(0) = 0.
#include <hip/hip_runtime.h>

#include "ssmhip_internal.h"

namespace ssmhip {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// real sizes
constexpr int kIn = 7, kH1 = 100, kH2 = 100, kH3 = 120, kOut = 1;
static_assert(kH1 <= 112 && kH2 <= 112, "the forward kernel skips output tile 7 of layers 1 and 2");
// parameter offsets in the flat FP32 master buffer (nn.Linear layout W[out][in])
constexpr int oW1 = 0, oB1 = oW1 + kH1 * kIn, oW2 = oB1 + kH1, oB2 = oW2 + kH2 * kH1, oW3 = oB2 + kH2,
              oB3 = oW3 + kH3 * kH2, oW4 = oB3 + kH3, oB4 = oW4 + kOut * kH3, kNParams = oB4 + kOut;
// packed fragment offsets (bf16 elements); each fragment = 64 lanes x 8
constexpr int kFrag = 64 * 8;
constexpr int fF1 = 0;                      // fwd L1: 8 tiles x 1 block
constexpr int fF2 = fF1 + 8 * 1 * kFrag;    // fwd L2: 8 x 4
constexpr int fF3 = fF2 + 8 * 4 * kFrag;    // fwd L3: 8 x 4
constexpr int fF4 = fF3 + 8 * 4 * kFrag;    // fwd L4: 1 x 4
constexpr int fB1 = fF4 + 1 * 4 * kFrag;    // bwd L1 (dX): 1 x 4
constexpr int fB2 = fB1 + 1 * 4 * kFrag;    // bwd L2: 8 x 4
constexpr int fB3 = fB2 + 8 * 4 * kFrag;    // bwd L3: 8 x 4
constexpr int fB4 = fB3 + 8 * 4 * kFrag;    // bwd L4: 8 x 1
constexpr int kNFrag = fB4 + 8 * 1 * kFrag;
// padded FP32 biases
constexpr int pB1 = 0, pB2 = 128, pB3 = 256, pB4 = 384, kNBias = 400;
// feature-major activation / delta buffers (bf16 [features][batch])
constexpr int kActRows = 8 + 128 * 3;     // H0 (8 inputs), H1, H2, H3
constexpr int kDeltaRows = 128 * 3 + 16;  // dZ1, dZ2, dZ3, dZ4

__device__ __forceinline__ int chain_feature(int b, int g, int j) { return 16 * (2 * b + (j >> 2)) + 4 * g + (j & 3); }

// Hidden-layer weights and biases are packed pre-multiplied by 2 log2(e), so the
// accumulator holds zs = 2 log2(e) z and tanh z = 1 - 2 / (2^zs + 1):
// v_exp_f32, add, v_rcp_f32, fma.
constexpr float kTanhScale = 2.8853900817779268f;
__device__ __forceinline__ float tanh_fast(float zs) {
  return __builtin_fmaf(-2.0f, __builtin_amdgcn_rcpf(__builtin_amdgcn_exp2f(zs) + 1.0f), 1.0f);
}

// Input transform: xnorm = shift[7], scale[7], rt mode, t index. Mode 1: input 5 (rt)
// enters as log(max(rt, 1e-4)); mode 2: as log(max(rt - x[t index], 1e-4)), the log of
// the time since the non-decision time; then every input as (v - shift) * scale.
constexpr float kRtFloor = 1e-4f;
__device__ __forceinline__ int t_index(const float* xnorm) {
  return min(max((int)xnorm[2 * kIn + 1], 0), kIn - 1);
}
__device__ __forceinline__ float input_value(const float* xnorm, int j, const float* xr) {
  float v = xr[j];
  if (!xnorm) return v;
  if (j == 5) {
    const float mode = xnorm[2 * kIn];
    if (mode == 2.f) v -= xr[t_index(xnorm)];
    if (mode != 0.f) v = logf(fmaxf(v, kRtFloor));
  }
  return (v - xnorm[j]) * xnorm[kIn + j];
}

__device__ __forceinline__ f32x4 mfma(bf16x8 a, bf16x8 b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// Keep weight fragments in memory (L1/L2): without this the row loop's
// invariant fragment loads are hoisted into hundreds of live registers.
__device__ __forceinline__ const __bf16* opaque(const __bf16* p) {
  asm volatile("" : "+s"(p));
  return p;
}

__device__ __forceinline__ bf16x8 load_frag(const __bf16* frags, int off, int idx, int lane) {
  return *reinterpret_cast<const bf16x8*>(frags + off + (idx * 64 + lane) * 8);
}

__device__ __forceinline__ bf16x8 pack_pair(f32x4 lo, f32x4 hi) {
  bf16x8 r;
  r[0] = (__bf16)lo[0]; r[1] = (__bf16)lo[1]; r[2] = (__bf16)lo[2]; r[3] = (__bf16)lo[3];
  r[4] = (__bf16)hi[0]; r[5] = (__bf16)hi[1]; r[6] = (__bf16)hi[2]; r[7] = (__bf16)hi[3];
  return r;
}

__device__ __forceinline__ f32x4 bias_tile(const float* bias, int off, int a, int g) {
  const float* p = bias + off + 16 * a + 4 * g;
  return f32x4{p[0], p[1], p[2], p[3]};
}

// ---- pack FP32 master weights into MFMA fragments -------------------------------
__device__ float w_at(const float* P, int layer, int out, int in) {
  switch (layer) {
    case 1: return (out < kH1 && in < kIn) ? P[oW1 + out * kIn + in] : 0.f;
    case 2: return (out < kH2 && in < kH1) ? P[oW2 + out * kH1 + in] : 0.f;
    case 3: return (out < kH3 && in < kH2) ? P[oW3 + out * kH2 + in] : 0.f;
    default: return (out < kOut && in < kH3) ? P[oW4 + out * kH3 + in] : 0.f;
  }
}

__global__ void mlp_pack_kernel(const float* __restrict__ P, __bf16* __restrict__ frags, float* __restrict__ bias) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;  // element index into frags
  if (e < kNFrag) {
    // locate region
    int layer, base, KB;
    bool bwd;
    if (e < fF2) { layer = 1; base = fF1; KB = 1; bwd = false; }
    else if (e < fF3) { layer = 2; base = fF2; KB = 4; bwd = false; }
    else if (e < fF4) { layer = 3; base = fF3; KB = 4; bwd = false; }
    else if (e < fB1) { layer = 4; base = fF4; KB = 4; bwd = false; }
    else if (e < fB2) { layer = 1; base = fB1; KB = 4; bwd = true; }
    else if (e < fB3) { layer = 2; base = fB2; KB = 4; bwd = true; }
    else if (e < fB4) { layer = 3; base = fB3; KB = 4; bwd = true; }
    else { layer = 4; base = fB4; KB = 1; bwd = true; }
    const int r = e - base;
    const int j = r & 7, lane = (r >> 3) & 63, fi = r >> 9;
    const int a = fi / KB, b = fi % KB, g = lane >> 4, m = 16 * a + (lane & 15);
    float v;
    if (!bwd) {
      // A[m = out][k]: k from raw inputs (layer 1) or the previous layer's tiles
      const int kin = (layer == 1) ? (8 * g + j) : chain_feature(b, g, j);
      v = w_at(P, layer, m, kin) * (layer < 4 ? kTanhScale : 1.0f);
    } else {
      // A[m = in][k = out]: k from this layer's output tiles
      v = w_at(P, layer, chain_feature(b, g, j), m);
    }
    frags[e] = (__bf16)v;
  }
  if (e < kNBias) {
    float v = 0.f;
    if (e < pB2) v = e < kH1 ? kTanhScale * P[oB1 + e] : 0.f;
    else if (e < pB3) v = (e - pB2) < kH2 ? kTanhScale * P[oB2 + e - pB2] : 0.f;
    else if (e < pB4) v = (e - pB3) < kH3 ? kTanhScale * P[oB3 + e - pB3] : 0.f;
    else v = (e - pB4) < kOut ? P[oB4 + e - pB4] : 0.f;
    bias[e] = v;
  }
}

// ---- forward (inference) --------------------------------------------------------------
// x: [n][7] fp32 (theta..., rt, choice), y: [n] fp32, 32 rows per wave and loop iteration.
// xnorm (optional): shift[7], scale[7]; the network sees (x - shift) * scale.
// Each block builds its image of the forward weights (68 KB of bf16) in LDS from the packed
// fragments once, so the per-row weight traffic is LDS bandwidth rather than L2 bandwidth.

// Inference carries each hidden unit as u = 1 / (2^zs + 1) = (1 - tanh z) / 2, one
// exp2, add and reciprocal, instead of tanh z = 1 - 2u (one FMA more per unit). The layer
// that consumes the units takes the affine map: with packed weights P (bf16) it computes
//   bias' + sum_k (-2 P_k) u_k,   bias' = bias + sum_k P_k,
// which equals bias + sum_k P_k (1 - 2 u_k) = bias + sum_k P_k tanh z_k. -2 P is exact in
// bf16, so the only change against carrying tanh is where bf16 rounds: u instead of tanh z.
// The block prologue builds -2 P and bias' in LDS from the packed weights.

// The forward kernel works on 32 x 32 tiles with v_mfma_f32_32x32x16_bf16 (A 32 x 16,
// B 16 x 32, D 32 x 32; lane l = 32 h + n):
//   A: lane l holds A[m = l % 32][k = 8 h + i], i < 8;   B: B[k = 8 h + i][n = l % 32];
//   D: 16 floats, register r = D[m = 8 (r / 4) + 4 h + r % 4][n = l % 32].
// Every layer computes Z^T = W H^T (M = units, N = 32 rows, K = input units), so register r
// of output tile t holds unit 32 t + 8 (r / 4) + 4 h + r % 4 of row n. Registers 0..7 and
// 8..15 of tile t are K blocks 2t and 2t+1 of the next layer as they stand (8 bf16 per lane):
// K block b lists, for lane half h and item i, unit fwd_unit(b, h, i), and the weights in the
// block's LDS image are laid out in that K order.
//
// Padding is skipped per register: the last tile of the 100-unit layers keeps registers
// 0..3 (units 96..103), that of the 120-unit layer registers 0..11 (units 96..119); the other
// registers are never computed and enter the next layer as 0. Against 16 x 16 tiles, where
// padding fills whole lanes, this computes 328 instead of 352 units per row, and one MFMA
// covers 32 rows instead of 16 (68 instead of 142 per 32 rows).
//
// The biases ride in the products. Input slots 7 and 8 of layer 1 are constants 1, with the
// bias split into bf16 high and low parts as their weights (hi + lo is the FP32 bias to
// 2^-16 relative). Units 100 and 101 of layers 1 and 2, padding of the network, are
// constant-one units: weight -200 from a constant input gives u = 1 / (2^-200 + 1) = 1
// exactly, and the next layer carries its bias' (high, low) as their weights. The output
// layer's bias' is added in FP32.
constexpr int kF32L1 = 0, kF32L2 = 4, kF32L3 = kF32L2 + 28, kF32L4 = kF32L3 + 28, kF32N = kF32L4 + 8;
constexpr int kImg32 = kF32N * kFrag;  // bf16 elements of a block's LDS image (69,632 B)
constexpr int kConstU0 = 100, kConstU1 = 101;
static_assert(kH1 <= kConstU0 && kH2 <= kConstU0, "the constant-one units sit in the padding of layers 1 and 2");
static_assert(kH3 <= 120, "the output layer reads registers 0..11 of layer 3's last tile");

typedef float f32x16 __attribute__((ext_vector_type(16)));

__device__ __forceinline__ f32x16 mfma32(bf16x8 a, bf16x8 b, f32x16 c) {
  return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ int fwd_unit(int b, int h, int i) { return 16 * b + 8 * (i >> 2) + 4 * h + (i & 3); }
__device__ __forceinline__ bf16x8 frag32(const __bf16* img, int f, int lane) {
  return *reinterpret_cast<const bf16x8*>(img + (f * 64 + lane) * 8);
}

// u = 1 / (2^zs + 1) of registers 0..R-1 of a tile
template <int R>
__device__ __forceinline__ void activate32(f32x16& c) {
#pragma unroll
  for (int r = 0; r < R; ++r) c[r] = __builtin_amdgcn_rcpf(__builtin_amdgcn_exp2f(c[r]) + 1.0f);
}
// K blocks 2t and 2t+1 of the next layer from tile t; registers from R on are padding: 0
template <int R>
__device__ __forceinline__ void pack32(const f32x16& c, bf16x8& lo, bf16x8& hi) {
#pragma unroll
  for (int i = 0; i < 8; ++i) lo[i] = i < R ? (__bf16)c[i] : (__bf16)0.f;
#pragma unroll
  for (int i = 0; i < 8; ++i) hi[i] = 8 + i < R ? (__bf16)c[8 + i] : (__bf16)0.f;
}

// One hidden layer, 4 output tiles from KB input K blocks in Bin. Bin's blocks 6 and 7 come
// from the previous layer's last tile `last` (RPREV real registers; 0: none pending), whose
// activation runs beside the MFMAs of K blocks 0..5 of this layer's first tile: only blocks
// 6 and 7 read it. The activation of tile t-1 runs after the MFMAs of tile t are issued. On
// return `last` holds tile 3, whose activation is left to the next layer.
template <int KB, int RPREV>
__device__ __forceinline__ void hidden32(const __bf16* img, int f0, int lane, bf16x8 (&Bin)[8], f32x16& last,
                                         bf16x8 (&Bout)[8]) {
  f32x16 prev;
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    f32x16 c = {};
#pragma unroll
    for (int b = 0; b < KB; ++b) {
      if constexpr (RPREV > 0) {
        if (t == 0 && b == 6) {
          activate32<RPREV>(last);
          pack32<RPREV>(last, Bin[6], Bin[7]);
        }
      }
      c = mfma32(frag32(img, f0 + t * KB + b, lane), Bin[b], c);
    }
    if (t > 0) {
      activate32<16>(prev);
      pack32<16>(prev, Bout[2 * (t - 1)], Bout[2 * (t - 1) + 1]);
    }
    prev = c;
  }
  last = prev;
}

// bf16 high and low parts of v: v - hi is exact in FP32
__device__ __forceinline__ void split_bf16(float v, __bf16& hi, __bf16& lo) {
  hi = (__bf16)v;
  lo = (__bf16)(v - (float)hi);
}

__global__ void __launch_bounds__(1024, 1) mlp_forward_kernel(const float* __restrict__ x, int64_t n,
                                                          const __bf16* __restrict__ gfrags,
                                                          const float* __restrict__ bias,
                                                          const float* __restrict__ xnorm, float* __restrict__ y) {
  __shared__ __attribute__((aligned(16))) __bf16 img[kImg32];
  __shared__ __attribute__((aligned(16))) float lbias[kNBias];
  const int lane = threadIdx.x & 63, h = lane >> 5, nr = lane & 31;
  {
    // 1. the image from the packed 16 x 16 fragments, four K slots (one 8-byte run of the
    //    source) at a time; layers 2..4 consume u, so they take -2 P
    for (int c = threadIdx.x; c < kF32N * 128; c += blockDim.x) {
      const int f = c >> 7, l = (c >> 1) & 63, i0 = 4 * (c & 1);
      const int m = l & 31, hl = l >> 5;
      int layer, t, b;
      if (f < kF32L2) { layer = 1; t = f; b = 0; }
      else if (f < kF32L3) { layer = 2; t = (f - kF32L2) / 7; b = (f - kF32L2) % 7; }
      else if (f < kF32L4) { layer = 3; t = (f - kF32L3) / 7; b = (f - kF32L3) % 7; }
      else { layer = 4; t = 0; b = f - kF32L4; }
      const int out = 32 * t + m;
      uint2 v = make_uint2(0u, 0u);
      if (layer == 1) {
        const int k0 = 8 * hl + i0;  // inputs k0..k0+3 (slot k = input k)
        v = *reinterpret_cast<const uint2*>(gfrags + fF1 + ((out >> 4) * 64 + (out & 15) + 16 * (k0 >> 3)) * 8 + (k0 & 7));
      } else if (layer < 4 || out < 16) {
        const int u0 = fwd_unit(b, hl, i0);  // units u0..u0+3
        const int base = layer == 2 ? fF2 : (layer == 3 ? fF3 : fF4);
        const int lsrc = (out & 15) + 16 * ((u0 >> 2) & 3), j0 = 4 * ((u0 >> 4) & 1);
        uint2 w = *reinterpret_cast<const uint2*>(gfrags + base + (((out >> 4) * 4 + (u0 >> 5)) * 64 + lsrc) * 8 + j0);
        __bf16 q[4];
        __builtin_memcpy(q, &w, 8);
#pragma unroll
        for (int k = 0; k < 4; ++k) q[k] = (__bf16)(-2.0f * (float)q[k]);
        __builtin_memcpy(&v, q, 8);
      }
      *reinterpret_cast<uint2*>(img + (f * 64 + l) * 8 + i0) = v;
    }
    // 2. bias' = bias + row sums of P for layers 2..4 (MFMAs against all-ones columns over
    //    the 16 x 16 fragments: 17 output tiles, one per wave in turn); layer 1 takes bias
    for (int i = threadIdx.x; i < pB2; i += blockDim.x) lbias[i] = bias[i];
    const int g = lane >> 4, r = lane & 15;
    bf16x8 ones;
#pragma unroll
    for (int j = 0; j < 8; ++j) ones[j] = (__bf16)1.0f;
    for (int task = threadIdx.x >> 6; task < 17; task += blockDim.x >> 6) {
      const int a = task & 7;
      const int fo = task < 8 ? fF2 : (task < 16 ? fF3 : fF4);
      const int bo = task < 8 ? pB2 : (task < 16 ? pB3 : pB4);
      f32x4 c = bias_tile(bias, bo, a, g);
#pragma unroll
      for (int b = 0; b < 4; ++b) c = mfma(load_frag(gfrags, fo, a * 4 + b, lane), ones, c);
      if (r == 0) {
#pragma unroll
        for (int i = 0; i < 4; ++i) lbias[bo + 16 * a + 4 * g + i] = c[i];
      }
    }
  }
  __syncthreads();
  // 3. the bias slots: layer 1's constant inputs 7 and 8, and the constant-one units 100
  //    and 101 as inputs of layers 2 and 3 (K block 6, lane half 1, items 0 and 1)
  for (int e = threadIdx.x; e < 3 * 128; e += blockDim.x) {
    const int layer = 1 + e / 128, out = e & 127, t = out >> 5, m = out & 31;
    const bool konst = layer < 3 && (out == kConstU0 || out == kConstU1);
    __bf16 hi, lo;
    if (konst) {
      hi = (__bf16)-200.0f;  // exact in bf16
      lo = (__bf16)0.0f;
    } else {
      split_bf16(lbias[(layer == 1 ? pB1 : (layer == 2 ? pB2 : pB3)) + out], hi, lo);
    }
    if (layer == 1) {
      img[(kF32L1 + t) * kFrag + m * 8 + 7] = hi;          // slot 7: lane m, item 7
      img[(kF32L1 + t) * kFrag + (32 + m) * 8 + 0] = lo;   // slot 8: lane 32 + m, item 0
    } else {
      const int f = (layer == 2 ? kF32L2 : kF32L3) + t * 7 + 6;
      img[f * kFrag + (32 + m) * 8 + 0] = hi;  // unit 100
      img[f * kFrag + (32 + m) * 8 + 1] = lo;  // unit 101
    }
  }
  __syncthreads();
  const float bias4 = lbias[pB4];
  // the input transform, read once per block (the identity when no transform is given:
  // (v - 0) * 1 = v)
  float shift[kIn], scale[kIn], mode = 0.f;
  int tix = 0;
#pragma unroll
  for (int j = 0; j < kIn; ++j) {
    shift[j] = 0.f;
    scale[j] = 1.f;
  }
  if (xnorm) {
#pragma unroll
    for (int j = 0; j < kIn; ++j) {
      shift[j] = xnorm[j];
      scale[j] = xnorm[kIn + j];
    }
    mode = xnorm[2 * kIn];
    tix = t_index(xnorm);
  }
  const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) >> 6;
  for (int64_t r0 = wave * 32; r0 < n; r0 += nwaves * 32) {
    // an offset the compiler cannot see through keeps the LDS fragment reads inside the row
    // loop (hoisting them would need hundreds of registers)
    int lo = 0;
    asm volatile("" : "+v"(lo));
    const __bf16* fi = img + lo;
    const int64_t row = r0 + nr;
    bf16x8 B[8], Bn[8];
    {
      // layer 1 input slots: lane half 0 holds inputs 0..6 of its row and the constant 1 of
      // slot 7, lane half 1 the constant 1 of slot 8 and zeros
      bf16x8 v = {};
      if (h == 0) {
        if (row < n) {
          const float* xr = x + row * kIn;
#pragma unroll
          for (int j = 0; j < kIn; ++j) {
            float q = xr[j];
            if (j == 5) {
              if (mode == 2.f) q -= xr[tix];
              if (mode != 0.f) q = logf(fmaxf(q, kRtFloor));
            }
            v[j] = (__bf16)((q - shift[j]) * scale[j]);
          }
        }
        v[7] = (__bf16)1.0f;
      } else {
        v[0] = (__bf16)1.0f;
      }
      B[0] = v;
    }
    f32x16 last;
    hidden32<1, 0>(fi, kF32L1, lane, B, last, Bn);
    hidden32<7, 4>(fi, kF32L2, lane, Bn, last, B);
    hidden32<7, 4>(fi, kF32L3, lane, B, last, Bn);
    // output layer: one tile (unit 0), K blocks 0..7; layer 3's last tile keeps registers 0..11
    f32x16 c = {};
#pragma unroll
    for (int b = 0; b < 8; ++b) {
      if (b == 6) {
        activate32<12>(last);
        pack32<12>(last, Bn[6], Bn[7]);
      }
      c = mfma32(frag32(fi, kF32L4 + b, lane), Bn[b], c);
    }
    if (h == 0 && row < n) y[row] = c[0] + bias4;
  }
}

// ---- fused forward + loss + backward-data (training and input gradients) -------------
// mode 0: dy = grad_y[row]; mode 1: Huber(y - target, delta) averaged with weight `scale`.
// Optional outputs: acts/deltas (feature-major bf16, ld = batch), dx [n][8] fp32, y, loss [n].
// TrainArgs is declared in ssmhip_internal.h

__device__ __forceinline__ void store_fm(__bf16* buf, int64_t ld, int feature, int64_t row, float v) {
  buf[(int64_t)feature * ld + row] = (__bf16)v;
}

__device__ __forceinline__ float bf_at(const bf16x8 (&hb)[4], int a, int i) {
  // element i of D tile a inside the packed operands: tile a sits in hb[a / 2] at (a & 1) * 4
  return (float)hb[a >> 1][((a & 1) << 2) + i];
}

__global__ void __launch_bounds__(256) mlp_train_kernel(TrainArgs A) {
  const int lane = threadIdx.x & 63, g = lane >> 4, r = lane & 15;
  const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) >> 6;
  for (int64_t r0 = wave * 16; r0 < A.n; r0 += nwaves * 16) {
    A.frags = opaque(A.frags);
    const int64_t row = r0 + r;
    const bool live = row < A.n;
    // hidden activations kept as the packed bf16 MFMA operands (the derivative uses them too)
    bf16x8 x0, h1[4], h2[4], h3[4];
    f32x4 acc[8];
    {
      bf16x8 v = {};
      if (g == 0 && live) {
#pragma unroll
        for (int jj = 0; jj < kIn; ++jj) {
          v[jj] = (__bf16)input_value(A.xnorm, jj, A.x + row * kIn);
        }
      }
      x0 = v;
      if (A.acts && g == 0 && live) {
#pragma unroll
        for (int jj = 0; jj < 8; ++jj) A.acts[(int64_t)jj * A.ld + row] = v[jj];
      }
    }
#pragma unroll
    for (int a = 0; a < 8; ++a) {
      __builtin_amdgcn_sched_barrier(0);
      acc[a] = mfma(load_frag(A.frags, fF1, a, lane), x0, bias_tile(A.bias, pB1, a, g));
#pragma unroll
      for (int i = 0; i < 4; ++i) acc[a][i] = tanh_fast(acc[a][i]);
    }
#pragma unroll
    for (int b = 0; b < 4; ++b) h1[b] = pack_pair(acc[2 * b], acc[2 * b + 1]);
#pragma unroll
    for (int a = 0; a < 8; ++a) {
      __builtin_amdgcn_sched_barrier(0);  // bound the fragment loads in flight (register pressure)
      f32x4 c = bias_tile(A.bias, pB2, a, g);
#pragma unroll
      for (int b = 0; b < 4; ++b) c = mfma(load_frag(A.frags, fF2, a * 4 + b, lane), h1[b], c);
#pragma unroll
      for (int i = 0; i < 4; ++i) acc[a][i] = tanh_fast(c[i]);
    }
#pragma unroll
    for (int b = 0; b < 4; ++b) h2[b] = pack_pair(acc[2 * b], acc[2 * b + 1]);
#pragma unroll
    for (int a = 0; a < 8; ++a) {
      __builtin_amdgcn_sched_barrier(0);  // bound the fragment loads in flight (register pressure)
      f32x4 c = bias_tile(A.bias, pB3, a, g);
#pragma unroll
      for (int b = 0; b < 4; ++b) c = mfma(load_frag(A.frags, fF3, a * 4 + b, lane), h2[b], c);
#pragma unroll
      for (int i = 0; i < 4; ++i) acc[a][i] = tanh_fast(c[i]);
    }
#pragma unroll
    for (int b = 0; b < 4; ++b) h3[b] = pack_pair(acc[2 * b], acc[2 * b + 1]);
    f32x4 yo = bias_tile(A.bias, pB4, 0, g);
#pragma unroll
    for (int b = 0; b < 4; ++b) yo = mfma(load_frag(A.frags, fF4, b, lane), h3[b], yo);
    // lanes of group 0 hold y of their row in element 0
    float dy = 0.f;
    if (g == 0 && live) {
      const float yv = yo[0];
      if (A.y) A.y[row] = yv;
      if (A.mode == 0) {
        dy = A.target_or_grad[row];
      } else {
        const float e = yv - A.target_or_grad[row];
        const float ae = fabsf(e);
        if (A.loss) A.loss[row] = ae <= A.delta ? 0.5f * e * e : A.delta * (ae - 0.5f * A.delta);
        dy = fminf(fmaxf(e, -A.delta), A.delta) * A.scale;
      }
    }
    if (A.acts && live) {
#pragma unroll
      for (int a = 0; a < 8; ++a)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const int f = 16 * a + 4 * g + i;
          A.acts[(int64_t)(8 + f) * A.ld + row] = h1[a >> 1][((a & 1) << 2) + i];
          A.acts[(int64_t)(8 + 128 + f) * A.ld + row] = h2[a >> 1][((a & 1) << 2) + i];
          A.acts[(int64_t)(8 + 256 + f) * A.ld + row] = h3[a >> 1][((a & 1) << 2) + i];
        }
    }
    // dZ4 tile: feature 0 only (lane group 0, element 0)
    const f32x4 dz4 = {g == 0 ? dy : 0.f, 0.f, 0.f, 0.f};
    if (A.deltas && live) {
#pragma unroll
      for (int i = 0; i < 4; ++i) store_fm(A.deltas, A.ld, 384 + 4 * g + i, row, dz4[i]);
    }
    bf16x8 B[4];
    // dh3 = W4^T dz4 (M = 128 inputs, K = one block: tile 0 = dz4, tile 1 = 0); dz3 = dh3 (1 - h3^2)
    const bf16x8 bd4 = pack_pair(dz4, f32x4{0.f, 0.f, 0.f, 0.f});
#pragma unroll
    for (int a = 0; a < 8; ++a) {
      __builtin_amdgcn_sched_barrier(0);
      acc[a] = mfma(load_frag(A.frags, fB4, a, lane), bd4, f32x4{0.f, 0.f, 0.f, 0.f});
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const float h = bf_at(h3, a, i);
        acc[a][i] *= 1.0f - h * h;
      }
    }
    if (A.deltas && live) {
#pragma unroll
      for (int a = 0; a < 8; ++a)
#pragma unroll
        for (int i = 0; i < 4; ++i) store_fm(A.deltas, A.ld, 256 + 16 * a + 4 * g + i, row, acc[a][i]);
    }
#pragma unroll
    for (int b = 0; b < 4; ++b) B[b] = pack_pair(acc[2 * b], acc[2 * b + 1]);
    // dh2 = W3^T dz3; dz2 = dh2 (1 - h2^2)
#pragma unroll
    for (int a = 0; a < 8; ++a) {
      __builtin_amdgcn_sched_barrier(0);  // bound the fragment loads in flight (register pressure)
      f32x4 c = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int b = 0; b < 4; ++b) c = mfma(load_frag(A.frags, fB3, a * 4 + b, lane), B[b], c);
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const float h = bf_at(h2, a, i);
        acc[a][i] = c[i] * (1.0f - h * h);
      }
    }
    if (A.deltas && live) {
#pragma unroll
      for (int a = 0; a < 8; ++a)
#pragma unroll
        for (int i = 0; i < 4; ++i) store_fm(A.deltas, A.ld, 128 + 16 * a + 4 * g + i, row, acc[a][i]);
    }
#pragma unroll
    for (int b = 0; b < 4; ++b) B[b] = pack_pair(acc[2 * b], acc[2 * b + 1]);
    // dh1 = W2^T dz2; dz1 = dh1 (1 - h1^2)
#pragma unroll
    for (int a = 0; a < 8; ++a) {
      __builtin_amdgcn_sched_barrier(0);  // bound the fragment loads in flight (register pressure)
      f32x4 c = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int b = 0; b < 4; ++b) c = mfma(load_frag(A.frags, fB2, a * 4 + b, lane), B[b], c);
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const float h = bf_at(h1, a, i);
        acc[a][i] = c[i] * (1.0f - h * h);
      }
    }
    if (A.deltas && live) {
#pragma unroll
      for (int a = 0; a < 8; ++a)
#pragma unroll
        for (int i = 0; i < 4; ++i) store_fm(A.deltas, A.ld, 16 * a + 4 * g + i, row, acc[a][i]);
    }
    if (A.dx) {
#pragma unroll
      for (int b = 0; b < 4; ++b) B[b] = pack_pair(acc[2 * b], acc[2 * b + 1]);
      f32x4 c = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int b = 0; b < 4; ++b) c = mfma(load_frag(A.frags, fB1, b, lane), B[b], c);
      // lane holds dx[row][4g + i]; inputs are features 0..6 (groups 0 and 1).
      // d/dx_raw = d/dx_norm * scale, times d(feature)/d(rt) = 1/u for input 5 with a log
      // transform (u = rt, or rt - t in mode 2, zero below the floor); in mode 2 the t input
      // also receives -1/u of it, taken from the row's group-1 lane.
      const float mode = A.xnorm ? A.xnorm[2 * kIn] : 0.f;
      float d5 = 0.f;
      if (mode != 0.f && live) {
        float u = A.x[row * kIn + 5];
        if (mode == 2.f) u -= A.x[row * kIn + t_index(A.xnorm)];
        d5 = u > kRtFloor ? c[1] * A.xnorm[kIn + 5] / u : 0.f;
      }
      const float d5_row = __shfl(d5, 16 + r, 64);
      if (live && g < 2) {
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const int k = 4 * g + i;
          float d = (A.xnorm && k < kIn) ? c[i] * A.xnorm[kIn + k] : c[i];
          if (k == 5 && mode != 0.f) d = d5;
          if (mode == 2.f && k == t_index(A.xnorm)) d -= d5_row;
          A.dx[row * 8 + k] = d;
        }
      }
    }
  }
}

// ---- the same forward + loss + backward-data, four waves per 16-row tile -------------
// Wave w of a block computes output tiles 2w and 2w+1 of every 128-feature layer. Packed
// as an MFMA operand, those two tiles are exactly K block w of the next layer's input
// (pack_pair keeps each lane's own values), so after each layer the four waves swap their
// blocks through LDS and continue. Every output element is accumulated in the same order
// as in mlp_train_kernel, so y, loss, acts and deltas are bit-identical to it; a batch of
// a few thousand rows then fills the chip four times more densely, with a quarter of the
// registers per wave. No input gradient (dx must be NULL).
__global__ void __launch_bounds__(256) mlp_train_team_kernel(TrainArgs A) {
  __shared__ bf16x8 xch[2][4][64];  // double-buffered operand exchange: [buffer][K block][lane]
  const int lane = threadIdx.x & 63, g = lane >> 4, r = lane & 15;
  const int w = threadIdx.x >> 6;
  const int a0 = 2 * w, a1 = 2 * w + 1;
  const int64_t ntiles = (A.n + 15) / 16;
  for (int64_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    A.frags = opaque(A.frags);
    const int64_t row = tile * 16 + r;
    const bool live = row < A.n;
    bf16x8 x0 = {};
    if (g == 0 && live) {
#pragma unroll
      for (int jj = 0; jj < kIn; ++jj) x0[jj] = (__bf16)input_value(A.xnorm, jj, A.x + row * kIn);
    }
    if (w == 0 && A.acts && g == 0 && live) {
#pragma unroll
      for (int jj = 0; jj < 8; ++jj) A.acts[(int64_t)jj * A.ld + row] = x0[jj];
    }
    bf16x8 B[4];
    // layer 1
    f32x4 c0 = mfma(load_frag(A.frags, fF1, a0, lane), x0, bias_tile(A.bias, pB1, a0, g));
    f32x4 c1 = mfma(load_frag(A.frags, fF1, a1, lane), x0, bias_tile(A.bias, pB1, a1, g));
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      c0[i] = tanh_fast(c0[i]);
      c1[i] = tanh_fast(c1[i]);
    }
    const bf16x8 h1w = pack_pair(c0, c1);
    xch[0][w][lane] = h1w;
    __syncthreads();
#pragma unroll
    for (int b = 0; b < 4; ++b) B[b] = xch[0][b][lane];
    // layer 2
    c0 = bias_tile(A.bias, pB2, a0, g);
    c1 = bias_tile(A.bias, pB2, a1, g);
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      c0 = mfma(load_frag(A.frags, fF2, a0 * 4 + b, lane), B[b], c0);
      c1 = mfma(load_frag(A.frags, fF2, a1 * 4 + b, lane), B[b], c1);
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      c0[i] = tanh_fast(c0[i]);
      c1[i] = tanh_fast(c1[i]);
    }
    const bf16x8 h2w = pack_pair(c0, c1);
    xch[1][w][lane] = h2w;
    __syncthreads();
#pragma unroll
    for (int b = 0; b < 4; ++b) B[b] = xch[1][b][lane];
    // layer 3
    c0 = bias_tile(A.bias, pB3, a0, g);
    c1 = bias_tile(A.bias, pB3, a1, g);
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      c0 = mfma(load_frag(A.frags, fF3, a0 * 4 + b, lane), B[b], c0);
      c1 = mfma(load_frag(A.frags, fF3, a1 * 4 + b, lane), B[b], c1);
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      c0[i] = tanh_fast(c0[i]);
      c1[i] = tanh_fast(c1[i]);
    }
    const bf16x8 h3w = pack_pair(c0, c1);
    xch[0][w][lane] = h3w;
    __syncthreads();
#pragma unroll
    for (int b = 0; b < 4; ++b) B[b] = xch[0][b][lane];
    // output layer (every wave forms y and dy of the tile's rows itself)
    f32x4 yo = bias_tile(A.bias, pB4, 0, g);
#pragma unroll
    for (int b = 0; b < 4; ++b) yo = mfma(load_frag(A.frags, fF4, b, lane), B[b], yo);
    float dy = 0.f;
    if (g == 0 && live) {
      const float yv = yo[0];
      if (w == 0 && A.y) A.y[row] = yv;
      if (A.mode == 0) {
        dy = A.target_or_grad[row];
      } else {
        const float e = yv - A.target_or_grad[row];
        const float ae = fabsf(e);
        if (w == 0 && A.loss) A.loss[row] = ae <= A.delta ? 0.5f * e * e : A.delta * (ae - 0.5f * A.delta);
        dy = fminf(fmaxf(e, -A.delta), A.delta) * A.scale;
      }
    }
    if (A.acts && live) {
#pragma unroll
      for (int q = 0; q < 2; ++q)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const int f = 16 * (a0 + q) + 4 * g + i;
          A.acts[(int64_t)(8 + f) * A.ld + row] = h1w[4 * q + i];
          A.acts[(int64_t)(8 + 128 + f) * A.ld + row] = h2w[4 * q + i];
          A.acts[(int64_t)(8 + 256 + f) * A.ld + row] = h3w[4 * q + i];
        }
    }
    const f32x4 dz4 = {g == 0 ? dy : 0.f, 0.f, 0.f, 0.f};
    if (w == 0 && A.deltas && live) {
#pragma unroll
      for (int i = 0; i < 4; ++i) store_fm(A.deltas, A.ld, 384 + 4 * g + i, row, dz4[i]);
    }
    // dz3 = (W4^T dz4) (1 - h3^2)
    const bf16x8 bd4 = pack_pair(dz4, f32x4{0.f, 0.f, 0.f, 0.f});
    c0 = mfma(load_frag(A.frags, fB4, a0, lane), bd4, f32x4{0.f, 0.f, 0.f, 0.f});
    c1 = mfma(load_frag(A.frags, fB4, a1, lane), bd4, f32x4{0.f, 0.f, 0.f, 0.f});
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const float p0 = (float)h3w[i], p1 = (float)h3w[4 + i];
      c0[i] *= 1.0f - p0 * p0;
      c1[i] *= 1.0f - p1 * p1;
    }
    if (A.deltas && live) {
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        store_fm(A.deltas, A.ld, 256 + 16 * a0 + 4 * g + i, row, c0[i]);
        store_fm(A.deltas, A.ld, 256 + 16 * a1 + 4 * g + i, row, c1[i]);
      }
    }
    xch[1][w][lane] = pack_pair(c0, c1);
    __syncthreads();
#pragma unroll
    for (int b = 0; b < 4; ++b) B[b] = xch[1][b][lane];
    // dz2 = (W3^T dz3) (1 - h2^2)
    c0 = f32x4{0.f, 0.f, 0.f, 0.f};
    c1 = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      c0 = mfma(load_frag(A.frags, fB3, a0 * 4 + b, lane), B[b], c0);
      c1 = mfma(load_frag(A.frags, fB3, a1 * 4 + b, lane), B[b], c1);
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const float p0 = (float)h2w[i], p1 = (float)h2w[4 + i];
      c0[i] = c0[i] * (1.0f - p0 * p0);
      c1[i] = c1[i] * (1.0f - p1 * p1);
    }
    if (A.deltas && live) {
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        store_fm(A.deltas, A.ld, 128 + 16 * a0 + 4 * g + i, row, c0[i]);
        store_fm(A.deltas, A.ld, 128 + 16 * a1 + 4 * g + i, row, c1[i]);
      }
    }
    xch[0][w][lane] = pack_pair(c0, c1);
    __syncthreads();
#pragma unroll
    for (int b = 0; b < 4; ++b) B[b] = xch[0][b][lane];
    // dz1 = (W2^T dz2) (1 - h1^2)
    c0 = f32x4{0.f, 0.f, 0.f, 0.f};
    c1 = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      c0 = mfma(load_frag(A.frags, fB2, a0 * 4 + b, lane), B[b], c0);
      c1 = mfma(load_frag(A.frags, fB2, a1 * 4 + b, lane), B[b], c1);
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const float p0 = (float)h1w[i], p1 = (float)h1w[4 + i];
      c0[i] = c0[i] * (1.0f - p0 * p0);
      c1[i] = c1[i] * (1.0f - p1 * p1);
    }
    if (A.deltas && live) {
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        store_fm(A.deltas, A.ld, 16 * a0 + 4 * g + i, row, c0[i]);
        store_fm(A.deltas, A.ld, 16 * a1 + 4 * g + i, row, c1[i]);
      }
    }
    // the next tile's first exchange writes buffer 0, which this tile last read before
    // its final barrier: no further barrier needed here
  }
}

// ---- weight gradients: dW_L = dZ_L^T H_{L-1} over the batch -------------------------------
// Tasks: L1 8x1 tiles, L2 8x8, L3 8x8, L4 1x8 (out tile a, in tile c); one wave per task and
// batch slice: blockIdx.y = s takes rows [s * chunk, min(n, (s + 1) * chunk)) (chunk a multiple
// of 64) and writes its partial sums to grad + s * kNParams (nn.Linear layout, written, not
// accumulated); the Adam kernel adds the slices in order, so the sum is deterministic.
__global__ void __launch_bounds__(64) mlp_wgrad_kernel(const __bf16* __restrict__ acts,
                                                      const __bf16* __restrict__ deltas, int64_t n, int64_t ld,
                                                      int64_t chunk, float* __restrict__ grad) {
  int task = blockIdx.x;
  grad += (int64_t)blockIdx.y * kNParams;
  const int64_t kb = (int64_t)blockIdx.y * chunk;
  const int64_t ke = kb + chunk < n ? kb + chunk : n;
  int layer, a, c;
  if (task < 8) { layer = 1; a = task; c = 0; }
  else if (task < 72) { layer = 2; a = (task - 8) / 8; c = (task - 8) % 8; }
  else if (task < 136) { layer = 3; a = (task - 72) / 8; c = (task - 72) % 8; }
  else { layer = 4; a = 0; c = task - 136; }
  const int lane = threadIdx.x, g = lane >> 4, m = lane & 15;
  const int drow = (layer == 1 ? 0 : layer == 2 ? 128 : layer == 3 ? 256 : 384) + 16 * a + m;  // dZ feature
  const int arow = (layer == 1 ? 0 : 8 + 128 * (layer - 2)) + 16 * c + m;                     // H feature
  const bool in_ok = layer > 1 || m < 8;  // H0 has 8 rows (7 inputs + zero)
  const __bf16* dz = deltas + (int64_t)drow * ld;
  const __bf16* hh = acts + (int64_t)(in_ok ? arow : 0) * ld;
  f32x4 acc0 = {0.f, 0.f, 0.f, 0.f}, acc1 = acc0;
  float bsum = 0.f;
  int64_t k = kb;
  for (; k + 64 <= ke; k += 64) {
    bf16x8 av0 = *reinterpret_cast<const bf16x8*>(dz + k + 8 * g);
    bf16x8 av1 = *reinterpret_cast<const bf16x8*>(dz + k + 32 + 8 * g);
    bf16x8 bv0 = {}, bv1 = {};
    if (in_ok) {
      bv0 = *reinterpret_cast<const bf16x8*>(hh + k + 8 * g);
      bv1 = *reinterpret_cast<const bf16x8*>(hh + k + 32 + 8 * g);
    }
    acc0 = mfma(av0, bv0, acc0);
    acc1 = mfma(av1, bv1, acc1);
    if (c == 0) {
#pragma unroll
      for (int j = 0; j < 8; ++j) bsum += (float)av0[j] + (float)av1[j];
    }
  }
  for (; k < ke; k += 32) {  // tail of the last slice (ld is padded to a multiple of 32 with zeros)
    bf16x8 av = *reinterpret_cast<const bf16x8*>(dz + k + 8 * g);
    bf16x8 bv = {};
    if (in_ok) bv = *reinterpret_cast<const bf16x8*>(hh + k + 8 * g);
    acc0 = mfma(av, bv, acc0);
    if (c == 0) {
#pragma unroll
      for (int j = 0; j < 8; ++j) bsum += (float)av[j];
    }
  }
  const f32x4 acc = acc0 + acc1;
  // D lane holds dW[out 16a + 4g + i][in 16c + m]
  const int nin = layer == 1 ? kIn : layer == 2 ? kH1 : layer == 3 ? kH2 : kH3;
  const int nout = layer == 1 ? kH1 : layer == 2 ? kH2 : layer == 3 ? kH3 : kOut;
  const int ow = layer == 1 ? oW1 : layer == 2 ? oW2 : layer == 3 ? oW3 : oW4;
  const int ob = layer == 1 ? oB1 : layer == 2 ? oB2 : layer == 3 ? oB3 : oB4;
  const int in = 16 * c + m;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int out = 16 * a + 4 * g + i;
    if (out < nout && in < nin) grad[ow + out * nin + in] = acc[i];
  }
  if (c == 0) {
    // bias: sum over the 4 lane groups of the same m
    bsum += __shfl_xor(bsum, 16);
    bsum += __shfl_xor(bsum, 32);
    const int out = 16 * a + m;
    if (g == 0 && out < nout) grad[ob + out] = bsum;
  }
}

// ---- Adam on the FP32 master weights ---------------------------------------------------------
// G holds nsplit gradient slices of n values each, summed here in slice order.
__global__ void mlp_adam_kernel(float* __restrict__ P, const float* __restrict__ G, int nsplit,
                                float* __restrict__ M, float* __restrict__ V, int n, float lr, float b1, float b2,
                                float eps, float bc1, float bc2, float wd) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float gs = G[i];
  for (int s = 1; s < nsplit; ++s) gs += G[(int64_t)s * n + i];
  const float g = gs + wd * P[i];
  const float m = b1 * M[i] + (1.f - b1) * g;
  const float v = b2 * V[i] + (1.f - b2) * g * g;
  M[i] = m;
  V[i] = v;
  P[i] -= lr * (m / bc1) / (sqrtf(v / bc2) + eps);
}

// Adam step fused with the repack: each parameter, once updated, is written straight to
// its forward and backward fragment slots (and its padded bias), the inverse of
// mlp_pack_kernel's mapping, with the same scaling and bf16 rounding. Padding entries are
// never touched (they stay zero from the initial pack).
__device__ __forceinline__ void decode_chain(int f, int& b, int& g, int& j) {
  // inverse of chain_feature: f = 16 (2b + j/4) + 4g + j%4
  const int t = f >> 4, w4 = f & 15;
  b = t >> 1;
  g = w4 >> 2;
  j = ((t & 1) << 2) | (w4 & 3);
}

__global__ void mlp_adam_pack_kernel(float* __restrict__ P, const float* __restrict__ G, int nsplit,
                                     float* __restrict__ M, float* __restrict__ V, float lr, float b1, float b2,
                                     float eps, float bc1, float bc2, float wd, __bf16* __restrict__ frags,
                                     float* __restrict__ bias) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= kNParams) return;
  // issue every slice's load before summing them in slice order
  float part[16];
#pragma unroll
  for (int s = 0; s < 16; ++s) part[s] = s < nsplit ? G[(int64_t)s * kNParams + i] : 0.f;
  float gs = part[0];
#pragma unroll
  for (int s = 1; s < 16; ++s)
    if (s < nsplit) gs += part[s];
  const float gr = gs + wd * P[i];
  const float m = b1 * M[i] + (1.f - b1) * gr;
  const float v = b2 * V[i] + (1.f - b2) * gr * gr;
  M[i] = m;
  V[i] = v;
  const float p = P[i] - lr * (m / bc1) / (sqrtf(v / bc2) + eps);
  P[i] = p;
  int layer, out, in = -1;
  if (i < oB1) { layer = 1; out = i / kIn; in = i % kIn; }
  else if (i < oW2) { layer = 1; out = i - oB1; }
  else if (i < oB2) { layer = 2; out = (i - oW2) / kH1; in = (i - oW2) % kH1; }
  else if (i < oW3) { layer = 2; out = i - oB2; }
  else if (i < oB3) { layer = 3; out = (i - oW3) / kH2; in = (i - oW3) % kH2; }
  else if (i < oW4) { layer = 3; out = i - oB3; }
  else if (i < oB4) { layer = 4; out = (i - oW4) / kH3; in = (i - oW4) % kH3; }
  else { layer = 4; out = i - oB4; }
  const float sc = layer < 4 ? kTanhScale : 1.0f;
  if (in < 0) {
    const int pb = layer == 1 ? pB1 : layer == 2 ? pB2 : layer == 3 ? pB3 : pB4;
    bias[pb + out] = sc * p;
    return;
  }
  // forward fragment: A[m = out][k = in]
  {
    const int base = layer == 1 ? fF1 : layer == 2 ? fF2 : layer == 3 ? fF3 : fF4;
    const int KB = layer == 1 ? 1 : 4;
    int b, g, j;
    if (layer == 1) { b = 0; g = in >> 3; j = in & 7; }
    else decode_chain(in, b, g, j);
    const int a = out >> 4, lane = 16 * g + (out & 15);
    frags[base + ((a * KB + b) * 64 + lane) * 8 + j] = (__bf16)(p * sc);
  }
  // backward fragment: A[m = in][k = out]
  {
    const int base = layer == 1 ? fB1 : layer == 2 ? fB2 : layer == 3 ? fB3 : fB4;
    const int KB = layer == 4 ? 1 : 4;
    int b, g, j;
    decode_chain(out, b, g, j);
    const int a = in >> 4, lane = 16 * g + (in & 15);
    frags[base + ((a * KB + b) * 64 + lane) * 8 + j] = (__bf16)p;
  }
}

// one MFMA with per-lane operands, to pin the operand / result lane layout
__global__ void mfma_probe_kernel(const bf16x8* a, const bf16x8* b, f32x4* d) {
  const int l = threadIdx.x;
  d[l] = mfma(a[l], b[l], f32x4{0.f, 0.f, 0.f, 0.f});
}

// the same for the inference kernel's 32 x 32 x 16 MFMA (result d[64][16])
__global__ void mfma32_probe_kernel(const bf16x8* a, const bf16x8* b, f32x16* d) {
  const int l = threadIdx.x;
  d[l] = mfma32(a[l], b[l], f32x16{});
}

// ---- host -------------------------------------------------------------------------------------
int32_t mfma_probe_enqueue(const void* a, const void* b, float* d, hipStream_t s) {
  if (!a || !b || !d) return fail(SSM_EINVAL, "NULL buffer");
  hipLaunchKernelGGL(mfma_probe_kernel, dim3(1), dim3(64), 0, s, (const bf16x8*)a, (const bf16x8*)b, (f32x4*)d);
  SSM_HIP(hipGetLastError());
  return SSM_OK;
}

int32_t mfma32_probe_enqueue(const void* a, const void* b, float* d, hipStream_t s) {
  if (!a || !b || !d) return fail(SSM_EINVAL, "NULL buffer");
  hipLaunchKernelGGL(mfma32_probe_kernel, dim3(1), dim3(64), 0, s, (const bf16x8*)a, (const bf16x8*)b, (f32x16*)d);
  SSM_HIP(hipGetLastError());
  return SSM_OK;
}

int32_t mlp_sizes(int32_t* n_params, int32_t* n_frag, int32_t* n_bias, int32_t* act_rows, int32_t* delta_rows) {
  if (n_params) *n_params = kNParams;
  if (n_frag) *n_frag = kNFrag;
  if (n_bias) *n_bias = kNBias;
  if (act_rows) *act_rows = kActRows;
  if (delta_rows) *delta_rows = kDeltaRows;
  return SSM_OK;
}

int32_t mlp_pack_enqueue(const float* params, __bf16* frags, float* bias, hipStream_t s) {
  if (!params || !frags || !bias) return fail(SSM_EINVAL, "NULL buffer");
  hipLaunchKernelGGL(mlp_pack_kernel, dim3((kNFrag + 255) / 256), dim3(256), 0, s, params, frags, bias);
  SSM_HIP(hipGetLastError());
  return SSM_OK;
}

int32_t mlp_forward_enqueue(const float* x, int64_t n, const __bf16* frags, const float* bias, const float* xnorm,
                            float* y, hipStream_t s) {
  if (n < 0) return fail(SSM_EINVAL, "n must be >= 0");
  if (n == 0) return SSM_OK;
  if (!x || !frags || !bias || !y) return fail(SSM_EINVAL, "NULL buffer");
  const int cus = device_cu_count();
  if (cus <= 0) return fail(SSM_EHIP, "no HIP device");
  // 1024-thread blocks, one per CU (71 KB of LDS image and biases), 32 rows per wave
  const int64_t waves_needed = (n + 31) / 32;
  int64_t blocks = (waves_needed + 15) / 16;
  const int64_t cap = (int64_t)cus;
  if (blocks > cap) blocks = cap;
  hipLaunchKernelGGL(mlp_forward_kernel, dim3((unsigned)blocks), dim3(1024), 0, s, x, n, frags, bias, xnorm, y);
  SSM_HIP(hipGetLastError());
  return SSM_OK;
}

int32_t mlp_train_enqueue(const TrainArgs& A, hipStream_t s) {
  if (A.n < 0) return fail(SSM_EINVAL, "n must be >= 0");
  if (A.n == 0) return SSM_OK;
  if (!A.x || !A.frags || !A.bias || !A.target_or_grad) return fail(SSM_EINVAL, "NULL buffer");
  if ((A.acts || A.deltas) && (A.ld < A.n || A.ld % 32)) return fail(SSM_EINVAL, "ld must be >= n and a multiple of 32");
  const int cus = device_cu_count();
  if (cus <= 0) return fail(SSM_EHIP, "no HIP device");
  const int64_t tiles = (A.n + 15) / 16;
  if (!A.dx) {
    // four waves per 16-row tile, up to 8 blocks per CU, the rest by grid stride
    const int64_t cap = (int64_t)cus * 8;
    const int64_t blocks = tiles < cap ? tiles : cap;
    hipLaunchKernelGGL(mlp_train_team_kernel, dim3((unsigned)blocks), dim3(256), 0, s, A);
  } else {
    // one-wave blocks (16 rows each) so that a batch of a few thousand rows spreads over
    // every CU; at most 16 waves per CU, the rest by grid stride
    const int64_t cap = (int64_t)cus * 16;
    const int64_t blocks = tiles < cap ? tiles : cap;
    hipLaunchKernelGGL(mlp_train_kernel, dim3((unsigned)blocks), dim3(64), 0, s, A);
  }
  SSM_HIP(hipGetLastError());
  return SSM_OK;
}

int32_t mlp_wgrad_enqueue(const __bf16* acts, const __bf16* deltas, int64_t n, int64_t ld, float* grad,
                          hipStream_t s) {
  if (n < 1 || ld < n || ld % 32) return fail(SSM_EINVAL, "need n >= 1 and ld >= n, ld % 32 == 0");
  if (!acts || !deltas || !grad) return fail(SSM_EINVAL, "NULL buffer");
  hipLaunchKernelGGL(mlp_wgrad_kernel, dim3(144, 1), dim3(64), 0, s, acts, deltas, n, ld, ld, grad);
  SSM_HIP(hipGetLastError());
  return SSM_OK;
}

int32_t mlp_adam_enqueue(float* P, const float* G, float* M, float* V, int32_t step, float lr, float b1, float b2,
                         float eps, float wd, hipStream_t s) {
  if (!P || !G || !M || !V) return fail(SSM_EINVAL, "NULL buffer");
  if (step < 1) return fail(SSM_EINVAL, "step counts from 1");
  const float bc1 = 1.f - powf(b1, (float)step), bc2 = 1.f - powf(b2, (float)step);
  hipLaunchKernelGGL(mlp_adam_kernel, dim3((kNParams + 255) / 256), dim3(256), 0, s, P, G, 1, M, V, kNParams, lr,
                     b1, b2, eps, bc1, bc2, wd);
  SSM_HIP(hipGetLastError());
  return SSM_OK;
}


// ---- a run of Adam steps on consecutive batches, enqueued from the host ----------------------
int64_t mlp_train_split(int64_t batch) {
  int64_t sp = batch / 512;
  return sp < 1 ? 1 : sp > 16 ? 16 : sp;
}
int64_t mlp_train_workspace_bytes(int64_t batch) {
  if (batch < 32 || batch % 32) return -1;
  return (int64_t)(kActRows + kDeltaRows) * batch * 2 + mlp_train_split(batch) * kNParams * 4;
}
int32_t mlp_train_steps_enqueue(const float* x, const float* target, int64_t batch, int32_t n_steps, int32_t step0,
                                const float* lr, float b1, float b2, float eps, float wd, float delta, float* P,
                                float* M, float* V, __bf16* frags, float* bias, const float* xnorm, void* ws,
                                int64_t ws_bytes, float* loss, hipStream_t s) {
  if (batch < 32 || batch % 32) return fail(SSM_EINVAL, "batch must be a positive multiple of 32");
  if (n_steps < 0 || step0 < 0) return fail(SSM_EINVAL, "n_steps and step0 must be >= 0");
  if (n_steps == 0) return SSM_OK;
  if (!x || !target || !lr || !P || !M || !V || !frags || !bias || !ws) return fail(SSM_EINVAL, "NULL buffer");
  if (ws_bytes < mlp_train_workspace_bytes(batch)) return fail(SSM_EINVAL, "workspace too small");
  if ((uintptr_t)ws % 16) return fail(SSM_EINVAL, "workspace must be 16-byte aligned");
  __bf16* acts = (__bf16*)ws;
  __bf16* deltas = acts + (int64_t)kActRows * batch;
  float* grad = (float*)(deltas + (int64_t)kDeltaRows * batch);
  const int64_t split = mlp_train_split(batch);
  const int64_t chunk = ((batch + split - 1) / split + 63) / 64 * 64;
  for (int32_t i = 0; i < n_steps; ++i) {
    const int64_t r0 = (int64_t)i * batch;
    TrainArgs A{x + r0 * kIn, target + r0, frags, bias, xnorm, acts, deltas, nullptr, nullptr,
                loss ? loss + r0 : nullptr, batch, batch, 1, delta, 1.0f / (float)batch};
    int32_t rc = mlp_train_enqueue(A, s);
    if (rc != SSM_OK) return rc;
    hipLaunchKernelGGL(mlp_wgrad_kernel, dim3(144, (unsigned)split), dim3(64), 0, s, acts, deltas, batch, batch,
                       chunk, grad);
    const int32_t step = step0 + i + 1;
    const float bc1 = 1.f - powf(b1, (float)step), bc2 = 1.f - powf(b2, (float)step);
    hipLaunchKernelGGL(mlp_adam_pack_kernel, dim3((kNParams + 255) / 256), dim3(256), 0, s, P, grad, (int)split, M,
                       V, lr[i], b1, b2, eps, bc1, bc2, wd, frags, bias);
    SSM_HIP(hipGetLastError());
  }
  return SSM_OK;
}

}  // namespace ssmhip
