// Persistent Euler–Maruyama first-passage kernel for gfx950.
//
// One trajectory per lane. Each wavefront pulls work from a global queue over
// the flattened (set, sample) index space; lanes whose walk has ended hold
// their result until enough lanes of the wave are idle, then the wave writes
// those results in one store instruction and refills the idle lanes from a
// wave-local index range. The range is topped up by one atomic on the global
// head for up to `queue_block` indices (fewer as the queue drains), so the
// single head address sees one atomic per several refills.  The walk follows the sequential loops of
//   ddm_flexbound  src/cssm/ddm_models.pyx:696-744
//   race_model     src/cssm/race_models.pyx:417-464
//   lca            src/cssm/race_models.pyx:825-879
//   full_ddm, ornstein_uhlenbeck, levy_flexbound (ddm_models.pyx:2284-2358,
//   ornstein_models.pyx:156-198, levy_models.pyx:160-220) on the DDM walk, and
//   racing_diffusion_model (race_models.pyx:1207-1248) on the accumulator walk
// in float32 with the same operation order (this file is compiled with
// -ffp-contract=off so no multiply-add is fused behind our back).
//
// Normals: native mode draws Philox4x32-10 + Box–Muller with counter
// (sample, set, block, stream); replay mode reads a caller buffer.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <map>
#include <mutex>
#include <tuple>
#include <type_traits>
#include <vector>

#include "ssmhip_internal.h"

namespace ssmhip {

struct EmArgs {
  const float* set_params;
  const int64_t* offsets;
  const int64_t* qoff;   // queue order: prefix of the set sizes in the order the queue visits sets
  const int32_t* order;  // queue position -> set (NULL: sets in index order, qoff == offsets)
  const float* table;
  const float* replay_normals;
  const float* replay_uniforms;
  float* rts;
  void* choices;
  float* traj;        // trajectory recording: positions of walk (set 0, sample 0), [traj_rows][n_particles]
  unsigned long long* ws;  // [0] queue head, [1] executed steps, [2] lane-step slots issued
  int64_t total;
  int64_t replay_stride;
  int32_t set_stride;  // floats per set record
  int32_t n_sets;
  int32_t n_steps;
  int32_t n_particles;
  int32_t smooth_unif;
  int32_t out_layout;
  int32_t choice_bytes;
  int32_t refill_min;
  int32_t queue_block;  // largest index range a wave claims from the global queue at once
  int32_t traj_rows;
  float dt;
  // t_fast[r] (r = 1 in replay mode): an upper bound on how far t can rise over the steps
  // of one walk-loop iteration; t_fast[2] the same over a long tail iteration (native mode);
  // NaN when n_steps > 2^24 (then no iteration runs fast)
  float t_fast[3];
  // t_grid[k]: t after k steps (k float32 additions of dt from 0), k = 0..n_steps; the
  // untested iterations read a walk's t from it instead of adding dt per step (NULL when
  // n_steps > 2^24, where no iteration is untested)
  const float* t_grid;
  // most walks a wave runs at once (64: every lane); fewer lets a small job's walks take the
  // short-latency tail modes from their first step
  int32_t max_walks;
  uint32_t k0, k1;
};

enum : int { kRunning = 0, kDone = 1, kEmpty = 2 };

// Boundary height at step index ix (t_s[ix] = float32(ix) * dt exactly, as
// np.arange(0, max_t + dt, dt).astype(float32) in _utils.pyx:636-640).
// fx is float32(ix), exact for ix < 2^24; the table bound reads row[ix].
template <int BOUND>
__device__ __forceinline__ float boundary_at_f(float fx, int ix, float dt, float a, float b1, float b2,
                                               const float* __restrict__ row) {
  if constexpr (BOUND == SSM_BOUNDARY_CONSTANT) {
    return a;
  } else if constexpr (BOUND == SSM_BOUNDARY_ANGLE) {
    // boundary_functions.py:57-72: a + t * (-sin(theta)/cos(theta)), slope precomputed in float32
    const float ts = fx * dt;
    return a + ts * b1;
  } else if constexpr (BOUND == SSM_BOUNDARY_WEIBULL) {
    // boundary_functions.py:103-122: a * exp(-(t/beta)^alpha); b1 = alpha, b2 = 1/beta
    const float ts = fx * dt;
    const float u = ts * b2;
    const float p = __builtin_amdgcn_exp2f(b1 * __builtin_amdgcn_logf(u));  // u == 0 -> 0
    return a * __builtin_amdgcn_exp2f(-1.4426950408889634f * p);
  } else {
    return row[ix];
  }
}
template <int BOUND>
__device__ __forceinline__ float boundary_at(int ix, float dt, float a, float b1, float b2,
                                             const float* __restrict__ row) {
  return boundary_at_f<BOUND>((float)ix, ix, dt, a, b1, b2, row);
}

__device__ __forceinline__ int find_set(const int64_t* __restrict__ off, int n_sets, int64_t g) {
  // largest k with off[k] <= g (off[n_sets] > g); sets with zero samples are skipped
  int lo = 0, hi = n_sets - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (off[mid] <= g) lo = mid; else hi = mid - 1;
  }
  return lo;
}

// Position increment of one step of the single-particle diffusions, in the reference's order:
//   DDM, full DDM   (v dt) + (sqrt_st g)             ddm_models.pyx:720, :2335
//   Ornstein        ((v - g y) dt) + (sqrt_st g)     ornstein_models.pyx:173
//   Levy            (v dt) + (dt_alpha x)            levy_models.pyx:192, x alpha-stable
template <int MODEL>
__device__ __forceinline__ float step_increment(float y, float vdt, float v, float leak, float dt, float sst,
                                                float nz) {
  if constexpr (MODEL == SSM_MODEL_ORNSTEIN) return ((v - (leak * y)) * dt) + (sst * nz);
  else return vdt + sst * nz;
}
// Native walks of the Gaussian diffusions draw their normals already scaled by sqrt_st
// (Box–Muller radius sqrt(-2 ln u1 sqrt_st^2)): the same law as sqrt_st g, one multiply
// less per step. Replay walks take the reference's product of the buffered draw.
template <int MODEL, bool REPLAY>
constexpr bool kScaledNormals = !REPLAY && MODEL != SSM_MODEL_LEVY;
template <int MODEL, bool REPLAY>
__device__ __forceinline__ float walk_increment(float y, float vdt, float v, float leak, float dt, float sst,
                                                float nz) {
  if constexpr (kScaledNormals<MODEL, REPLAY>) {
    if constexpr (MODEL == SSM_MODEL_ORNSTEIN) return ((v - (leak * y)) * dt) + nz;
    else return vdt + nz;
  } else {
    return step_increment<MODEL>(y, vdt, v, leak, dt, sst, nz);
  }
}

// ---- float32 forms of the reference's double-precision expressions ----------------
// A few reference lines mix float operands with C double literals (0.5, 2.0) or libc's
// double fmax, so C evaluates them in double and rounds to float on assignment. The
// operands are floats, so the exact value is a short sum of float products; it is formed
// here without FP64 by error-free transformations and rounded once. This equals the double
// expression whenever its double intermediates are exact, which holds for the operand
// ranges of these lines except in cases where rounding to double and then to float lands
// differently from rounding once (relative probability ~2^-29 per value).

// Knuth's TwoSum: s + e == a + b exactly (needs -ffp-contract=off, as this file is built).
__device__ __forceinline__ float two_sum(float a, float b, float& e) {
  const float s = a + b;
  const float bb = s - a;
  e = (a - (s - bb)) + (b - bb);
  return s;
}

// y + 2 (u - 0.5) s, the full DDM's start and non-decision-time jitter
// (ddm_models.pyx:2295-2302, float y, u, s and double literals).
__device__ __forceinline__ float add_unit_shift(float y, float u, float s) {
  float wl;
  const float wh = two_sum(u + u, -1.0f, wl);                 // 2u - 1 == wh + wl exactly
  const float ph = wh * s;
  const float pl = __builtin_fmaf(wh, s, -ph) + wl * s;       // wh s - ph exactly, plus wl s
  float e;
  const float r = two_sum(y, ph, e);
  return r + (e + pl);
}

// compute_smooth_unif (_utils.pyx:668-695): u 0.5 dt at t == 0 (u 0.5 is exact and the float
// product of two floats is the rounded exact product, as in double), (0.5 - u) dt before the
// deadline (one rounding of the exact value by the FMA), else 0.
__device__ __forceinline__ float smooth_unif_shift(float u, float t, float dl_tmp, float dt) {
  if (t == 0.f) return (u * 0.5f) * dt;
  if (t < dl_tmp) return __builtin_fmaf(-u, dt, 0.5f * dt);
  return 0.f;
}

// a + b + c rounded once (parallel_models.pyx:538: fmax(t_h, t_l) + t + smooth_u is a double sum)
__device__ __forceinline__ float sum3_rounded(float a, float b, float c) {
  float e1, e2;
  const float s1 = two_sum(a, b, e1);
  const float s2 = two_sum(s1, c, e2);
  return s2 + (e1 + e2);
}

// the single-particle diffusions that share the DDM walk
template <int MODEL>
constexpr bool kDiffusion = MODEL == SSM_MODEL_DDM || MODEL == SSM_MODEL_FULL_DDM || MODEL == SSM_MODEL_ORNSTEIN ||
                            MODEL == SSM_MODEL_LEVY || MODEL == SSM_MODEL_PAR2 || MODEL == SSM_MODEL_SEQ2 ||
                            MODEL == SSM_MODEL_DDM_LEGACY;

// The walk continues while y lies inside the bounds: [-b, b], or [0, b] for the legacy DDM
// (ddm_models.pyx:463: y <= a and y >= 0), both inclusive. -b <= y <= b is tested as
// |y| <= b, one compare with an abs source modifier; the two agree for every float y and b,
// signed zeros, infinities and NaN (false either way) included.
template <int MODEL>
__device__ __forceinline__ bool inside(float y, float b) {
  if constexpr (MODEL == SSM_MODEL_DDM_LEGACY) return (y <= b) && (y >= 0.f);
  else return fabsf(y) <= b;
}
// Lane shifts read the source lane whether or not this lane's result is used, so every call
// sits where all 64 lanes are active: DPP treats an inactive source lane as out of range.
// Lane i of each row of 16 lanes reads lane i - 1 of its row; lane 0 of a row reads 0
// (DPP row_shr:1 with bound_ctrl; the compiler folds it into the add that consumes it).
__device__ __forceinline__ float row_shr1(float v) {
  return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0x111, 0xf, 0xf, true));
}
// Lane i of each row of 16 lanes reads lane i - 1 of its row; lane 0 of a row keeps `old`.
__device__ __forceinline__ float row_shr1_or(float old, float v) {
  return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(__builtin_bit_cast(int, old), __builtin_bit_cast(int, v),
                                                               0x111, 0xf, 0xf, false));
}
// Every lane of a row of 16 lanes reads lane 7 of its row (DPP row_newbcast:7).
__device__ __forceinline__ float row_bcast7(float v) {
  return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0x157, 0xf, 0xf, false));
}
// Keeps a lane-shift result where it is formed: without it the compiler may move the shift
// into the branch of a select that uses it, where only some lanes are active.
__device__ __forceinline__ void dpp_fence(float& v) { asm volatile("" : "+v"(v)); }
// Every lane of a row of 16 lanes reads lane 15 of its row (DPP row_newbcast:15).
__device__ __forceinline__ float row_bcast15(float v) {
  return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0x15F, 0xf, 0xf, false));
}
// hierarchical two-choice DDMs: a high walk, then a low walk chosen by its outcome
template <int MODEL>
constexpr bool kTwoStage = MODEL == SSM_MODEL_PAR2 || MODEL == SSM_MODEL_SEQ2;

// Work per iteration of the walk loop, whose own work (the state ballots and the refill
// test) is then paid once for it: blocks of four steps for the diffusions, single steps
// (one Philox block each) for up to 4 accumulators. More accumulators keep one step per
// iteration (two would cost them a wave per SIMD in registers), and replay mode keeps
// few, since its gathers would otherwise multiply the code size.
template <bool REPLAY>
constexpr int kSubBlocks = REPLAY ? 2 : 8;
template <bool REPLAY, int MAXP>
constexpr int kAccSteps = (REPLAY || MAXP > 4) ? 1 : 8;
// Trajectory recording (RECORD) runs one walk; its kernels keep the shortest loop bodies.
template <bool REPLAY, bool RECORD>
constexpr int kSubBlocksR = RECORD ? 1 : kSubBlocks<REPLAY>;
template <bool REPLAY, int MAXP, bool RECORD>
constexpr int kAccStepsR = RECORD ? 1 : kAccSteps<REPLAY, MAXP>;
// Native single-stage Gaussian diffusions share out the Philox blocks of their last walks
// across the wave's idle lanes once at most kHelperWalks walks are left in it.
template <int MODEL, bool REPLAY, bool RECORD>
constexpr bool kHelperTail = !REPLAY && !RECORD &&
                             (MODEL == SSM_MODEL_DDM || MODEL == SSM_MODEL_FULL_DDM ||
                              MODEL == SSM_MODEL_ORNSTEIN || MODEL == SSM_MODEL_DDM_LEGACY);
constexpr int kHelperWalks = 32;
// With at most kLoneWalks walks left, each walk takes a group of 16 or 8 lanes, which form
// its positions and times that many steps at a time; one iteration takes at most kLoneSteps
// steps.
constexpr int kLoneWalks = 8;
constexpr int kLoneSteps = 64;
// walks per wave of a small job (at most 16 walks per SIMD): within reach of the chains;
// of a job of at most 40 walks per SIMD: within reach of the shared Philox blocks
constexpr int kSmallJobWalks = 4;
constexpr int kMidJobWalks = 16;

// Choice rule at a two-stage walk's exit (parallel_models.pyx:474-479, :547-552): a fair coin
// when the bound has collapsed (b <= 0), else u <= (y + b) / (2 b) in float32.
__device__ __forceinline__ bool two_stage_coin(float u, float y, float b) {
  return b <= 0.f ? u <= 0.5f : u <= (y + b) / (2.0f * b);
}

// Symmetric alpha-stable variate by Chambers, Mallows & Stuck (1976), the
// representation the reference samples with (draw_random_stable, _utils.pyx:369-470):
//   X = sin(alpha U) / |cos U|^(1/alpha) * (|cos((1 - alpha) U)| / W)^((1 - alpha)/alpha)
// with U ~ U(-pi/2 + 1e-7, pi/2 - 1e-7), W ~ Exp(1) floored at 1e-10, the result
// clipped to +-1e10 and NaN mapped to 0. Angles are kept in revolutions for
// v_sin_f32; cos y is taken as sin(1/4 - |y|) so that it keeps its relative
// accuracy where it approaches 0 (|U| -> pi/2 is the heavy tail). alpha == 1
// gives tan U through the same formula (the second power is x^0 = 1).
__device__ __forceinline__ float stable_cms(uint32_t bu, uint32_t bw, float alpha, float alpha_inv, float scale) {
  const float urev = (u32_to_unit(bu) - 0.5f) * 0.49999996817f;  // U / (2 pi), |U| <= pi/2 - 1e-7
  const float w = fmaxf(-0.69314718055994531f * __builtin_amdgcn_logf(u32_to_open01(bw)), 1e-10f);
  const float cu = fmaxf(__builtin_amdgcn_sinf(0.25f - fabsf(urev)), 1e-10f);                 // |cos U|
  const float ca = fmaxf(__builtin_amdgcn_sinf(0.25f - fabsf((1.0f - alpha) * urev)), 1e-10f);  // |cos((1-alpha)U)|
  const float lg = scale * (__builtin_amdgcn_logf(ca) - __builtin_amdgcn_logf(w)) - alpha_inv * __builtin_amdgcn_logf(cu);
  float x = __builtin_amdgcn_sinf(alpha * urev) * __builtin_amdgcn_exp2f(lg);
  x = fminf(fmaxf(x, -1e10f), 1e10f);
  return x == x ? x : 0.0f;
}

// RECORD: the launch runs walk (set 0, sample 0) only and stores its positions, start included,
// after every step as the reference's sequential loops do for metadata['trajectory']
// (ddm_models.pyx:714-716, :726-728; race_models.pyx:433-436, :449-452, :1223-1226, :1250-1254);
// it writes no rts or choices.
template <int MODEL, int BOUND, bool REPLAY, int MAXP, bool RECORD>
__global__ void __launch_bounds__(1024) em_kernel(EmArgs A) {
  const int P = kDiffusion<MODEL> ? 1 : A.n_particles;
  const float dt = A.dt;

  int state = kEmpty;
  int64_t g = 0;
  int k = 0, n = 0;
  float x[MAXP];
  float t = 0.f;
  int ix = 0;
  // per-trajectory constants
  float vdt[MAXP], sst[MAXP], vraw[MAXP];
  float bmc = kBoxMullerLog2;  // native Gaussian diffusions: -2 ln 2 sqrt_st^2 (normals scaled by sqrt_st)
  float a = 0.f, b1 = 0.f, b2 = 0.f, t0 = 0.f, deadline = 0.f, dl_tmp = 0.f, leak = 0.f, inh = 0.f;
  float leak_dt = 0.f, inh_dt = 0.f;  // native LCA walks: g dt and b dt
  float alpha = 2.f, alpha_inv = 0.5f, alpha_scale = -0.5f;  // Levy stability and its two exponents
  int win = -1;  // racing diffusion: first accumulator past the threshold in its last step
  // two-stage models: stage (0 high, 1 low), choice so far, low choice already taken, the
  // high walk's time, the boundary index the current stage started at, the high walk's draw
  // count (the low walk's replay draws follow it), the four per-trajectory uniforms
  int stage = 0, ch2 = 0, ix0 = 0, nh = 0;
  bool decided = false;
  float t_h = 0.f, vl1 = 0.f, vl2 = 0.f, zl1 = 0.f, zl2 = 0.f, max_t = 0.f;
  float uu[4] = {0.f, 0.f, 0.f, 0.f};
  const float* row = nullptr;
  unsigned long long my_steps = 0;
  unsigned long long iters = 0;  // wave-uniform: step iterations issued
  bool exhausted = false;        // wave-uniform: no work left for this wave
  bool queue_done = false;       // wave-uniform: the global queue is drained
  int64_t q_lo = 0, q_hi = 0;    // wave-uniform local range [q_lo, q_hi) of unclaimed indices
  int q_k = 0;                   // wave-uniform: set of q_lo (valid while q_lo < q_hi)
  const int64_t waves_total = ((int64_t)gridDim.x * blockDim.x) >> 6;

#pragma unroll
  for (int j = 0; j < MAXP; ++j) { x[j] = 0.f; vdt[j] = 0.f; sst[j] = 0.f; vraw[j] = 0.f; }

  while (true) {
    // ---------------- flush finished lanes / refill idle lanes ----------------
    const uint64_t idle = __ballot(state != kRunning);
    const int n_idle = __popcll(idle);
    bool maint;
    if (exhausted) maint = __ballot(state == kDone) != 0;
    else if (A.max_walks < 64) maint = 64 - n_idle < A.max_walks;  // room for another walk
    else maint = n_idle >= A.refill_min;
    if (maint) {
      if (state == kDone) {
        // result of the finished walk (ddm_models.pyx:735-744, race_models.pyx:454-464)
        float su = 0.f;
        const float t_end = MODEL == SSM_MODEL_PAR2 ? fmaxf(t_h, t) : t;  // par2: rt from max(t_h, t_l)
        if (A.smooth_unif) {
          float u;
          if constexpr (kTwoStage<MODEL>) {
            u = uu[2];
          } else if constexpr (REPLAY) {
            // full DDM rows carry (u_t, u_z, u_smooth), racing-diffusion rows (u_start..., u_smooth);
            // the others one value per trajectory
            if constexpr (MODEL == SSM_MODEL_FULL_DDM) u = A.replay_uniforms[g * 3 + 2];
            else if constexpr (MODEL == SSM_MODEL_RDM) u = A.replay_uniforms[g * (P + 1) + P];
            else u = A.replay_uniforms[g];
          } else {
            u32x4 r = philox4x32_10(u32x4{(uint32_t)n, (uint32_t)k, 0u, 1u}, A.k0, A.k1);
            u = u32_to_unit(r.x);
          }
          su = smooth_unif_shift(u, t_end, dl_tmp, dt);
        }
        float rt;
        if constexpr (MODEL == SSM_MODEL_PAR2)
          rt = sum3_rounded(t_end, t0, su);
        else
          rt = t + t0 + su;
        if (rt >= deadline || deadline <= 0.f) rt = -999.f;  // enforce_deadline, _utils.pyx:697-710
        int choice;
        if constexpr (kTwoStage<MODEL>) {
          // the low choice from the low walk's exit unless it was taken at the transition
          if (!decided) {
            const float bx = ix < A.n_steps ? boundary_at<BOUND>(ix, dt, a, b1, b2, row) : 0.f;
            ch2 += two_stage_coin(uu[3], x[0], bx) ? 1 : 0;
          }
          choice = ch2;
        } else if constexpr (kDiffusion<MODEL>) {
          choice = (x[0] > 0.f) - (x[0] < 0.f);  // sign(), _utils.pyx:474-484
        } else if constexpr (MODEL == SSM_MODEL_RDM) {
          // race_models.pyx:1240-1248: no winner or past the deadline -> (-999, -1)
          choice = win;
          if (rt == -999.f || win < 0) { rt = -999.f; choice = -1; }
        } else {
          choice = 0;  // np.argmax: first maximum
          float best = x[0];
#pragma unroll
          for (int j = 1; j < MAXP; ++j)
            if (j < P && x[j] > best) { best = x[j]; choice = j; }
        }
        const int64_t o = A.out_layout == 0 ? g : (int64_t)n * A.n_sets + k;
        if (!RECORD && o < A.total) {  // layout 1 assumes uniform offsets; never write past the buffers
          A.rts[o] = rt;
          if constexpr (MODEL == SSM_MODEL_PAR2) {  // rts_high, rts_low (parallel_models.pyx:539-540)
            A.rts[A.total + o] = t_h + t0;
            A.rts[2 * A.total + o] = t + t0;
          }
          if (A.choice_bytes == 1) ((int8_t*)A.choices)[o] = (int8_t)choice;
          else ((int32_t*)A.choices)[o] = choice;
        }
        if constexpr (kTwoStage<MODEL>) my_steps += (unsigned long long)(nh + ix - ix0);
        else my_steps += (unsigned long long)ix;
        if constexpr (MODEL == SSM_MODEL_RDM) my_steps += win >= 0;  // the winning step leaves t, ix as they were
        state = kEmpty;
      }
      if (!exhausted) {
        uint64_t want = __ballot(state == kEmpty);
        if (A.max_walks < 64) {  // the first empty lanes up to the cap take work; the rest stay empty
          const int room = A.max_walks - __popcll(__ballot(state == kRunning));
          want = __ballot(state == kEmpty && lanes_below(want) < room);
        }
        const int cnt = __popcll(want);
        // the first r ranked lanes take the wave's local range, the rest a fresh block
        const int64_t lo_old = q_lo;
        const int k_old = q_k;
        const int64_t r = q_hi - q_lo;
        const int64_t from_local = r < cnt ? r : cnt;
        int64_t nb_base = 0, nb_end = 0;
        int nb_k = 0;
        if (r < cnt && !queue_done) {
          // block size: queue_block while the queue is long, shrinking to one refill near its end
          int64_t blk = (int64_t)A.queue_block;
          const int64_t left = A.total - q_hi;
          if (left / (waves_total * 8) < blk) blk = left / (waves_total * 8);
          if (blk < cnt - from_local) blk = cnt - from_local;
          const int leader = __ffsll((unsigned long long)want) - 1;
          unsigned long long b = 0;
          if (lane_id() == leader) b = atomicAdd(A.ws, (unsigned long long)blk);
          nb_base = (int64_t)readlane_u64(b, leader);
          if (nb_base + blk >= A.total) queue_done = true;
          nb_end = nb_base + blk < A.total ? nb_base + blk : A.total;
          if (nb_base < nb_end) nb_k = find_set(A.qoff, A.n_sets, nb_base);
        }
        const int64_t need_new = cnt - from_local;
        if (r >= cnt) {
          q_lo = lo_old + cnt;
        } else if (nb_base < nb_end) {
          q_lo = nb_base + need_new < nb_end ? nb_base + need_new : nb_end;
          q_hi = nb_end;
          q_k = nb_k;
        } else {
          q_lo = q_hi;
        }
        if (q_lo < q_hi)
          while (q_k < A.n_sets - 1 && A.qoff[q_k + 1] <= q_lo) ++q_k;  // keep q_k = queue position of q_lo
        if (q_lo >= q_hi && queue_done) exhausted = true;
        if ((want >> lane_id()) & 1ull) {
          const int j = lanes_below(want);
          int64_t gi = A.total;
          int kk = 0;
          if (j < from_local) {
            gi = lo_old + j;
            kk = k_old;
          } else if (nb_base + (j - from_local) < nb_end) {
            gi = nb_base + (j - from_local);
            kk = nb_k;
          }
          if (gi < A.total) {
            int jq = kk;  // queue position of gi
            while (jq < A.n_sets - 1 && A.qoff[jq + 1] <= gi) ++jq;
            n = (int)(gi - A.qoff[jq]);
            k = A.order ? A.order[jq] : jq;
            g = A.order ? A.offsets[k] + n : gi;  // the trajectory's index in the set-major outputs
            const float* sp = A.set_params + (int64_t)k * A.set_stride;
            t0 = sp[SSM_P_T0];
            deadline = sp[SSM_P_DEADLINE];
            dl_tmp = sp[SSM_P_DEADLINE_TMP];
            a = sp[SSM_P_A];
            b1 = sp[SSM_P_B1];
            b2 = sp[SSM_P_B2];
            if constexpr (BOUND == SSM_BOUNDARY_WEIBULL) b2 = 1.0f / b2;
            if constexpr (BOUND == SSM_BOUNDARY_TABLE) row = A.table + (int64_t)k * A.n_steps;
            const float b0 = boundary_at<BOUND>(0, dt, a, b1, b2, row);
            if constexpr (kDiffusion<MODEL>) {
              const float z = sp[SSM_P_Z];
              vdt[0] = sp[SSM_P_V] * dt;
              sst[0] = sp[SSM_P_SQRT_ST];
              if constexpr (kScaledNormals<MODEL, REPLAY>) bmc = kBoxMullerLog2 * (sst[0] * sst[0]);
              if constexpr (MODEL == SSM_MODEL_DDM_LEGACY) {
                x[0] = z * b0;  // ddm_models.pyx:457: start z a, in float
              } else if constexpr (MODEL == SSM_MODEL_FULL_DDM) {
                // ddm_models.pyx:2305: the literal 2.0 is a C double, so the sum is the exact
                // -b0 + 2 z b0 rounded to float once: one FMA (2 z is exact)
                x[0] = __builtin_fmaf(z * 2.0f, b0, -b0);
              } else {
                // ddm_models.pyx:709 (and ornstein_models.pyx:165, levy_models.pyx:183,
                // parallel_models.pyx:452, sequential_models.pyx:460): the literal 2 is an int,
                // so C evaluates (-b0) + ((z 2) b0) in float with two roundings
                x[0] = (-b0) + ((z * 2.0f) * b0);
              }
              if constexpr (MODEL == SSM_MODEL_ORNSTEIN || MODEL == SSM_MODEL_LEVY || kTwoStage<MODEL>) {
                leak = sp[SSM_P_G];
                vraw[0] = sp[SSM_P_V];
                if constexpr (MODEL == SSM_MODEL_LEVY) {
                  alpha = sp[SSM_P_ALPHA];
                  alpha_inv = 1.0f / alpha;
                  alpha_scale = (1.0f - alpha) / alpha;
                }
                if constexpr (kTwoStage<MODEL>) {
                  stage = 0; ch2 = 0; ix0 = 0; nh = 0; decided = false; t_h = 0.f;
                  vl1 = sp[SSM_P_VJ + 0]; vl2 = sp[SSM_P_VJ + 1];
                  zl1 = sp[SSM_P_ZJ + 0]; zl2 = sp[SSM_P_ZJ + 1];
                  max_t = sp[SSM_P_MAX_T];
                  if constexpr (REPLAY) {
#pragma unroll
                    for (int i = 0; i < 4; ++i) uu[i] = A.replay_uniforms[g * 4 + i];
                  } else {
                    const u32x4 r = philox4x32_10(u32x4{(uint32_t)n, (uint32_t)k, 0u, 5u}, A.k0, A.k1);
                    uu[0] = u32_to_unit(r.x); uu[1] = u32_to_unit(r.y);
                    uu[2] = u32_to_unit(r.z); uu[3] = u32_to_unit(r.w);
                  }
                }
              }
              if constexpr (MODEL == SSM_MODEL_FULL_DDM) {
                // trial-to-trial variability, ddm_models.pyx:2294-2313: drift v + sv g,
                // non-decision time t + 2 (u_t - 1/2) st, start + 2 (u_z - 1/2) sz; the
                // uniform terms are double expressions in the reference (0.5 is a C double)
                float gv, ut, uz;
                if constexpr (REPLAY) {
                  gv = A.replay_normals[g * A.replay_stride];
                  ut = A.replay_uniforms[g * 3 + 0];
                  uz = A.replay_uniforms[g * 3 + 1];
                } else {
                  const u32x4 r = philox4x32_10(u32x4{(uint32_t)n, (uint32_t)k, 0u, 3u}, A.k0, A.k1);
                  float z0, z1;
                  box_muller(r.x, r.y, z0, z1);
                  gv = z0;
                  ut = u32_to_unit(r.z);
                  uz = u32_to_unit(r.w);
                }
                vdt[0] = (sp[SSM_P_V] + sp[SSM_P_SV] * gv) * dt;
                t0 = add_unit_shift(t0, ut, sp[SSM_P_ST]);
                x[0] = add_unit_shift(x[0], uz, sp[SSM_P_SZ]);
              }
            } else {
              leak = sp[SSM_P_G];
              inh = sp[SSM_P_BINH];
              leak_dt = leak * dt;
              inh_dt = inh * dt;
#pragma unroll
              for (int j = 0; j < MAXP; ++j) {
                if (j < P) {
                  vraw[j] = sp[SSM_P_VJ + j];
                  vdt[j] = vraw[j] * dt;
                  sst[j] = sp[SSM_P_SSTJ + j];
                  x[j] = sp[SSM_P_ZJ + j] * b0;  // race_models.pyx:420-421
                } else {
                  vraw[j] = 0.f; vdt[j] = 0.f; sst[j] = 0.f; x[j] = -1.f;
                }
              }
              if constexpr (MODEL == SSM_MODEL_RDM) {
                // race_models.pyx:1207-1208: start u_j * A with u_j uniform on [0, 1]
                win = -1;
#pragma unroll
                for (int q = 0; q < (MAXP + 3) / 4; ++q) {
                  if (4 * q < P) {
                    float u4[4];
                    if constexpr (REPLAY) {
#pragma unroll
                      for (int i = 0; i < 4; ++i)
                        u4[i] = 4 * q + i < P ? A.replay_uniforms[g * (P + 1) + 4 * q + i] : 0.f;
                    } else {
                      const u32x4 r = philox4x32_10(u32x4{(uint32_t)n, (uint32_t)k, (uint32_t)q, 3u}, A.k0, A.k1);
                      u4[0] = u32_to_unit(r.x); u4[1] = u32_to_unit(r.y);
                      u4[2] = u32_to_unit(r.z); u4[3] = u32_to_unit(r.w);
                    }
#pragma unroll
                    for (int i = 0; i < 4; ++i)
                      if (4 * q + i < P) x[4 * q + i] = u4[i] * sp[SSM_P_ZJ + 4 * q + i];
                  }
                }
              }
            }
            t = 0.f;
            ix = 0;
            state = kRunning;
            if constexpr (RECORD) {
              if (g == 0 && A.traj_rows > 0) {
#pragma unroll
                for (int j = 0; j < MAXP; ++j)
                  if (j < P) A.traj[j] = x[j];
              }
            }
          }
        }
      }
    }
    if (__ballot(state == kRunning) == 0) {
      if (exhausted && __ballot(state == kDone) == 0) break;
      continue;
    }

    // ---------------- advance running lanes ----------------
    ++iters;
    if constexpr (kDiffusion<MODEL>) {
      // up to kSubBlocks blocks of four steps per loop iteration, one Philox block each;
      // every walk starts block-aligned, so the draws do not depend on the grouping
      bool alive = state == kRunning;
      // An iteration in which no walking lane can reach the end of t_s (ix + all its steps
      // <= n_steps) or its deadline (t + t_fast <= dl_tmp, t_fast bounding the rise of t over
      // those steps) needs neither the index guard nor the t <= dl_tmp test: both hold at
      // every step it can take. It also counts the steps in float32, exact below 2^24, so
      // that the boundary time needs no conversion per step.
      bool fast = false;
      if constexpr (!RECORD)
        fast = __ballot(alive && !(ix + 4 * kSubBlocks<REPLAY> <= A.n_steps &&
                                   t + A.t_fast[REPLAY ? 1 : 0] <= dl_tmp)) == 0;
      if (fast) {
        // fx counts the steps as float32(ix): it rises by one per step taken and stays at the
        // index of the step whose test ended the walk
        float fx = (float)ix;
        bool helped = false;
        if constexpr (kHelperTail<MODEL, REPLAY, RECORD>) {
          const uint64_t run = __ballot(alive);
          const int K = __popcll(run);  // >= 1: some lane is walking
          if (K <= kHelperWalks) {
            // Few walks left in the wave (the end of the queue): every lane computes one
            // Philox block of one walk, lane = rank + kp * j for block j of the walk of that
            // rank (kp = K rounded up to a power of two), and each walking lane reads its
            // normals from those lanes. A walk's 32 steps then wait for one to four Philox
            // passes instead of eight; up to 8 walks take their steps from chains of lane
            // shifts. Same counters (sample, set, step / 4), same draws.
            helped = true;
            const int lane = lane_id();
            const int r = __popcll(run & ((1ull << lane) - 1ull));  // walking lanes below this one
            // at most kLoneWalks walks left, far enough from the end of t_s and every deadline
            const bool lone = K <= kLoneWalks &&
                              __ballot(alive && !(ix + kLoneSteps <= A.n_steps && t + A.t_fast[2] <= dl_tmp)) == 0;
            // rank -> lane: walking lanes send their lane id to lane r, the others fill K..63
            const int rank_lane = __builtin_amdgcn_ds_permute((alive ? r : K + lane - r) << 2, lane);
            // Chains of lane shifts: each walk takes a group of G lanes (G = 16 or 8 for up to 4
            // or 8 walks) and steps G steps per chunk, lane i holding the position after i steps
            // of the chunk. It comes from a chain y(i) = y(i - 1) + e(i), e(i) the increment of
            // step i - 1, y(0) the chunk's start: after G - 1 links every lane's value is formed
            // by the same float additions, in the same order, as the walk's own steps. The G tests then run at once and the walk keeps the state at
            // its first failing step. One Philox pass (block j of rank w in lane j (64 / G) + w)
            // serves four chunks; the iteration ends after the chunk in which a walk ends.
            // A group's first lane: for G = 16 the row shift reads 0 there, and 0 + y == y but
            // for y = -0, whose sign never reaches a result; for G = 8 a select restores it
            // after every link (the row shift reads across the group boundary at lane 8).
            auto chain_tail = [&](auto gtag) {
              constexpr int G = decltype(gtag)::value;
              constexpr int NW = 64 / G;  // walks one pass serves
              const int gi = lane & (G - 1), grp = lane / G;
              const int pw = lane % NW;
              const int ol = __shfl(rank_lane, pw < K ? pw : 0);
              const uint32_t on = (uint32_t)__shfl(n, ol), ok = (uint32_t)__shfl(k, ol);
              const uint32_t oblk = (uint32_t)__shfl(ix, ol) >> 2;  // walks are block-aligned
              const float obmc = __shfl(bmc, ol);
              float hz[4];
              philox_normals4_c(on, ok, oblk + (uint32_t)(lane / NW), 0u, A.k0, A.k1, obmc, hz);
              // the walk of this group (groups past the last walk repeat walk 0 and never end)
              const bool gwalk = grp < K;
              const int wl = __shfl(rank_lane, gwalk ? grp : 0);
              float Y = __shfl(x[0], wl);
              const float F = __shfl(fx, wl);
              const float gvdt = __shfl(vdt[0], wl);
              float gv = 0.f, gleak = 0.f;
              if constexpr (MODEL == SSM_MODEL_ORNSTEIN) {
                gv = __shfl(vraw[0], wl);
                gleak = __shfl(leak, wl);
              }
              const float ga = __shfl(a, wl), gb1 = __shfl(b1, wl), gb2 = __shfl(b2, wl);
              const float* grow = row;
              if constexpr (BOUND == SSM_BOUNDARY_TABLE)
                grow = (const float*)(uintptr_t)(((uint64_t)(uint32_t)__shfl((int)(uint32_t)(uintptr_t)row, wl)) |
                                                 ((uint64_t)(uint32_t)__shfl((int)(uint32_t)((uintptr_t)row >> 32), wl) << 32));
              const int gix = (int)F;
              const bool first = gi == 0;
              // the normals of this lane's step in each chunk, all read before the first chunk
              float zc[4];
#pragma unroll
              for (int c = 0; c < 4; ++c) {
                const int src = ((c * G + gi) >> 2) * NW + (gwalk ? grp : 0);
                const float z0 = __shfl(hz[0], src), z1 = __shfl(hz[1], src), z2 = __shfl(hz[2], src),
                            z3 = __shfl(hz[3], src);
                const float za = (gi & 1) ? z1 : z0, zb = (gi & 1) ? z3 : z2;
                zc[c] = (gi & 2) ? zb : za;
              }
              float yf = Y, ff = F;
              bool gexit = false;
#pragma unroll
              for (int c = 0; c < 4; ++c) {
                iters += c ? G / 4 : G / 4 - 1;  // blocks of four steps; the loop counted one
                const int q = c * G + gi;  // this lane's step within the pass
                const float z = zc[c];     // the normal of step q
                float V;
                if constexpr (MODEL == SSM_MODEL_ORNSTEIN) {
                  // the increment depends on the position: one shift and one step per link
                  float zs = row_shr1(z);
                  dpp_fence(zs);
                  V = Y;
#pragma unroll
                  for (int s = 1; s < G; ++s) {
                    float vs = row_shr1(V);
                    dpp_fence(vs);
                    const float vn = vs + walk_increment<MODEL, REPLAY>(vs, gvdt, gv, gleak, dt, 0.f, zs);
                    V = first ? Y : vn;
                  }
                } else {
                  const float e = row_shr1_or(Y, walk_increment<MODEL, REPLAY>(0.f, gvdt, gv, gleak, dt, 0.f, z));
                  V = e;
                  if constexpr (G == 8) V = first ? Y : V;
#pragma unroll
                  for (int s = 1; s < G; ++s) {
                    V = row_shr1(V) + e;
                    if constexpr (G == 8) {
                      dpp_fence(V);
                      V = first ? Y : V;
                    }
                  }
                }
                const bool okq = !gwalk || inside<MODEL>(V, boundary_at_f<BOUND>(F + (float)q, gix + q, dt, ga, gb1, gb2, grow));
                // the state after this lane's step; lane G - 1's starts the next chunk
                const float Vn = V + walk_increment<MODEL, REPLAY>(V, gvdt, gv, gleak, dt, 0.f, z);
                const uint64_t fail = __ballot(!okq);
                if (fail != 0) {
                  const uint32_t fg = (uint32_t)(fail >> (grp * G)) & ((1u << G) - 1u);
                  gexit = fg != 0;
                  const int j = gexit ? __builtin_ctz(fg) : G - 1;
                  const int sl = grp * G + j;
                  yf = __shfl(gexit ? V : Vn, sl);
                  ff = F + (float)(c * G + (gexit ? j : G));
                  break;
                }
                if constexpr (G == 16) {
                  Y = row_bcast15(Vn);
                } else {
                  float y7 = row_bcast7(Vn), y15 = row_bcast15(Vn);
                  dpp_fence(y7);
                  dpp_fence(y15);
                  Y = (lane & 8) ? y15 : y7;
                }
                yf = Y;
                ff = F + (float)((c + 1) * G);
              }
              const int sg = alive ? r * G : 0;
              const float ny = __shfl(yf, sg), nf = __shfl(ff, sg);
              const bool nex = __shfl((int)gexit, sg) != 0;
              if (alive) {  // t follows from the step count after the iteration
                x[0] = ny;
                fx = nf;
                alive = !nex;
              }
            };
            if (lone) {
              if (K <= 4) chain_tail(std::integral_constant<int, 16>{});
              else chain_tail(std::integral_constant<int, 8>{});
            } else {
            const int lgk = K <= 1 ? 0 : 32 - __clz(K - 1);
            const int kp = 1 << lgk;
            const int nb = kSubBlocks<false> < (64 >> lgk) ? kSubBlocks<false> : (64 >> lgk);  // blocks per pass
            const int ol = __shfl(rank_lane, lane & (kp - 1));  // the walk this lane draws for
            const uint32_t on = (uint32_t)__shfl(n, ol), ok = (uint32_t)__shfl(k, ol);
            const uint32_t oblk = (uint32_t)__shfl(ix, ol) >> 2;  // walks are block-aligned
            const float obmc = __shfl(bmc, ol);
            float hz[4];
            bool more = true;
            for (int p0 = 0; p0 < kSubBlocks<false> && more; p0 += nb) {
              philox_normals4_c(on, ok, oblk + (uint32_t)(p0 + (lane >> lgk)), 0u, A.k0, A.k1, obmc, hz);
              // this walk's block p0 + j sits in lane r + kp j; the next block is read while
              // the current one steps
              float cur[4], nxt[4];
#pragma unroll
              for (int i = 0; i < 4; ++i) cur[i] = __shfl(hz[i], r);
#pragma unroll
              for (int j = 0; j < kSubBlocks<false>; ++j) {
                if (j >= nb) break;
                if (p0 + j > 0) {
                  if (__ballot(alive) == 0) {
                    more = false;
                    break;
                  }
                  ++iters;
                }
                if (j + 1 < nb) {
#pragma unroll
                  for (int i = 0; i < 4; ++i) nxt[i] = __shfl(hz[i], r + ((j + 1) << lgk));
                }
                if (alive) {
                  // A wave with few walks left waits on each walk's step latency rather than on
                  // issue slots. The block's four positions are the same float sums whether or
                  // not a test fails, so they are formed first; the four tests then run
                  // independently and the walk keeps the state at its first failing test
                  // (fx + i is exact; t comes from the table after the iteration). Same steps,
                  // same rounding, a chain of four adds and four selects per block instead of
                  // four test-and-branch rounds.
                  const int ixb = (int)fx;  // this block's first step index
                  float yv[5];
                  yv[0] = x[0];
#pragma unroll
                  for (int i = 0; i < 4; ++i)
                    yv[i + 1] = yv[i] + walk_increment<MODEL, REPLAY>(yv[i], vdt[0], vraw[0], leak, dt, sst[0], cur[i]);
                  bool okb[4];
#pragma unroll
                  for (int i = 0; i < 4; ++i)
                    okb[i] = inside<MODEL>(yv[i], boundary_at_f<BOUND>(i ? fx + (float)i : fx, ixb + i, dt, a, b1, b2, row));
                  float yf = yv[4], ff = fx + 4.0f;
#pragma unroll
                  for (int i = 3; i >= 0; --i) {
                    yf = okb[i] ? yf : yv[i];
                    ff = okb[i] ? ff : (i ? fx + (float)i : fx);
                  }
                  alive = okb[0] && okb[1] && okb[2] && okb[3];
                  x[0] = yf;
                  fx = ff;
                }
#pragma unroll
                for (int i = 0; i < 4; ++i) cur[i] = nxt[i];
              }
            }
            }
          }
        }
        if (!helped) {
#pragma unroll
        for (int sb = 0; sb < kSubBlocks<REPLAY>; ++sb) {
          if (sb > 0) {
            if (__ballot(alive) == 0) break;
            ++iters;
          }
          if (alive) {
            const int ixb = (int)fx;  // this block's first step index
            float nz[4];
            if constexpr (REPLAY) {
              constexpr int off = MODEL == SSM_MODEL_FULL_DDM ? 1 : 0;
              const float* rb = A.replay_normals + g * A.replay_stride;
              const int64_t d0 = kTwoStage<MODEL> ? (int64_t)nh + (ixb - ix0) : (int64_t)(off + ixb);
#pragma unroll
              for (int i = 0; i < 4; ++i) nz[i] = rb[(d0 + i) % A.replay_stride];
            } else if constexpr (MODEL == SSM_MODEL_LEVY) {
              const u32x4 r0 = philox4x32_10(u32x4{(uint32_t)n, (uint32_t)k, (uint32_t)(ixb >> 2), 0u}, A.k0, A.k1);
              const u32x4 r1 = philox4x32_10(u32x4{(uint32_t)n, (uint32_t)k, (uint32_t)(ixb >> 2), 2u}, A.k0, A.k1);
              nz[0] = stable_cms(r0.x, r0.y, alpha, alpha_inv, alpha_scale);
              nz[1] = stable_cms(r0.z, r0.w, alpha, alpha_inv, alpha_scale);
              nz[2] = stable_cms(r1.x, r1.y, alpha, alpha_inv, alpha_scale);
              nz[3] = stable_cms(r1.z, r1.w, alpha, alpha_inv, alpha_scale);
            } else if constexpr (kTwoStage<MODEL>) {
              philox_normals4_c((uint32_t)n, (uint32_t)k, (uint32_t)((ixb - ix0) >> 2), stage ? 4u : 0u, A.k0, A.k1, bmc, nz);
            } else {
              philox_normals4_c((uint32_t)n, (uint32_t)k, (uint32_t)(ixb >> 2), 0u, A.k0, A.k1, bmc, nz);
            }
#pragma unroll
            for (int i = 0; i < 4; ++i) {
              if (alive) {
                const float b = boundary_at_f<BOUND>(fx, ixb + i, dt, a, b1, b2, row);
                if (inside<MODEL>(x[0], b)) {
                  // keeps the step a branch: speculated as selects, the first step of each
                  // block costs three VALU instructions more than the exec-mask branch
                  asm volatile("");
                  x[0] += walk_increment<MODEL, REPLAY>(x[0], vdt[0], vraw[0], leak, dt, sst[0], nz[i]);
                  fx += 1.0f;
                } else {
                  alive = false;
                }
              }
            }
          }
        }
        }
        // t after ix steps, from the table of the same sequential float32 sums
        if (state == kRunning) {
          ix = (int)fx;
          t = A.t_grid[ix];
        }
      } else {
#pragma unroll
      for (int sb = 0; sb < kSubBlocksR<REPLAY, RECORD>; ++sb) {
        if (sb > 0) {
          if (__ballot(alive) == 0) break;  // wave-uniform: no lane left to step
          ++iters;
        }
        if (alive) {
          float nz[4];
          if constexpr (REPLAY) {
            // a full-DDM row starts with the drift draw, the walk's draws follow it; a two-stage
            // row holds the high walk's draws, then the low walk's
            constexpr int off = MODEL == SSM_MODEL_FULL_DDM ? 1 : 0;
            const float* rb = A.replay_normals + g * A.replay_stride;
            const int64_t d0 = kTwoStage<MODEL> ? (int64_t)nh + (ix - ix0) : (int64_t)(off + ix);
#pragma unroll
            for (int i = 0; i < 4; ++i) nz[i] = rb[(d0 + i) % A.replay_stride];
          } else if constexpr (MODEL == SSM_MODEL_LEVY) {
            // two Philox blocks per four steps: (U, W) pairs for four stable variates
            const u32x4 r0 = philox4x32_10(u32x4{(uint32_t)n, (uint32_t)k, (uint32_t)(ix >> 2), 0u}, A.k0, A.k1);
            const u32x4 r1 = philox4x32_10(u32x4{(uint32_t)n, (uint32_t)k, (uint32_t)(ix >> 2), 2u}, A.k0, A.k1);
            nz[0] = stable_cms(r0.x, r0.y, alpha, alpha_inv, alpha_scale);
            nz[1] = stable_cms(r0.z, r0.w, alpha, alpha_inv, alpha_scale);
            nz[2] = stable_cms(r1.x, r1.y, alpha, alpha_inv, alpha_scale);
            nz[3] = stable_cms(r1.z, r1.w, alpha, alpha_inv, alpha_scale);
          } else if constexpr (kTwoStage<MODEL>) {
            // each stage's draws start block-aligned in a stream of their own (0 high, 4 low)
            philox_normals4_c((uint32_t)n, (uint32_t)k, (uint32_t)((ix - ix0) >> 2), stage ? 4u : 0u, A.k0, A.k1, bmc, nz);
          } else {
            philox_normals4_c((uint32_t)n, (uint32_t)k, (uint32_t)(ix >> 2), 0u, A.k0, A.k1, bmc, nz);
          }
          // Steps use exec-mask branches rather than selects: with every lane usually on the
          // same side, a branch costs scalar instructions while a select costs VALU issue slots.
          if (__ballot(ix + 4 > A.n_steps) == 0) {
            // common case: no lane can reach the end of t_s in this block, no index guard needed
#pragma unroll
            for (int i = 0; i < 4; ++i) {
              if (alive) {
                const float b = boundary_at<BOUND>(ix, dt, a, b1, b2, row);
                if (inside<MODEL>(x[0], b) && (t <= dl_tmp)) {
                  x[0] += walk_increment<MODEL, REPLAY>(x[0], vdt[0], vraw[0], leak, dt, sst[0], nz[i]);
                  t += dt;
                  ++ix;
                  if constexpr (RECORD) {
                    if (g == 0 && ix < A.traj_rows) A.traj[ix] = x[0];
                  }
                } else {
                  alive = false;
                }
              }
            }
          } else {
#pragma unroll
            for (int i = 0; i < 4; ++i) {
              if (alive) {
                bool go = ix < A.n_steps;  // never read past a boundary table
                if (go) {
                  const float b = boundary_at<BOUND>(ix, dt, a, b1, b2, row);
                  go = inside<MODEL>(x[0], b) && (t <= dl_tmp);
                }
                if (go) {
                  x[0] += walk_increment<MODEL, REPLAY>(x[0], vdt[0], vraw[0], leak, dt, sst[0], nz[i]);
                  t += dt;
                  ++ix;
                  if constexpr (RECORD) {
                    if (g == 0 && ix < A.traj_rows) A.traj[ix] = x[0];
                  }
                } else {
                  alive = false;
                }
              }
            }
          }
        }
      }
      }
      if (state == kRunning) {
        if (!alive) {
          if constexpr (kTwoStage<MODEL>) {
            if (stage == 0) {
              // the high walk ended: its choice by the coin rule at the exit bound, then the low
              // walk of that choice (parallel_models.pyx:474-489, sequential_models.pyx:474-533)
              const float bx = ix < A.n_steps ? boundary_at<BOUND>(ix, dt, a, b1, b2, row) : 0.f;
              const bool hi = two_stage_coin(uu[0], x[0], bx);
              ch2 = hi ? 2 : 0;
              const float zl = hi ? zl2 : zl1;
              vdt[0] = (hi ? vl2 : vl1) * dt;
              t_h = t;
              nh = ix;
              stage = 1;
              alive = true;
              if constexpr (MODEL == SSM_MODEL_PAR2) {
                // the low walk starts afresh at t = 0 from b(0)
                const float b0 = boundary_at<BOUND>(0, dt, a, b1, b2, row);
                x[0] = (-b0) + ((zl * 2.0f) * b0);
                t = 0.f;
                ix = 0;
                ix0 = 0;
              } else {
                ix0 = ix;
                if (t >= max_t) {
                  // no time left: the low choice from its prior, no low walk (:474-499)
                  ch2 += uu[1] <= zl ? 1 : 0;
                  decided = true;
                  alive = false;
                } else {
                  // the low walk continues from the high walk's end, from b(ix); a start
                  // outside the bounds takes the prior choice (u < zl) and does not walk
                  x[0] = (-bx) + ((zl * 2.0f) * bx);
                  if ((x[0] >= bx) || (x[0] <= -bx)) {
                    ch2 += uu[1] < zl ? 1 : 0;
                    decided = true;
                  }
                }
              }
            }
          }
          if (!alive) state = kDone;
        }
      }
    } else {
      // up to kAccSteps steps per loop iteration (the loop's own work is paid once for them).
      // Native LCA: an iteration in which no walking lane can reach the end of t_s or its
      // deadline skips the per-step index and deadline tests, as the diffusions' do (measured:
      // LCA-4 faster, race-4 slower with it)
      constexpr int kIter = kAccStepsR<REPLAY, MAXP, RECORD>;
      bool untested = false;
      if constexpr (MODEL == SSM_MODEL_LCA && !REPLAY && !RECORD)
        untested = __ballot(state == kRunning && !(ix + kIter <= A.n_steps && t + A.t_fast[0] <= dl_tmp)) == 0;
#pragma unroll
      for (int sb = 0; sb < kIter; ++sb) {
        if (sb > 0) {
          if (__ballot(state == kRunning) == 0) break;  // wave-uniform: no lane left to step
          ++iters;
        }
        if (state == kRunning) {
          bool go = true;
          if (!untested) go = ix < A.n_steps && t <= dl_tmp;
          if (MODEL != SSM_MODEL_RDM && go) {  // racing diffusion tests the threshold inside the step
            const float b = boundary_at<BOUND>(ix, dt, a, b1, b2, row);
#pragma unroll
            for (int j = 0; j < MAXP; ++j)
              if (j < P && x[j] > b) go = false;  // check_finished, race_models.pyx:62-77
          }
          if (!go) {
            state = kDone;
          } else {
            float nz[MAXP];
            if constexpr (REPLAY) {
              const float* rb = A.replay_normals + g * A.replay_stride;
              const int64_t d0 = (int64_t)ix * P;
#pragma unroll
              for (int j = 0; j < MAXP; ++j)
                nz[j] = (j < P) ? rb[(d0 + j) % A.replay_stride] : 0.f;
            } else {
#pragma unroll
              for (int q = 0; q < (MAXP + 3) / 4; ++q) {
                if (4 * q < P) {
                  float tmp[4];
                  philox_normals4((uint32_t)n, (uint32_t)k, (uint32_t)ix, (uint32_t)(2 * q), A.k0, A.k1, tmp);
#pragma unroll
                  for (int i = 0; i < 4; ++i)
                    if (4 * q + i < MAXP) nz[4 * q + i] = tmp[i];
                }
              }
            }
            if constexpr (MODEL == SSM_MODEL_RDM) {
            } else if constexpr (MODEL == SSM_MODEL_RACE) {
#pragma unroll
              for (int j = 0; j < MAXP; ++j) {
                if (j < P) {
                  const float xn = x[j] + (vdt[j] + sst[j] * nz[j]);  // race_models.pyx:441
                  x[j] = fmaxf(0.0f, xn);
                }
              }
            } else {
              float sum = 0.f;  // csum, _utils.pyx:486-502 (sequential float32 sum)
#pragma unroll
              for (int j = 0; j < MAXP; ++j)
                if (j < P) sum += x[j];
              float xnew[MAXP];
#pragma unroll
              for (int j = 0; j < MAXP; ++j) {
                if (j < P) {
                  const float red = (-x[j]) + sum;  // race_models.pyx:851
                  if constexpr (REPLAY) {
                    const float drift = ((vraw[j] - (leak * x[j])) - (inh * red)) * dt;
                    xnew[j] = fmaxf(0.0f, x[j] + (drift + sst[j] * nz[j]));  // :852-854
                  } else {
                    // native walks: the same update, (v - g x - b red) dt + s sqrt(dt) z, with dt
                    // taken into the rates once per walk and each product fused with its sum
                    const float dr = __builtin_fmaf(-leak_dt, x[j], __builtin_fmaf(-inh_dt, red, vdt[j]));
                    xnew[j] = fmaxf(0.0f, x[j] + __builtin_fmaf(sst[j], nz[j], dr));
                  }
                } else {
                  xnew[j] = x[j];
                }
              }
#pragma unroll
              for (int j = 0; j < MAXP; ++j) x[j] = xnew[j];
            }
            if constexpr (MODEL == SSM_MODEL_RDM) {
              // race_models.pyx:1221-1237: accumulators move in order and the first to reach
              // b ends the walk before the rest move and before t advances; the later ones'
              // positions are never read, so all move here and the first crossing wins
              const float b = boundary_at<BOUND>(ix, dt, a, b1, b2, row);
#pragma unroll
              for (int j = MAXP - 1; j >= 0; --j) {
                if (j < P) {
                  x[j] = x[j] + (vdt[j] + sst[j] * nz[j]);
                  if (x[j] >= b) win = j;
                }
              }
              if (win >= 0) state = kDone;
            }
            if (MODEL != SSM_MODEL_RDM || win < 0) {
              t += dt;
              ++ix;
              if constexpr (RECORD) {
                if (g == 0 && ix < A.traj_rows) {
#pragma unroll
                  for (int j = 0; j < MAXP; ++j)
                    if (j < P) A.traj[(int64_t)ix * P + j] = x[j];
                }
              }
            }
          }
        }
      }
    }
  }

  // one atomic per wavefront for each counter
  unsigned long long s = my_steps;
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
  if (lane_id() == 0) {
    if (s) atomicAdd(A.ws + 1, s);
    atomicAdd(A.ws + 2, iters * (kDiffusion<MODEL> ? 256ull : 64ull));  // diffusions: 4 steps per iteration
  }
}

__global__ void ws_reset_kernel(unsigned long long* ws) {
  if (threadIdx.x < SSM_EM_WORKSPACE_WORDS) ws[threadIdx.x] = 0ull;
}

// --------------------------------------------------------------------------------
using KernelFn = void (*)(EmArgs);

template <int MODEL, int BOUND, bool REPLAY, bool RECORD>
static KernelFn pick_p(int p) {
  if constexpr (kDiffusion<MODEL>) {
    return em_kernel<MODEL, BOUND, REPLAY, 1, RECORD>;
  } else {
    if (p <= 4) return em_kernel<MODEL, BOUND, REPLAY, 4, RECORD>;
    if (p <= 8) return em_kernel<MODEL, BOUND, REPLAY, 8, RECORD>;
    return em_kernel<MODEL, BOUND, REPLAY, 16, RECORD>;
  }
}

template <int MODEL, bool REPLAY, bool RECORD>
static KernelFn pick_bound(int bound, int p) {
  switch (bound) {
    case SSM_BOUNDARY_CONSTANT: return pick_p<MODEL, SSM_BOUNDARY_CONSTANT, REPLAY, RECORD>(p);
    case SSM_BOUNDARY_ANGLE: return pick_p<MODEL, SSM_BOUNDARY_ANGLE, REPLAY, RECORD>(p);
    case SSM_BOUNDARY_WEIBULL: return pick_p<MODEL, SSM_BOUNDARY_WEIBULL, REPLAY, RECORD>(p);
    default: return pick_p<MODEL, SSM_BOUNDARY_TABLE, REPLAY, RECORD>(p);
  }
}

template <int MODEL, bool RECORD>
static KernelFn pick_mode(int bound, int p, bool replay) {
  return replay ? pick_bound<MODEL, true, RECORD>(bound, p) : pick_bound<MODEL, false, RECORD>(bound, p);
}

static KernelFn pick_kernel(int model, int bound, int p, bool replay) {
  switch (model) {
    case SSM_MODEL_DDM: return pick_mode<SSM_MODEL_DDM, false>(bound, p, replay);
    case SSM_MODEL_FULL_DDM: return pick_mode<SSM_MODEL_FULL_DDM, false>(bound, p, replay);
    case SSM_MODEL_ORNSTEIN: return pick_mode<SSM_MODEL_ORNSTEIN, false>(bound, p, replay);
    case SSM_MODEL_LEVY: return pick_mode<SSM_MODEL_LEVY, false>(bound, p, replay);
    case SSM_MODEL_DDM_LEGACY: return pick_mode<SSM_MODEL_DDM_LEGACY, false>(bound, p, replay);
    case SSM_MODEL_PAR2: return pick_mode<SSM_MODEL_PAR2, false>(bound, p, replay);
    case SSM_MODEL_SEQ2: return pick_mode<SSM_MODEL_SEQ2, false>(bound, p, replay);
    case SSM_MODEL_RDM: return pick_mode<SSM_MODEL_RDM, false>(bound, p, replay);
    case SSM_MODEL_RACE: return pick_mode<SSM_MODEL_RACE, false>(bound, p, replay);
    default: return pick_mode<SSM_MODEL_LCA, false>(bound, p, replay);
  }
}

// trajectory-recording kernels: every model but the two-stage ones
static KernelFn pick_record_kernel(int model, int bound, int p, bool replay) {
  switch (model) {
    case SSM_MODEL_DDM: return pick_mode<SSM_MODEL_DDM, true>(bound, p, replay);
    case SSM_MODEL_FULL_DDM: return pick_mode<SSM_MODEL_FULL_DDM, true>(bound, p, replay);
    case SSM_MODEL_ORNSTEIN: return pick_mode<SSM_MODEL_ORNSTEIN, true>(bound, p, replay);
    case SSM_MODEL_LEVY: return pick_mode<SSM_MODEL_LEVY, true>(bound, p, replay);
    case SSM_MODEL_DDM_LEGACY: return pick_mode<SSM_MODEL_DDM_LEGACY, true>(bound, p, replay);
    case SSM_MODEL_RDM: return pick_mode<SSM_MODEL_RDM, true>(bound, p, replay);
    case SSM_MODEL_RACE: return pick_mode<SSM_MODEL_RACE, true>(bound, p, replay);
    case SSM_MODEL_LCA: return pick_mode<SSM_MODEL_LCA, true>(bound, p, replay);
    default: return nullptr;
  }
}

// t after `steps` additions of dt from 0 in float32, as the walk accumulates it; the value of
// the last (steps, dt) asked for is kept, since every launch of a job asks the same
static float t_after_steps(int32_t steps, float dt) {
  static thread_local int32_t c_steps = -1;
  static thread_local uint32_t c_dt = 0;
  static thread_local float c_t = 0.f;
  uint32_t dt_bits;
  std::memcpy(&dt_bits, &dt, sizeof dt_bits);
  if (steps == c_steps && dt_bits == c_dt) return c_t;
  float t = 0.f;
  for (int32_t i = 0; i < steps; ++i) t = t + dt;
  c_steps = steps;
  c_dt = dt_bits;
  c_t = t;
  return t;
}

// Bounds on the rise of t over `steps` - 1 further steps of t += dt in float32, for walk-loop
// iterations of `steps` steps. A walk's t never exceeds T = t after n_steps - 1 steps, so
// every rounded step adds at most dt + u/2 with u = ulp(2 max(T, n_steps dt)); the bound
// takes dt + u per step plus 2u for the rounding of t + bound in the kernel's test, and is
// rounded up to float.
static void fill_t_fast(const ssm_em_problem* P, EmArgs& A) {
  const int steps[3] = {4 * kSubBlocks<false>, 4 * kSubBlocks<true>, kLoneSteps};
  if (P->n_steps > (1 << 24)) {
    A.t_fast[0] = A.t_fast[1] = A.t_fast[2] = std::numeric_limits<float>::quiet_NaN();
    return;
  }
  const double top = 2.0 * std::max((double)t_after_steps(P->n_steps - 1, P->dt), (double)P->n_steps * P->dt);
  int e = 0;
  std::frexp(top, &e);
  const double u = std::ldexp(1.0, e - 24);
  for (int r = 0; r < 3; ++r) {
    const double b = (steps[r] - 1) * ((double)P->dt + u) + 2.0 * u;
    A.t_fast[r] = std::nextafter((float)b, std::numeric_limits<float>::infinity());
  }
}

// The walks' times after k steps, k = 0..n_steps, one table per (device, dt, n_steps), built
// on the host (the same sequential float32 additions as the walk) and kept on the device for
// the life of the process: 4 (n_steps + 1) bytes each. The first launch of a (dt, n_steps)
// on a device allocates and uploads it, so it may not run inside a stream capture.
int32_t em_time_grid(float dt, int32_t n_steps, hipStream_t stream, const float** out) {
  *out = nullptr;
  if (n_steps > (1 << 24)) return SSM_OK;  // no untested iterations: the table is never read
  int dev = 0;
  SSM_HIP(hipGetDevice(&dev));
  uint32_t dt_bits;
  std::memcpy(&dt_bits, &dt, 4);
  static std::mutex mu;
  static std::map<std::tuple<int, uint32_t, int32_t>, float*> grids;
  std::lock_guard<std::mutex> lock(mu);
  const auto key = std::make_tuple(dev, dt_bits, n_steps);
  const auto it = grids.find(key);
  if (it != grids.end()) {
    *out = it->second;
    return SSM_OK;
  }
  if (stream) {
    hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
    SSM_HIP(hipStreamIsCapturing(stream, &st));
    if (st != hipStreamCaptureStatusNone)
      return fail(SSM_EINVAL, "the first launch of this (delta_t, n_steps) on this device is inside a stream "
                              "capture: launch it once outside the capture (it builds the walks' time table)");
  }
  std::vector<float> h((size_t)n_steps + 1);
  float t = 0.f;
  for (int32_t k = 0; k <= n_steps; ++k) {
    h[k] = t;
    t = t + dt;
  }
  float* d = nullptr;
  SSM_HIP(hipMalloc(&d, h.size() * sizeof(float)));
  const hipError_t e = hipMemcpy(d, h.data(), h.size() * sizeof(float), hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    (void)hipFree(d);
    return hip_fail(e, "hipMemcpy (time table)");
  }
  grids.emplace(key, d);
  *out = d;
  return SSM_OK;
}

int32_t em_validate(const ssm_em_problem* P, const float* set_params, const int64_t* offsets,
                    const float* table, const float* rn, const float* ru, float* rts, void* choices,
                    uint64_t* ws) {
  if (!P) return fail(SSM_EINVAL, "problem descriptor is NULL");
  if (P->model < SSM_MODEL_DDM || P->model > SSM_MODEL_DDM_LEGACY) return fail(SSM_EINVAL, "unknown model");
  if (P->boundary < SSM_BOUNDARY_CONSTANT || P->boundary > SSM_BOUNDARY_TABLE) return fail(SSM_EINVAL, "unknown boundary");
  const bool diffusion = P->model == SSM_MODEL_DDM || P->model == SSM_MODEL_FULL_DDM ||
                         P->model == SSM_MODEL_ORNSTEIN || P->model == SSM_MODEL_LEVY ||
                         P->model == SSM_MODEL_PAR2 || P->model == SSM_MODEL_SEQ2 ||
                         P->model == SSM_MODEL_DDM_LEGACY;
  if (diffusion && P->n_particles != 1) return fail(SSM_EINVAL, "single-particle models require n_particles == 1");
  if (P->n_particles < 1 || P->n_particles > SSM_MAX_PARTICLES)
    return fail(SSM_EINVAL, "n_particles must be in [1, 16]");
  if (P->n_sets < 0 || P->n_total < 0) return fail(SSM_EINVAL, "negative sizes");
  if (P->n_total > 0 && P->n_sets == 0) return fail(SSM_EINVAL, "n_total > 0 with no sets");
  if (P->n_steps < 1) return fail(SSM_EINVAL, "n_steps must be >= 1");
  if (!(P->dt > 0.f)) return fail(SSM_EINVAL, "delta_t must be positive");
  if (P->out_layout != 0 && P->out_layout != 1) return fail(SSM_EINVAL, "out_layout must be 0 or 1");
  if (P->choice_bytes != 1 && P->choice_bytes != 4) return fail(SSM_EINVAL, "choice_bytes must be 1 or 4");
  if (P->replay_stride < 0) return fail(SSM_EINVAL, "replay_stride must be >= 0");
  const bool two_stage = P->model == SSM_MODEL_PAR2 || P->model == SSM_MODEL_SEQ2;
  const int32_t min_stride = (diffusion && !two_stage) ? SSM_SET_STRIDE_SINGLE : SSM_SET_STRIDE;
  if (P->set_stride != 0 && (P->set_stride < min_stride || P->set_stride > (1 << 16)))
    return fail(SSM_EINVAL, (diffusion && !two_stage) ? "set_stride must be 0 or in [16, 65536] for this model"
                                                      : "set_stride must be 0 or in [64, 65536] for this model");
  if (P->out_layout == 1 && P->n_sets > 0 && P->n_total % P->n_sets != 0)
    return fail(SSM_EINVAL, "out_layout 1 needs the same n_samples for every set");
  if (P->n_total > 0) {
    if (!set_params || !offsets || !rts || !choices || !ws) return fail(SSM_EINVAL, "NULL buffer");
    if (P->boundary == SSM_BOUNDARY_TABLE && !table) return fail(SSM_EINVAL, "TABLE boundary needs a table");
    if (P->replay_stride > 0 && !rn) return fail(SSM_EINVAL, "replay mode needs a normals buffer");
    if (P->replay_stride > 0 && P->smooth_unif && !ru) return fail(SSM_EINVAL, "replay + smooth_unif needs uniforms");
    if (P->replay_stride > 0 && P->model == SSM_MODEL_FULL_DDM && !ru)
      return fail(SSM_EINVAL, "full-DDM replay needs uniforms (u_t, u_z, u_smooth per trajectory)");
    if (P->replay_stride > 0 && (P->model == SSM_MODEL_PAR2 || P->model == SSM_MODEL_SEQ2) && !ru)
      return fail(SSM_EINVAL, "two-stage replay needs uniforms (four per trajectory)");
    if (P->replay_stride > 0 && P->model == SSM_MODEL_RDM && !ru)
      return fail(SSM_EINVAL, "racing-diffusion replay needs uniforms (start draws and u_smooth per trajectory)");
  }
  return SSM_OK;
}

int32_t em_enqueue(const ssm_em_problem* P, const ssm_launch_config* L, const float* set_params,
                   const int64_t* offsets, const float* table, const float* rn, const float* ru,
                   float* rts, void* choices, uint64_t* ws, hipStream_t stream) {
  int block = 256, per_cu = 8, refill = 8;
  if (L) {
    if (L->block_size) block = L->block_size;
    if (L->blocks_per_cu) per_cu = L->blocks_per_cu;
    if (L->refill_min) refill = L->refill_min;
  }
  if (block < 64 || block > 1024 || block % 64) return fail(SSM_EINVAL, "block_size must be a multiple of 64 in [64, 1024]");
  if (per_cu < 1 || per_cu > 64) return fail(SSM_EINVAL, "blocks_per_cu must be in [1, 64]");
  if (refill < 1 || refill > 64) return fail(SSM_EINVAL, "refill_min must be in [1, 64]");
  if (L && (L->queue_block < 0 || L->queue_block > 65536))
    return fail(SSM_EINVAL, "queue_block must be in [0, 65536]");
  if (L && (L->set_order != nullptr) != (L->queue_offsets != nullptr))
    return fail(SSM_EINVAL, "set_order and queue_offsets go together");
  const float* t_grid = nullptr;
  if (P->n_total > 0) {
    const int32_t rc = em_time_grid(P->dt, P->n_steps, stream, &t_grid);
    if (rc) return rc;
  }
  hipLaunchKernelGGL(ws_reset_kernel, dim3(1), dim3(64), 0, stream, (unsigned long long*)ws);
  SSM_HIP(hipGetLastError());
  if (P->n_total == 0) return SSM_OK;

  EmArgs A;
  A.t_grid = t_grid;
  A.set_params = set_params;
  A.set_stride = P->set_stride ? P->set_stride : SSM_SET_STRIDE;
  A.offsets = offsets;
  A.order = (L && L->set_order) ? L->set_order : nullptr;
  A.qoff = A.order ? L->queue_offsets : offsets;
  A.table = table;
  A.replay_normals = rn;
  A.replay_uniforms = ru;
  A.rts = rts;
  A.choices = choices;
  A.traj = nullptr;
  A.traj_rows = 0;
  A.ws = (unsigned long long*)ws;
  A.total = P->n_total;
  A.replay_stride = P->replay_stride;
  A.n_sets = P->n_sets;
  A.n_steps = P->n_steps;
  A.n_particles = P->n_particles;
  A.smooth_unif = P->smooth_unif;
  A.out_layout = P->out_layout;
  A.choice_bytes = P->choice_bytes;
  A.refill_min = refill;
  A.queue_block = (L && L->queue_block) ? L->queue_block : 64;
  A.dt = P->dt;
  fill_t_fast(P, A);
  A.k0 = (uint32_t)P->seed;
  A.k1 = (uint32_t)(P->seed >> 32);

  const int cus = device_cu_count();
  if (cus <= 0) return fail(SSM_EHIP, "no HIP device");
  // A job of at most 16 walks per SIMD is bound by its longest walk's latency, not by issue:
  // the native Gaussian diffusions then run 4 walks per wave, so every walk steps in the
  // lane-shift chains (~55 cycles per step instead of ~170); up to 40 walks per SIMD they run
  // 16 per wave (the shared Philox blocks, ~150 cycles). Larger jobs keep full waves, whose
  // steps cost fewer instructions (measured, one set: 1e3-1e4 walks 0.40 -> 0.20-0.26 ms,
  // 2e4-3e4 walks 0.40 -> 0.33-0.36 ms; 6e4 walks and more slower capped). The other models
  // and replay have no tail modes and keep full waves.
  const bool chains = P->replay_stride == 0 && (P->model == SSM_MODEL_DDM || P->model == SSM_MODEL_FULL_DDM ||
                                                P->model == SSM_MODEL_ORNSTEIN || P->model == SSM_MODEL_DDM_LEGACY);
  const int64_t per_simd = (P->n_total + (int64_t)cus * 4 - 1) / ((int64_t)cus * 4);
  A.max_walks = !chains ? 64 : per_simd <= 16 ? kSmallJobWalks : per_simd <= 40 ? kMidJobWalks : 64;
  int64_t grid = (int64_t)cus * per_cu;
  const int64_t needed = (P->n_total * (64 / A.max_walks) + block - 1) / block;
  if (grid > needed) grid = needed;
  if (grid < 1) grid = 1;
  KernelFn fn = pick_kernel(P->model, P->boundary, P->n_particles, P->replay_stride > 0);
  hipLaunchKernelGGL(fn, dim3((unsigned)grid), dim3(block), 0, stream, A);
  SSM_HIP(hipGetLastError());
  return SSM_OK;
}

// Walk (set 0, sample 0) of a problem once more, one wave of one block, storing its positions:
// traj[ix][j] for ix = 0 (start) up to the last step, at most traj_rows rows; later rows are left
// as the caller filled them. The draws are the problem's own (the Philox counter of native mode
// depends only on (sample, set); replay mode reads row 0), so the walk is the one the full launch
// produced for that trajectory.
int32_t em_trajectory_enqueue(const ssm_em_problem* P, const float* set_params, const int64_t* offsets,
                              const float* table, const float* rn, const float* ru, float* traj, int32_t traj_rows,
                              uint64_t* ws, hipStream_t stream) {
  if (!traj) return fail(SSM_EINVAL, "traj is NULL");
  if (traj_rows < 1) return fail(SSM_EINVAL, "traj_rows must be >= 1");
  // rts / choices are never written by the recording kernels: the trajectory buffer stands in
  // for them in the shared validation
  int32_t rc = em_validate(P, set_params, offsets, table, rn, ru, traj, traj, ws);
  if (rc) return rc;
  if (!ws) return fail(SSM_EINVAL, "workspace is NULL");
  if (P->n_total < 1 || P->n_sets < 1) return fail(SSM_EINVAL, "the problem has no trajectory to record");
  KernelFn fn = pick_record_kernel(P->model, P->boundary, P->n_particles, P->replay_stride > 0);
  if (!fn) return fail(SSM_EUNSUPPORTED, "trajectory recording is not available for the two-stage models");
  hipLaunchKernelGGL(ws_reset_kernel, dim3(1), dim3(64), 0, stream, (unsigned long long*)ws);
  SSM_HIP(hipGetLastError());
  EmArgs A;
  A.t_grid = nullptr;  // one walk, every step tested
  A.set_params = set_params;
  A.set_stride = P->set_stride ? P->set_stride : SSM_SET_STRIDE;
  A.offsets = offsets;
  A.order = nullptr;
  A.qoff = offsets;
  A.table = table;
  A.replay_normals = rn;
  A.replay_uniforms = ru;
  A.rts = nullptr;
  A.choices = nullptr;
  A.traj = traj;
  A.traj_rows = traj_rows;
  A.ws = (unsigned long long*)ws;
  A.total = 1;  // queue index 0 only: set 0, sample 0 (set 0 must hold a sample)
  A.replay_stride = P->replay_stride;
  A.n_sets = 1;
  A.n_steps = P->n_steps;
  A.n_particles = P->n_particles;
  A.smooth_unif = P->smooth_unif;
  A.out_layout = 0;
  A.choice_bytes = P->choice_bytes;
  A.refill_min = 1;
  A.max_walks = 64;
  A.queue_block = 1;
  A.dt = P->dt;
  fill_t_fast(P, A);
  A.k0 = (uint32_t)P->seed;
  A.k1 = (uint32_t)(P->seed >> 32);
  hipLaunchKernelGGL(fn, dim3(1), dim3(64), 0, stream, A);
  SSM_HIP(hipGetLastError());
  return SSM_OK;
}

}  // namespace ssmhip
