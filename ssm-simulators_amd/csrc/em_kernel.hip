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
// in float32 with the same operation order (this file is compiled with
// -ffp-contract=off so no multiply-add is fused behind our back).
//
// Normals: native mode draws Philox4x32-10 + Box–Muller with counter
// (sample, set, block, stream); replay mode reads a caller buffer.

#include <hip/hip_runtime.h>

#include "ssmhip_internal.h"

namespace ssmhip {

struct EmArgs {
  const float* set_params;
  const int64_t* offsets;
  const float* table;
  const float* replay_normals;
  const float* replay_uniforms;
  float* rts;
  void* choices;
  unsigned long long* ws;  // [0] queue head, [1] executed steps, [2] lane-step slots issued
  int64_t total;
  int64_t replay_stride;
  int32_t n_sets;
  int32_t n_steps;
  int32_t n_particles;
  int32_t smooth_unif;
  int32_t out_layout;
  int32_t choice_bytes;
  int32_t refill_min;
  int32_t queue_block;  // largest index range a wave claims from the global queue at once
  float dt;
  uint32_t k0, k1;
};

enum : int { kRunning = 0, kDone = 1, kEmpty = 2 };

// Boundary height at step index ix (t_s[ix] = float32(ix) * dt exactly, as
// np.arange(0, max_t + dt, dt).astype(float32) in _utils.pyx:636-640).
template <int BOUND>
__device__ __forceinline__ float boundary_at(int ix, float dt, float a, float b1, float b2,
                                             const float* __restrict__ row) {
  if constexpr (BOUND == SSM_BOUNDARY_CONSTANT) {
    return a;
  } else if constexpr (BOUND == SSM_BOUNDARY_ANGLE) {
    // boundary_functions.py:57-72: a + t * (-sin(theta)/cos(theta)), slope precomputed in float32
    const float ts = (float)ix * dt;
    return a + ts * b1;
  } else if constexpr (BOUND == SSM_BOUNDARY_WEIBULL) {
    // boundary_functions.py:103-122: a * exp(-(t/beta)^alpha); b1 = alpha, b2 = 1/beta
    const float ts = (float)ix * dt;
    const float u = ts * b2;
    const float p = __builtin_amdgcn_exp2f(b1 * __builtin_amdgcn_logf(u));  // u == 0 -> 0
    return a * __builtin_amdgcn_exp2f(-1.4426950408889634f * p);
  } else {
    return row[ix];
  }
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
template <int MODEL>
__device__ __forceinline__ float step_increment(float y, float vdt, float v, float leak, float dt, float sst,
                                                float nz) {
  if constexpr (MODEL == SSM_MODEL_ORNSTEIN) return ((v - (leak * y)) * dt) + (sst * nz);
  else return vdt + sst * nz;
}

// the single-particle diffusions that share the DDM walk
template <int MODEL>
constexpr bool kDiffusion = MODEL == SSM_MODEL_DDM || MODEL == SSM_MODEL_FULL_DDM || MODEL == SSM_MODEL_ORNSTEIN;

template <int MODEL, int BOUND, bool REPLAY, int MAXP>
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
  float a = 0.f, b1 = 0.f, b2 = 0.f, t0 = 0.f, deadline = 0.f, dl_tmp = 0.f, leak = 0.f, inh = 0.f;
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
    else maint = n_idle >= A.refill_min;
    if (maint) {
      if (state == kDone) {
        // result of the finished walk (ddm_models.pyx:735-744, race_models.pyx:454-464)
        float su = 0.f;
        if (A.smooth_unif) {
          float u;
          if constexpr (REPLAY) {
            // full DDM rows carry (u_t, u_z, u_smooth); the others one value per trajectory
            u = MODEL == SSM_MODEL_FULL_DDM ? A.replay_uniforms[g * 3 + 2] : A.replay_uniforms[g];
          } else {
            u32x4 r = philox4x32_10(u32x4{(uint32_t)n, (uint32_t)k, 0u, 1u}, A.k0, A.k1);
            u = u32_to_unit(r.x);
          }
          // compute_smooth_unif, _utils.pyx:668-695
          // the 0.5 literals are C doubles: formed in double, rounded to float once
          if (t == 0.f) su = (float)((double)u * 0.5 * (double)dt);
          else if (t < dl_tmp) su = (float)((0.5 - (double)u) * (double)dt);
          else su = 0.f;
        }
        float rt = t + t0 + su;
        if (rt >= deadline || deadline <= 0.f) rt = -999.f;  // enforce_deadline, _utils.pyx:697-710
        int choice;
        if constexpr (kDiffusion<MODEL>) {
          choice = (x[0] > 0.f) - (x[0] < 0.f);  // sign(), _utils.pyx:474-484
        } else {
          choice = 0;  // np.argmax: first maximum
          float best = x[0];
#pragma unroll
          for (int j = 1; j < MAXP; ++j)
            if (j < P && x[j] > best) { best = x[j]; choice = j; }
        }
        const int64_t o = A.out_layout == 0 ? g : (int64_t)n * A.n_sets + k;
        if (o < A.total) {  // layout 1 assumes uniform offsets; never write past the buffers
          A.rts[o] = rt;
          if (A.choice_bytes == 1) ((int8_t*)A.choices)[o] = (int8_t)choice;
          else ((int32_t*)A.choices)[o] = choice;
        }
        my_steps += (unsigned long long)ix;
        state = kEmpty;
      }
      if (!exhausted) {
        const uint64_t want = __ballot(state == kEmpty);
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
          if (nb_base < nb_end) nb_k = find_set(A.offsets, A.n_sets, nb_base);
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
          while (q_k < A.n_sets - 1 && A.offsets[q_k + 1] <= q_lo) ++q_k;  // keep q_k = set of q_lo
        if (q_lo >= q_hi && queue_done) exhausted = true;
        if (state == kEmpty) {
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
            g = gi;
            k = kk;
            while (k < A.n_sets - 1 && A.offsets[k + 1] <= gi) ++k;
            n = (int)(gi - A.offsets[k]);
            const float* sp = A.set_params + (int64_t)k * SSM_SET_STRIDE;
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
              if constexpr (MODEL == SSM_MODEL_ORNSTEIN) {
                // ornstein_models.pyx:165: the literal 2 is an int, so this stays in float
                x[0] = (-b0) + ((z * 2.0f) * b0);
                leak = sp[SSM_P_G];
                vraw[0] = sp[SSM_P_V];
              } else {
                // ddm_models.pyx:719 / :2291: the literal 2.0 is a C double, so the sum is
                // formed in double and rounded to float once
                x[0] = (float)((double)(-b0) + (double)z * 2.0 * (double)b0);
              }
              if constexpr (MODEL == SSM_MODEL_FULL_DDM) {
                // trial-to-trial variability, ddm_models.pyx:2294-2313: drift v + sv g,
                // non-decision time t + 2 (u_t - 1/2) st, start + 2 (u_z - 1/2) sz; the
                // uniform terms are formed in double (0.5 is a C double)
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
                t0 = (float)((double)t0 + (2.0 * ((double)ut - 0.5)) * (double)sp[SSM_P_ST]);
                x[0] = (float)((double)x[0] + (2.0 * ((double)uz - 0.5)) * (double)sp[SSM_P_SZ]);
              }
            } else {
              leak = sp[SSM_P_G];
              inh = sp[SSM_P_BINH];
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
            }
            t = 0.f;
            ix = 0;
            state = kRunning;
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
    if (state == kRunning) {
      if constexpr (kDiffusion<MODEL>) {
        // four steps per Philox block; every walk starts block-aligned
        float nz[4];
        if constexpr (REPLAY) {
          // a full-DDM row starts with the drift draw, the walk's draws follow it
          constexpr int off = MODEL == SSM_MODEL_FULL_DDM ? 1 : 0;
          const float* rb = A.replay_normals + g * A.replay_stride;
#pragma unroll
          for (int i = 0; i < 4; ++i) nz[i] = rb[(int64_t)(off + ix + i) % A.replay_stride];
        } else {
          philox_normals4((uint32_t)n, (uint32_t)k, (uint32_t)(ix >> 2), 0u, A.k0, A.k1, nz);
        }
        bool alive = true;
        // Steps use exec-mask branches rather than selects: with every lane usually on the
        // same side, a branch costs scalar instructions while a select costs VALU issue slots.
        if (__ballot(ix + 4 > A.n_steps) == 0) {
          // common case: no lane can reach the end of t_s in this block, no index guard needed
#pragma unroll
          for (int i = 0; i < 4; ++i) {
            if (alive) {
              const float b = boundary_at<BOUND>(ix, dt, a, b1, b2, row);
              if ((x[0] >= -b) && (x[0] <= b) && (t <= dl_tmp)) {
                x[0] += step_increment<MODEL>(x[0], vdt[0], vraw[0], leak, dt, sst[0], nz[i]);
                t += dt;
                ++ix;
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
                go = (x[0] >= -b) && (x[0] <= b) && (t <= dl_tmp);
              }
              if (go) {
                x[0] += step_increment<MODEL>(x[0], vdt[0], vraw[0], leak, dt, sst[0], nz[i]);
                t += dt;
                ++ix;
              } else {
                alive = false;
              }
            }
          }
        }
        if (!alive) state = kDone;
      } else {
        bool go = ix < A.n_steps && t <= dl_tmp;
        if (go) {
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
          if constexpr (MODEL == SSM_MODEL_RACE) {
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
                const float drift = ((vraw[j] - (leak * x[j])) - (inh * red)) * dt;
                xnew[j] = fmaxf(0.0f, x[j] + (drift + sst[j] * nz[j]));  // :852-854
              } else {
                xnew[j] = x[j];
              }
            }
#pragma unroll
            for (int j = 0; j < MAXP; ++j) x[j] = xnew[j];
          }
          t += dt;
          ++ix;
        }
      }
    }
  }

  // one atomic per wavefront for each counter
  unsigned long long s = my_steps;
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
  if (lane_id() == 0) {
    if (s) atomicAdd(A.ws + 1, s);
    atomicAdd(A.ws + 2, iters * (MODEL == SSM_MODEL_DDM ? 256ull : 64ull));
  }
}

__global__ void ws_reset_kernel(unsigned long long* ws) {
  if (threadIdx.x < SSM_EM_WORKSPACE_WORDS) ws[threadIdx.x] = 0ull;
}

// --------------------------------------------------------------------------------
using KernelFn = void (*)(EmArgs);

template <int MODEL, int BOUND, bool REPLAY>
static KernelFn pick_p(int p) {
  if constexpr (kDiffusion<MODEL>) {
    return em_kernel<MODEL, BOUND, REPLAY, 1>;
  } else {
    if (p <= 4) return em_kernel<MODEL, BOUND, REPLAY, 4>;
    if (p <= 8) return em_kernel<MODEL, BOUND, REPLAY, 8>;
    return em_kernel<MODEL, BOUND, REPLAY, 16>;
  }
}

template <int MODEL, bool REPLAY>
static KernelFn pick_bound(int bound, int p) {
  switch (bound) {
    case SSM_BOUNDARY_CONSTANT: return pick_p<MODEL, SSM_BOUNDARY_CONSTANT, REPLAY>(p);
    case SSM_BOUNDARY_ANGLE: return pick_p<MODEL, SSM_BOUNDARY_ANGLE, REPLAY>(p);
    case SSM_BOUNDARY_WEIBULL: return pick_p<MODEL, SSM_BOUNDARY_WEIBULL, REPLAY>(p);
    default: return pick_p<MODEL, SSM_BOUNDARY_TABLE, REPLAY>(p);
  }
}

static KernelFn pick_kernel(int model, int bound, int p, bool replay) {
  if (model == SSM_MODEL_DDM) return replay ? pick_bound<SSM_MODEL_DDM, true>(bound, p) : pick_bound<SSM_MODEL_DDM, false>(bound, p);
  if (model == SSM_MODEL_FULL_DDM)
    return replay ? pick_bound<SSM_MODEL_FULL_DDM, true>(bound, p) : pick_bound<SSM_MODEL_FULL_DDM, false>(bound, p);
  if (model == SSM_MODEL_ORNSTEIN)
    return replay ? pick_bound<SSM_MODEL_ORNSTEIN, true>(bound, p) : pick_bound<SSM_MODEL_ORNSTEIN, false>(bound, p);
  if (model == SSM_MODEL_RACE) return replay ? pick_bound<SSM_MODEL_RACE, true>(bound, p) : pick_bound<SSM_MODEL_RACE, false>(bound, p);
  return replay ? pick_bound<SSM_MODEL_LCA, true>(bound, p) : pick_bound<SSM_MODEL_LCA, false>(bound, p);
}

int32_t em_validate(const ssm_em_problem* P, const float* set_params, const int64_t* offsets,
                    const float* table, const float* rn, const float* ru, float* rts, void* choices,
                    uint64_t* ws) {
  if (!P) return fail(SSM_EINVAL, "problem descriptor is NULL");
  if (P->model < SSM_MODEL_DDM || P->model > SSM_MODEL_ORNSTEIN) return fail(SSM_EINVAL, "unknown model");
  if (P->boundary < SSM_BOUNDARY_CONSTANT || P->boundary > SSM_BOUNDARY_TABLE) return fail(SSM_EINVAL, "unknown boundary");
  const bool diffusion = P->model == SSM_MODEL_DDM || P->model == SSM_MODEL_FULL_DDM || P->model == SSM_MODEL_ORNSTEIN;
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
  if (P->out_layout == 1 && P->n_sets > 0 && P->n_total % P->n_sets != 0)
    return fail(SSM_EINVAL, "out_layout 1 needs the same n_samples for every set");
  if (P->n_total > 0) {
    if (!set_params || !offsets || !rts || !choices || !ws) return fail(SSM_EINVAL, "NULL buffer");
    if (P->boundary == SSM_BOUNDARY_TABLE && !table) return fail(SSM_EINVAL, "TABLE boundary needs a table");
    if (P->replay_stride > 0 && !rn) return fail(SSM_EINVAL, "replay mode needs a normals buffer");
    if (P->replay_stride > 0 && P->smooth_unif && !ru) return fail(SSM_EINVAL, "replay + smooth_unif needs uniforms");
    if (P->replay_stride > 0 && P->model == SSM_MODEL_FULL_DDM && !ru)
      return fail(SSM_EINVAL, "full-DDM replay needs uniforms (u_t, u_z, u_smooth per trajectory)");
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
  hipLaunchKernelGGL(ws_reset_kernel, dim3(1), dim3(64), 0, stream, (unsigned long long*)ws);
  SSM_HIP(hipGetLastError());
  if (P->n_total == 0) return SSM_OK;

  EmArgs A;
  A.set_params = set_params;
  A.offsets = offsets;
  A.table = table;
  A.replay_normals = rn;
  A.replay_uniforms = ru;
  A.rts = rts;
  A.choices = choices;
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
  A.k0 = (uint32_t)P->seed;
  A.k1 = (uint32_t)(P->seed >> 32);

  const int cus = device_cu_count();
  if (cus <= 0) return fail(SSM_EHIP, "no HIP device");
  int64_t grid = (int64_t)cus * per_cu;
  const int64_t needed = (P->n_total + block - 1) / block;
  if (grid > needed) grid = needed;
  if (grid < 1) grid = 1;
  KernelFn fn = pick_kernel(P->model, P->boundary, P->n_particles, P->replay_stride > 0);
  hipLaunchKernelGGL(fn, dim3((unsigned)grid), dim3(block), 0, stream, A);
  SSM_HIP(hipGetLastError());
  return SSM_OK;
}

}  // namespace ssmhip
