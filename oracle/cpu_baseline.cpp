// CPU baseline and fast CPU oracle (test infrastructure only).
//
// C++/OpenMP restatement of the reference's parallel EM path:
//   ddm_flexbound parallel branch  src/cssm/ddm_models.pyx:750-842
//   race_model parallel branch     src/cssm/race_models.pyx:206-316
//   lca parallel branch            src/cssm/race_models.pyx:595-717
// A flattened (set, sample) loop under `omp for schedule(dynamic)`; every
// trajectory re-seeds a per-thread MT19937 with the seed mix of
// src/cssm/gsl_rng.h:97-102 and draws normals with a Ziggurat written from
// Marsaglia & Tsang, "The Ziggurat Method for Generating Random Variables",
// J. Stat. Software 5(8), 2000 (no GSL code). Boundaries come from a
// precomputed per-set table, as the reference's parallel branch does.
// Walk semantics follow the sequential loop (inclusive bounds; the parallel
// branch's break test is the same set), with optional smooth_unif.
// With replay_normals != NULL the same walk consumes a caller buffer
// (draw d of trajectory i = normals[i * stride + d % stride]); that mode is
// pinned bit-exactly against oracle/em_numpy.py.
#include <math.h>
#include <omp.h>
#include <stdint.h>
#include <string.h>

#include <random>

namespace {

uint64_t mix_seed(uint64_t base, uint64_t t1, uint64_t t2) {
  uint64_t z = base + t1 * 0x9E3779B97F4A7C15ULL + t2 * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

// Ziggurat tables for the standard normal, 128 layers (Marsaglia & Tsang 2000).
struct ZigTables {
  uint32_t kn[128];
  float wn[128], fn[128];
  ZigTables() {
    const double m1 = 2147483648.0;
    double dn = 3.442619855899, tn = dn;
    const double vn = 9.91256303526217e-3;
    const double q = vn / exp(-.5 * dn * dn);
    kn[0] = (uint32_t)((dn / q) * m1);
    kn[1] = 0;
    wn[0] = (float)(q / m1);
    wn[127] = (float)(dn / m1);
    fn[0] = 1.0f;
    fn[127] = (float)exp(-.5 * dn * dn);
    for (int i = 126; i >= 1; --i) {
      dn = sqrt(-2. * log(vn / dn + exp(-.5 * dn * dn)));
      kn[i + 1] = (uint32_t)((dn / tn) * m1);
      tn = dn;
      fn[i] = (float)exp(-.5 * dn * dn);
      wn[i] = (float)(dn / m1);
    }
  }
};
const ZigTables& zig() {
  static ZigTables t;
  return t;
}

struct Rng {
  std::mt19937 mt;
  const ZigTables* T;
  void seed(uint64_t s) {
    uint32_t s32 = (uint32_t)s;  // gsl_rng_set(mt19937) keeps the low 32 bits; 0 -> 4357
    mt.seed(s32 ? s32 : 4357u);
  }
  float uni() { return ((float)(mt() >> 8) + 0.5f) * (1.0f / 16777216.0f); }  // (0,1)
  float normal() {
    const ZigTables& Z = *T;
    const float r = 3.442620f;
    int32_t hz = (int32_t)mt();
    uint32_t iz = hz & 127;
    if ((uint32_t)(hz < 0 ? -(int64_t)hz : hz) < Z.kn[iz]) return hz * Z.wn[iz];
    for (;;) {
      float x = hz * Z.wn[iz];
      if (iz == 0) {
        float xx, y;
        do {
          xx = -logf(uni()) * 0.2904764f;
          y = -logf(uni());
        } while (y + y < xx * xx);
        return hz > 0 ? r + xx : -r - xx;
      }
      if (Z.fn[iz] + uni() * (Z.fn[iz - 1] - Z.fn[iz]) < expf(-.5f * x * x)) return x;
      hz = (int32_t)mt();
      iz = hz & 127;
      if ((uint32_t)(hz < 0 ? -(int64_t)hz : hz) < Z.kn[iz]) return hz * Z.wn[iz];
    }
  }
};

}  // namespace

extern "C" {

// model: 0 DDM, 1 race, 2 LCA. Arrays are per set: v/z/sst [n_sets * P],
// t0/deadline/dl_tmp/g/b [n_sets], table [n_sets * n_steps].
// Outputs are set-major: rts[offsets[k] + n]. Returns total executed steps.
int64_t cpu_em_simulate(int model, int P, int n_sets, const int64_t* offsets, const float* v, const float* z,
                        const float* sst, const float* t0, const float* deadline, const float* dl_tmp,
                        const float* g, const float* b, const float* table, int n_steps, float dt,
                        int smooth_unif, uint64_t seed, const float* replay_normals, int64_t stride,
                        const float* replay_uniforms, float* rts, int32_t* choices, int n_threads) {
  const int64_t total = offsets[n_sets];
  int64_t steps_total = 0;
  if (n_threads > 0) omp_set_num_threads(n_threads);
#pragma omp parallel reduction(+ : steps_total)
  {
    Rng rng;
    rng.T = &zig();
    float x[16], xn[16];
    int k = 0;
#pragma omp for schedule(dynamic, 64)
    for (int64_t i = 0; i < total; ++i) {
      while (k > 0 && offsets[k] > i) --k;
      while (k < n_sets - 1 && offsets[k + 1] <= i) ++k;
      const int64_t n = i - offsets[k];
      if (!replay_normals) rng.seed(mix_seed(seed, (uint64_t)k, (uint64_t)n));
      const float* row = table + (int64_t)k * n_steps;
      const float b0 = row[0];
      const float dlt = dl_tmp[k];
      float t = 0.f;
      int ix = 0;
      int64_t d = 0;
      auto draw = [&]() -> float {
        if (replay_normals) {
          const float val = replay_normals[i * stride + (d % stride)];
          ++d;
          return val;
        }
        return rng.normal();
      };
      int choice = 0;
      if (model == 0) {
        const float vdt = v[k] * dt, s = sst[k];
        // replay mode restates the sequential loop, whose start (ddm_models.pyx:709, int 2) is
        // float arithmetic; the native mode restates the parallel branch, whose start (:798,
        // double literal 2.0) is formed in double and rounded once
        const float y_seq = (-b0) + ((z[k] * 2.0f) * b0);
        const float y_par = (float)((double)(-b0) + (double)z[k] * 2.0 * (double)b0);
        float y = replay_normals ? y_seq : y_par;
        for (;;) {
          if (ix >= n_steps) break;
          const float bb = row[ix];
          if (!(y >= -bb && y <= bb && t <= dlt)) break;
          const float nz = draw();
          y = y + (vdt + s * nz);
          t = t + dt;
          ++ix;
        }
        choice = (y > 0.f) - (y < 0.f);
      } else {
        const float* vk = v + (int64_t)k * P;
        const float* sk = sst + (int64_t)k * P;
        const float* zk = z + (int64_t)k * P;
        for (int j = 0; j < P; ++j) x[j] = zk[j] * b0;
        for (;;) {
          if (ix >= n_steps) break;
          const float bb = row[ix];
          bool crossed = false;
          for (int j = 0; j < P; ++j) crossed |= x[j] > bb;
          if (crossed || !(t <= dlt)) break;
          if (model == 1) {
            for (int j = 0; j < P; ++j) {
              const float nz = draw();
              x[j] = fmaxf(0.0f, x[j] + (vk[j] * dt + sk[j] * nz));
            }
          } else {
            float sum = 0.f;
            for (int j = 0; j < P; ++j) sum += x[j];
            for (int j = 0; j < P; ++j) {
              const float nz = draw();
              const float red = (-x[j]) + sum;
              const float drift = ((vk[j] - (g[k] * x[j])) - (b[k] * red)) * dt;
              xn[j] = fmaxf(0.0f, x[j] + (drift + sk[j] * nz));
            }
            for (int j = 0; j < P; ++j) x[j] = xn[j];
          }
          t = t + dt;
          ++ix;
        }
        float best = x[0];
        choice = 0;
        for (int j = 1; j < P; ++j)
          if (x[j] > best) { best = x[j]; choice = j; }
      }
      float su = 0.f;
      if (smooth_unif) {
        const float u = replay_uniforms ? replay_uniforms[i] : rng.uni();
        if (t == 0.f) su = (float)((double)u * 0.5 * (double)dt);  // 0.5 is a C double
        else if (t < dlt) su = (float)((0.5 - (double)u) * (double)dt);
      }
      float rt = t + t0[k] + su;
      if (rt >= deadline[k] || deadline[k] <= 0.f) rt = -999.f;
      rts[i] = rt;
      choices[i] = choice;
      steps_total += ix;
    }
  }
  return steps_total;
}

// n standard normals from the Ziggurat generator seeded like one trajectory.
void cpu_ziggurat_normals(uint64_t seed, uint64_t k, uint64_t n, int64_t count, float* out) {
  Rng rng;
  rng.T = &zig();
  rng.seed(mix_seed(seed, k, n));
  for (int64_t i = 0; i < count; ++i) out[i] = rng.normal();
}

uint64_t cpu_mix_seed(uint64_t base, uint64_t t1, uint64_t t2) { return mix_seed(base, t1, t2); }

// Exact float64 Gaussian-kernel log-sum for KDE evaluation:
// out[i] = log sum_j exp(-(x_i - s_j)^2 / (2 h^2)) (max-shifted).
void cpu_kde_logsum(const double* x, int64_t m, const double* s, int64_t n, double h, double* out, int n_threads) {
  if (n_threads > 0) omp_set_num_threads(n_threads);
  const double c = -0.5 / (h * h);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < m; ++i) {
    double mx = -1e300;
    for (int64_t j = 0; j < n; ++j) {
      const double d = x[i] - s[j];
      const double e = c * d * d;
      if (e > mx) mx = e;
    }
    double acc = 0.0;
    for (int64_t j = 0; j < n; ++j) {
      const double d = x[i] - s[j];
      acc += exp(c * d * d - mx);
    }
    out[i] = (n > 0) ? mx + log(acc) : -1e300;
  }
}

}  // extern "C"
