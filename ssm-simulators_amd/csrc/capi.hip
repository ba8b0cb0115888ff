// extern "C" entry points of libssmhip (declared in include/ssmhip.h).
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>
#include <vector>

#include "ssmhip_internal.h"

namespace ssmhip {

static thread_local std::string g_last_error;

void set_error(const std::string& msg) { g_last_error = msg; }

int32_t fail(int32_t code, const std::string& msg) {
  set_error(msg);
  return code;
}

int32_t hip_fail(hipError_t e, const char* what) {
  return fail(SSM_EHIP, std::string(what) + ": " + hipGetErrorString(e));
}

int device_cu_count() {
  static std::mutex mu;
  static std::vector<int> cache;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 0;
  std::lock_guard<std::mutex> lock(mu);
  if ((int)cache.size() <= dev) cache.resize(dev + 1, 0);
  if (cache[dev] == 0) {
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 0;
    cache[dev] = cus;
  }
  return cache[dev];
}

int32_t em_validate(const ssm_em_problem* P, const float* set_params, const int64_t* offsets,
                    const float* table, const float* rn, const float* ru, float* rts, void* choices,
                    uint64_t* ws);
int32_t em_enqueue(const ssm_em_problem* P, const ssm_launch_config* L, const float* set_params,
                   const int64_t* offsets, const float* table, const float* rn, const float* ru,
                   float* rts, void* choices, uint64_t* ws, hipStream_t stream);
int32_t em_time_grid(float dt, int32_t n_steps, hipStream_t stream, const float** out);
int32_t em_trajectory_enqueue(const ssm_em_problem* P, const float* set_params, const int64_t* offsets,
                              const float* table, const float* rn, const float* ru, float* traj, int32_t traj_rows,
                              uint64_t* ws, hipStream_t stream);
int32_t lba_enqueue(int32_t nact, int32_t n_sets, int64_t n_total, int32_t out_layout, int32_t choice_bytes,
                    uint64_t seed, const float* params, const int64_t* offsets, const float* replay,
                    float* rts, void* choices, hipStream_t stream);
int32_t philox_enqueue(uint64_t seed, uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, int64_t n,
                       uint32_t* raw, float* normals, hipStream_t stream);

int32_t probe_enqueue(int32_t kind, int64_t iters, int64_t n_threads, float* out, hipStream_t stream);

int32_t kde_fit_enqueue(int32_t n_sets, int32_t n_choices, const int32_t* choice_values, const float* rts,
                        const void* choices, int32_t choice_bytes, const int64_t* offsets, float* samples,
                        int32_t* kde_count, int32_t* choice_total, float* bandwidth, int64_t* kde_start,
                        hipStream_t stream);
int32_t kde_sample_enqueue(int32_t n_sets, int32_t n_choices, int32_t n_kde, int32_t n_unif, float max_t,
                           float unif_lo, uint64_t seed, const float* samples, const int32_t* kde_count,
                           const int32_t* choice_total, const float* bandwidth, const int64_t* kde_start,
                           const int64_t* offsets, const int32_t* n_by_override, float* pts_rt,
                           int8_t* pts_choice, int64_t* grp_begin, int32_t* grp_count, hipStream_t stream);
int32_t kde_eval_enqueue(int32_t n_sets, int32_t n_choices, const int64_t* grp_begin, const int32_t* grp_count,
                         const float* pts_rt, const float* samples, const int32_t* kde_count,
                         const int32_t* choice_total, const float* bandwidth, const int64_t* kde_start,
                         const int64_t* offsets, float max_t, float lb, float eps, int64_t max_samples,
                         float* workspace, float* out_ll, hipStream_t stream);
int32_t aux_labels_enqueue(int32_t n_sets, int32_t n_choices, const int32_t* choice_values_dev,
                           int32_t max_choice, const float* rts, const void* choices, int32_t choice_bytes,
                           const int64_t* offsets, double max_t, int32_t* counts, int32_t* nogo, int32_t* hist128,
                           int32_t* hist256, hipStream_t s);
int32_t theta_uniform_enqueue(int32_t n_sets, int32_t n_params, const double* lo, const double* hi, uint64_t seed,
                              float* theta, hipStream_t s);
int32_t filter_workspace_bytes(int32_t n_sets, int64_t n_total, size_t* bytes);
int32_t filter_stats_enqueue(int32_t n_sets, int32_t n_choices, const int32_t* choice_values, const float* rts,
                             const void* choices, int32_t choice_bytes, const int64_t* offsets, int64_t n_total,
                             void* ws, size_t ws_bytes, double* stats, hipStream_t stream);
int32_t kde_eval_noalloc_enqueue(int32_t n_sets, int32_t n_choices, const int64_t* grp_begin,
                                 const int32_t* grp_count, const float* pts_rt, const float* samples,
                                 const int32_t* kde_count, const int32_t* choice_total, const float* bandwidth,
                                 const int64_t* kde_start, const int64_t* offsets, float max_t, float lb, float eps,
                                 int64_t max_samples, int64_t n_points, float* out_ll, hipStream_t stream);
int32_t lan_assemble_enqueue(int32_t n_sets, int32_t n_theta, int32_t m, int32_t n_neg, int32_t n_choices,
                             float neg_ll, int32_t shuffle, uint64_t seed, const float* theta, const float* pts_rt,
                             const int8_t* pts_choice, const float* ll, const int32_t* choice_values, float* rows,
                             hipStream_t stream);
int32_t copy_rows_enqueue(int64_t n_rows, int64_t row_bytes, const void* src, const int64_t* src_idx, void* dst,
                          const int64_t* dst_idx, hipStream_t stream);
int32_t em_records_enqueue(int32_t n_sets, int32_t set_stride, int32_t n_slots, const int32_t* slots,
                           const float* columns, float* records, hipStream_t stream);
int32_t uniform_offsets_enqueue(int32_t n_sets, int64_t n_samples, int64_t* offsets, hipStream_t stream);

int32_t mlp_sizes(int32_t* n_params, int32_t* n_frag, int32_t* n_bias, int32_t* act_rows, int32_t* delta_rows);
int32_t mlp_pack_enqueue(const float* params, __bf16* frags, float* bias, hipStream_t s);
int32_t mlp_forward_enqueue(const float* x, int64_t n, const __bf16* frags, const float* bias, const float* xnorm,
                            float* y, hipStream_t s);
int32_t mlp_train_enqueue(const TrainArgs& A, hipStream_t s);
int32_t mlp_wgrad_enqueue(const __bf16* acts, const __bf16* deltas, int64_t n, int64_t ld, float* grad,
                          hipStream_t s);
int64_t mlp_train_workspace_bytes(int64_t batch);
int32_t mlp_train_steps_enqueue(const float* x, const float* target, int64_t batch, int32_t n_steps, int32_t step0,
                                const float* lr, float b1, float b2, float eps, float wd, float delta, float* P,
                                float* M, float* V, __bf16* frags, float* bias, const float* xnorm, void* ws,
                                int64_t ws_bytes, float* loss, hipStream_t s);
int32_t mlp_adam_enqueue(float* P, const float* G, float* M, float* V, int32_t step, float lr, float b1, float b2,
                         float eps, float wd, hipStream_t s);
int32_t mfma_probe_enqueue(const void* a, const void* b, float* d, hipStream_t s);
int32_t mfma32_probe_enqueue(const void* a, const void* b, float* d, hipStream_t s);


struct GraphHandle {
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

}  // namespace ssmhip

using namespace ssmhip;

extern "C" {

int32_t ssm_abi_version(void) { return 6; }

const char* ssm_last_error(void) { return g_last_error.c_str(); }

int32_t ssm_device_cu_count(void) {
  const int c = device_cu_count();
  if (c <= 0) set_error("no HIP device available");
  return c;
}

int32_t ssm_em_simulate(const ssm_em_problem* prob, const ssm_launch_config* launch, const float* set_params_dev,
                        const int64_t* offsets_dev, const float* table_dev, const float* replay_normals_dev,
                        const float* replay_uniforms_dev, float* rts_dev, void* choices_dev,
                        uint64_t* workspace_dev, void* stream) {
  int32_t rc = em_validate(prob, set_params_dev, offsets_dev, table_dev, replay_normals_dev,
                           replay_uniforms_dev, rts_dev, choices_dev, workspace_dev);
  if (rc) return rc;
  if (!workspace_dev) return fail(SSM_EINVAL, "workspace is NULL");
  return em_enqueue(prob, launch, set_params_dev, offsets_dev, table_dev, replay_normals_dev, replay_uniforms_dev,
                    rts_dev, choices_dev, workspace_dev, (hipStream_t)stream);
}

int32_t ssm_em_trajectory(const ssm_em_problem* prob, const float* set_params_dev, const int64_t* offsets_dev,
                          const float* table_dev, const float* replay_normals_dev, const float* replay_uniforms_dev,
                          float* traj_dev, int32_t traj_rows, uint64_t* workspace_dev, void* stream) {
  return em_trajectory_enqueue(prob, set_params_dev, offsets_dev, table_dev, replay_normals_dev, replay_uniforms_dev,
                               traj_dev, traj_rows, workspace_dev, (hipStream_t)stream);
}

int32_t ssm_em_graph_create(const ssm_em_problem* prob, const ssm_launch_config* launch,
                            const float* set_params_dev, const int64_t* offsets_dev, const float* table_dev,
                            const float* replay_normals_dev, const float* replay_uniforms_dev, float* rts_dev,
                            void* choices_dev, uint64_t* workspace_dev, void** graph_out) {
  if (!graph_out) return fail(SSM_EINVAL, "graph_out is NULL");
  *graph_out = nullptr;
  int32_t rc = em_validate(prob, set_params_dev, offsets_dev, table_dev, replay_normals_dev,
                           replay_uniforms_dev, rts_dev, choices_dev, workspace_dev);
  if (rc) return rc;
  if (!workspace_dev) return fail(SSM_EINVAL, "workspace is NULL");
  if (prob->n_total > 0) {  // the launch's time table is built here, not inside the capture
    const float* t_grid = nullptr;
    rc = em_time_grid(prob->dt, prob->n_steps, nullptr, &t_grid);
    if (rc) return rc;
  }
  hipStream_t s = nullptr;
  SSM_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  GraphHandle* h = new GraphHandle();
  hipError_t e = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal);
  if (e != hipSuccess) {
    delete h;
    (void)hipStreamDestroy(s);
    return hip_fail(e, "hipStreamBeginCapture");
  }
  rc = em_enqueue(prob, launch, set_params_dev, offsets_dev, table_dev, replay_normals_dev, replay_uniforms_dev,
                  rts_dev, choices_dev, workspace_dev, s);
  e = hipStreamEndCapture(s, &h->graph);
  if (rc == SSM_OK && e != hipSuccess) rc = hip_fail(e, "hipStreamEndCapture");
  if (rc == SSM_OK) {
    e = hipGraphInstantiate(&h->exec, h->graph, nullptr, nullptr, 0);
    if (e != hipSuccess) rc = hip_fail(e, "hipGraphInstantiate");
  }
  (void)hipStreamDestroy(s);
  if (rc != SSM_OK) {
    if (h->exec) (void)hipGraphExecDestroy(h->exec);
    if (h->graph) (void)hipGraphDestroy(h->graph);
    delete h;
    return rc;
  }
  *graph_out = h;
  return SSM_OK;
}

int32_t ssm_graph_launch(void* graph, void* stream) {
  if (!graph) return fail(SSM_EINVAL, "graph is NULL");
  GraphHandle* h = (GraphHandle*)graph;
  SSM_HIP(hipGraphLaunch(h->exec, (hipStream_t)stream));
  return SSM_OK;
}

int32_t ssm_graph_destroy(void* graph) {
  if (!graph) return SSM_OK;
  GraphHandle* h = (GraphHandle*)graph;
  if (h->exec) (void)hipGraphExecDestroy(h->exec);
  if (h->graph) (void)hipGraphDestroy(h->graph);
  delete h;
  return SSM_OK;
}

int32_t ssm_lba_simulate(int32_t nact, int32_t n_sets, int64_t n_total, int32_t out_layout, int32_t choice_bytes,
                         uint64_t seed, const float* params_dev, const int64_t* offsets_dev,
                         const float* replay_dev, float* rts_dev, void* choices_dev, void* stream) {
  return lba_enqueue(nact, n_sets, n_total, out_layout, choice_bytes, seed, params_dev, offsets_dev, replay_dev,
                     rts_dev, choices_dev, (hipStream_t)stream);
}

int32_t ssm_philox_fill(uint64_t seed, uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, int64_t n,
                        uint32_t* raw_dev, float* normals_dev, void* stream) {
  return philox_enqueue(seed, c0, c1, c2, c3, n, raw_dev, normals_dev, (hipStream_t)stream);
}

int32_t ssm_kde_fit(int32_t n_sets, int32_t n_choices, const int32_t* choice_values_dev, const float* rts_dev,
                    const void* choices_dev, int32_t choice_bytes, const int64_t* offsets_dev, float* samples_dev,
                    int32_t* kde_count_dev, int32_t* choice_total_dev, float* bandwidth_dev,
                    int64_t* kde_start_dev, void* stream) {
  return kde_fit_enqueue(n_sets, n_choices, choice_values_dev, rts_dev, choices_dev, choice_bytes, offsets_dev,
                         samples_dev, kde_count_dev, choice_total_dev, bandwidth_dev, kde_start_dev,
                         (hipStream_t)stream);
}

int32_t ssm_kde_sample(int32_t n_sets, int32_t n_choices, int32_t n_kde, int32_t n_unif, float max_t, float unif_lo,
                       uint64_t seed, const float* samples_dev, const int32_t* kde_count_dev,
                       const int32_t* choice_total_dev, const float* bandwidth_dev, const int64_t* kde_start_dev,
                       const int64_t* offsets_dev, const int32_t* n_by_override_dev, float* pts_rt_dev,
                       int8_t* pts_choice_dev, int64_t* grp_begin_dev, int32_t* grp_count_dev, void* stream) {
  return kde_sample_enqueue(n_sets, n_choices, n_kde, n_unif, max_t, unif_lo, seed, samples_dev, kde_count_dev,
                            choice_total_dev, bandwidth_dev, kde_start_dev, offsets_dev, n_by_override_dev,
                            pts_rt_dev, pts_choice_dev, grp_begin_dev, grp_count_dev, (hipStream_t)stream);
}

int32_t ssm_kde_eval(int32_t n_sets, int32_t n_choices, const int64_t* grp_begin_dev, const int32_t* grp_count_dev,
                     const float* pts_rt_dev, const float* samples_dev, const int32_t* kde_count_dev,
                     const int32_t* choice_total_dev, const float* bandwidth_dev, const int64_t* kde_start_dev,
                     const int64_t* offsets_dev, float max_t, float lb, float eps, int64_t max_samples,
                     float* workspace_dev, float* out_ll_dev, void* stream) {
  return kde_eval_enqueue(n_sets, n_choices, grp_begin_dev, grp_count_dev, pts_rt_dev, samples_dev, kde_count_dev,
                          choice_total_dev, bandwidth_dev, kde_start_dev, offsets_dev, max_t, lb, eps, max_samples,
                          workspace_dev, out_ll_dev, (hipStream_t)stream);
}

int32_t ssm_kde_eval_noalloc(int32_t n_sets, int32_t n_choices, const int64_t* grp_begin_dev,
                             const int32_t* grp_count_dev, const float* pts_rt_dev, const float* samples_dev,
                             const int32_t* kde_count_dev, const int32_t* choice_total_dev,
                             const float* bandwidth_dev, const int64_t* kde_start_dev, const int64_t* offsets_dev,
                             float max_t, float lb, float eps, int64_t max_samples, int64_t n_points,
                             float* out_ll_dev, void* stream) {
  return kde_eval_noalloc_enqueue(n_sets, n_choices, grp_begin_dev, grp_count_dev, pts_rt_dev, samples_dev,
                                  kde_count_dev, choice_total_dev, bandwidth_dev, kde_start_dev, offsets_dev, max_t,
                                  lb, eps, max_samples, n_points, out_ll_dev, (hipStream_t)stream);
}

int32_t ssm_lan_assemble(int32_t n_sets, int32_t n_theta, int32_t m, int32_t n_neg, int32_t n_choices,
                         float neg_ll, int32_t shuffle, uint64_t seed, const float* theta_dev,
                         const float* pts_rt_dev, const int8_t* pts_choice_dev, const float* ll_dev,
                         const int32_t* choice_values_dev, float* rows_dev, void* stream) {
  return lan_assemble_enqueue(n_sets, n_theta, m, n_neg, n_choices, neg_ll, shuffle, seed, theta_dev, pts_rt_dev,
                              pts_choice_dev, ll_dev, choice_values_dev, rows_dev, (hipStream_t)stream);
}

int32_t ssm_em_records(int32_t n_sets, int32_t set_stride, int32_t n_slots, const int32_t* slots,
                       const float* columns_dev, float* records_dev, void* stream) {
  return em_records_enqueue(n_sets, set_stride, n_slots, slots, columns_dev, records_dev, (hipStream_t)stream);
}

int32_t ssm_uniform_offsets(int32_t n_sets, int64_t n_samples, int64_t* offsets_dev, void* stream) {
  return uniform_offsets_enqueue(n_sets, n_samples, offsets_dev, (hipStream_t)stream);
}

int32_t ssm_copy_rows(int64_t n_rows, int64_t row_bytes, const void* src_dev, const int64_t* src_idx_dev,
                      void* dst_dev, const int64_t* dst_idx_dev, void* stream) {
  return copy_rows_enqueue(n_rows, row_bytes, src_dev, src_idx_dev, dst_dev, dst_idx_dev, (hipStream_t)stream);
}

int32_t ssm_aux_labels(int32_t n_sets, int32_t n_choices, const int32_t* choice_values_dev, int32_t max_choice,
                       const float* rts_dev, const void* choices_dev, int32_t choice_bytes,
                       const int64_t* offsets_dev, double max_t, int32_t* counts_dev, int32_t* nogo_dev,
                       int32_t* hist128_dev, int32_t* hist256_dev, void* stream) {
  return aux_labels_enqueue(n_sets, n_choices, choice_values_dev, max_choice, rts_dev, choices_dev, choice_bytes,
                            offsets_dev, max_t, counts_dev, nogo_dev, hist128_dev, hist256_dev,
                            (hipStream_t)stream);
}

int32_t ssm_theta_uniform(int32_t n_sets, int32_t n_params, const double* lo_dev, const double* hi_dev,
                          uint64_t seed, float* theta_dev, void* stream) {
  return theta_uniform_enqueue(n_sets, n_params, lo_dev, hi_dev, seed, theta_dev, (hipStream_t)stream);
}

int32_t ssm_filter_workspace_bytes(int32_t n_sets, int64_t n_total, size_t* bytes) {
  return filter_workspace_bytes(n_sets, n_total, bytes);
}

int32_t ssm_filter_stats(int32_t n_sets, int32_t n_choices, const int32_t* choice_values_dev, const float* rts_dev,
                         const void* choices_dev, int32_t choice_bytes, const int64_t* offsets_dev, int64_t n_total,
                         void* workspace_dev, size_t workspace_bytes, double* stats_dev, void* stream) {
  return filter_stats_enqueue(n_sets, n_choices, choice_values_dev, rts_dev, choices_dev, choice_bytes, offsets_dev,
                              n_total, workspace_dev, workspace_bytes, stats_dev, (hipStream_t)stream);
}

int32_t ssm_mlp_sizes(int32_t* n_params, int32_t* n_frag, int32_t* n_bias, int32_t* act_rows,
                      int32_t* delta_rows) {
  return mlp_sizes(n_params, n_frag, n_bias, act_rows, delta_rows);
}

int32_t ssm_mlp_pack(const float* params_dev, void* frags_dev, float* bias_dev, void* stream) {
  return mlp_pack_enqueue(params_dev, (__bf16*)frags_dev, bias_dev, (hipStream_t)stream);
}

int32_t ssm_mlp_forward(const float* x_dev, int64_t n, const void* frags_dev, const float* bias_dev,
                        const float* xnorm_dev, float* y_dev, void* stream) {
  return mlp_forward_enqueue(x_dev, n, (const __bf16*)frags_dev, bias_dev, xnorm_dev, y_dev, (hipStream_t)stream);
}

int32_t ssm_mlp_backward(const float* x_dev, const float* target_or_grad_dev, int64_t n, int32_t mode, float delta,
                         float scale, const void* frags_dev, const float* bias_dev, const float* xnorm_dev,
                         void* acts_dev, void* deltas_dev, int64_t ld, float* dx_dev, float* y_dev, float* loss_dev,
                         void* stream) {
  if (mode != 0 && mode != 1) return fail(SSM_EINVAL, "mode must be 0 (upstream gradient) or 1 (Huber)");
  TrainArgs A{x_dev, target_or_grad_dev, (const __bf16*)frags_dev, bias_dev, xnorm_dev, (__bf16*)acts_dev,
              (__bf16*)deltas_dev, dx_dev, y_dev, loss_dev, n, ld, mode, delta, scale};
  return mlp_train_enqueue(A, (hipStream_t)stream);
}

int32_t ssm_mlp_wgrad(const void* acts_dev, const void* deltas_dev, int64_t n, int64_t ld, float* grad_dev,
                      void* stream) {
  return mlp_wgrad_enqueue((const __bf16*)acts_dev, (const __bf16*)deltas_dev, n, ld, grad_dev, (hipStream_t)stream);
}

int32_t ssm_mlp_adam(float* params_dev, const float* grad_dev, float* m_dev, float* v_dev, int32_t step, float lr,
                     float beta1, float beta2, float eps, float weight_decay, void* stream) {
  return mlp_adam_enqueue(params_dev, grad_dev, m_dev, v_dev, step, lr, beta1, beta2, eps, weight_decay,
                          (hipStream_t)stream);
}

int64_t ssm_mlp_train_workspace_bytes(int64_t batch) { return mlp_train_workspace_bytes(batch); }

int32_t ssm_mlp_train_steps(const float* x_dev, const float* target_dev, int64_t batch, int32_t n_steps,
                            int32_t step0, const float* lr_host, float beta1, float beta2, float eps,
                            float weight_decay, float delta, float* params_dev, float* m_dev, float* v_dev,
                            void* frags_dev, float* bias_dev, const float* xnorm_dev, void* workspace_dev,
                            int64_t workspace_bytes, float* loss_dev, void* stream) {
  return mlp_train_steps_enqueue(x_dev, target_dev, batch, n_steps, step0, lr_host, beta1, beta2, eps,
                                 weight_decay, delta, params_dev, m_dev, v_dev, (__bf16*)frags_dev, bias_dev,
                                 xnorm_dev, workspace_dev, workspace_bytes, loss_dev, (hipStream_t)stream);
}

int32_t ssm_mfma_probe(const void* a_dev, const void* b_dev, float* d_dev, void* stream) {
  return mfma_probe_enqueue(a_dev, b_dev, d_dev, (hipStream_t)stream);
}

int32_t ssm_mfma32_probe(const void* a_dev, const void* b_dev, float* d_dev, void* stream) {
  return mfma32_probe_enqueue(a_dev, b_dev, d_dev, (hipStream_t)stream);
}

int32_t ssm_valu_probe(int32_t kind, int64_t iters, int64_t n_threads, float* out_dev, void* stream) {
  return probe_enqueue(kind, iters, n_threads, out_dev, (hipStream_t)stream);
}

}  // extern "C"
