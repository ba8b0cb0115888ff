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
int32_t lba_enqueue(int32_t nact, int32_t n_sets, int64_t n_total, int32_t out_layout, int32_t choice_bytes,
                    uint64_t seed, const float* params, const int64_t* offsets, const float* replay,
                    float* rts, void* choices, hipStream_t stream);
int32_t philox_enqueue(uint64_t seed, uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, int64_t n,
                       uint32_t* raw, float* normals, hipStream_t stream);

int32_t probe_enqueue(int32_t kind, int64_t iters, int64_t n_threads, float* out, hipStream_t stream);

struct GraphHandle {
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

}  // namespace ssmhip

using namespace ssmhip;

extern "C" {

int32_t ssm_abi_version(void) { return 1; }

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
  hipStream_t s = nullptr;
  SSM_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  GraphHandle* h = new GraphHandle();
  hipError_t e = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal);
  if (e != hipSuccess) {
    delete h;
    hipStreamDestroy(s);
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
  hipStreamDestroy(s);
  if (rc != SSM_OK) {
    if (h->exec) hipGraphExecDestroy(h->exec);
    if (h->graph) hipGraphDestroy(h->graph);
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
  if (h->exec) hipGraphExecDestroy(h->exec);
  if (h->graph) hipGraphDestroy(h->graph);
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

int32_t ssm_valu_probe(int32_t kind, int64_t iters, int64_t n_threads, float* out_dev, void* stream) {
  return probe_enqueue(kind, iters, n_threads, out_dev, (hipStream_t)stream);
}

}  // extern "C"
