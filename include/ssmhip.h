/*
 * ssmhip.h — C ABI of the MI355X (gfx950) simulators, KDE labelling and
 * likelihood-surrogate kernels.
 *
 * Conventions
 *   - Every pointer argument named *_dev is a device pointer (hipMalloc or a
 *     torch.empty(device="cuda") tensor); the library never takes ownership.
 *   - `stream` is a hipStream_t passed as void* (NULL = default stream).
 *   - Every entry point returns 0 on success or a negative SSM_E* status and
 *     never throws; ssm_last_error() returns the calling thread's message.
 *   - Calls are asynchronous on `stream` unless stated otherwise.
 *
 * Reference interfaces replaced (paths relative to the ssm-simulators tree):
 *   src/cssm/ddm_models.pyx:589-878   ddm_flexbound  -> ssm_em_simulate(model=DDM)
 *   src/cssm/race_models.pyx:141-506  race_model     -> ssm_em_simulate(model=RACE)
 *   src/cssm/race_models.pyx:509-920  lca            -> ssm_em_simulate(model=LCA)
 *   src/cssm/lba_models.pyx:36-250    lba_vanilla / lba_angle -> ssm_lba_simulate
 *   ssms/support_utils/kde_class.py:89-186,438-553  LogKDE fit -> ssm_kde_fit
 *   ssms/support_utils/kde_class.py:188-329         kde_eval   -> ssm_kde_eval
 *   ssms/support_utils/kde_class.py:331-434         kde_sample -> ssm_kde_sample
 */
#ifndef SSMHIP_H
#define SSMHIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes --------------------------------------------------------- */
#define SSM_OK 0
#define SSM_EINVAL (-1)   /* invalid argument (maps to Python ValueError)   */
#define SSM_EHIP (-2)     /* HIP runtime error (maps to RuntimeError)       */
#define SSM_ENOMEM (-3)   /* allocation failure                              */
#define SSM_EUNSUPPORTED (-4)

/* ---- model / boundary selectors ------------------------------------------ */
#define SSM_MODEL_DDM 0   /* ddm_flexbound: bounds +-b(t), start -b0 + 2 z b0 */
#define SSM_MODEL_RACE 1  /* race_model: N accumulators clamped at 0           */
#define SSM_MODEL_LCA 2   /* lca: race + leak g and lateral inhibition b       */
#define SSM_MODEL_FULL_DDM 3 /* full_ddm: DDM + per-trajectory v, z, t variability
                                (src/cssm/ddm_models.pyx:2177-2360)            */
#define SSM_MODEL_ORNSTEIN 4 /* ornstein_uhlenbeck: DDM with leak, step
                                ((v - g y) dt) + sqrt_st g (ornstein_models.pyx:47-200) */
#define SSM_MODEL_LEVY 5     /* levy_flexbound: DDM walk with symmetric alpha-stable
                                increments (v dt) + (s dt^(1/alpha)) x (levy_models.pyx:47-220) */
#define SSM_MODEL_RDM 6      /* racing_diffusion_model: N unclamped accumulators from
                                U(0, A) starts, first past b wins (race_models.pyx:925-1260) */
#define SSM_MODEL_PAR2 7     /* ddm_flexbound_par2: high walk, then a low walk from t = 0;
                                choice in 0..3, rt = max(t_h, t_l) (parallel_models.pyx:57-560) */
#define SSM_MODEL_SEQ2 8     /* ddm_flexbound_seq2: the low walk continues where the high
                                walk ended; choice in 0..3 (sequential_models.pyx:55-640) */
#define SSM_MODEL_DDM_LEGACY 9 /* ddm: bounds 0 and a, start z a (ddm_models.pyx:349-582) */

#define SSM_BOUNDARY_CONSTANT 0 /* b(t) = a                                  */
#define SSM_BOUNDARY_ANGLE 1    /* b(t) = a + t_s * slope, slope = -sin/cos  */
#define SSM_BOUNDARY_WEIBULL 2  /* b(t) = a * exp(-(t_s/beta)^alpha)         */
#define SSM_BOUNDARY_TABLE 3    /* b(t) = table[set][ix] (any Python boundary)*/

#define SSM_MAX_PARTICLES 16

/* Per-set parameter record: float32[SSM_SET_STRIDE] per parameter set (the
 * single-particle diffusions may use compact records, see set_stride).      */
#define SSM_SET_STRIDE 64
#define SSM_SET_STRIDE_SINGLE 16
#define SSM_P_T0 0           /* non-decision time t                          */
#define SSM_P_DEADLINE 1     /* deadline (999 when the model has none)       */
#define SSM_P_DEADLINE_TMP 2 /* min(max_t, deadline - t), float32            */
#define SSM_P_A 3            /* boundary a                                    */
#define SSM_P_B1 4           /* angle: slope (-sin th/cos th); weibull: alpha */
#define SSM_P_B2 5           /* weibull: beta                                 */
#define SSM_P_Z 6            /* DDM relative start z                          */
#define SSM_P_V 7            /* DDM drift v                                   */
#define SSM_P_SQRT_ST 8      /* DDM float32(sqrt(dt)) * s                     */
#define SSM_P_G 9            /* LCA / Ornstein leak g                         */
#define SSM_P_BINH 10        /* LCA inhibition b                              */
#define SSM_P_SZ 11          /* full DDM start variability sz                 */
#define SSM_P_SV 12          /* full DDM drift variability sv                 */
#define SSM_P_ST 13          /* full DDM non-decision-time variability st     */
#define SSM_P_ALPHA 14       /* Levy stability alpha (0 < alpha <= 2); the scale
                                s dt^(1/alpha) goes in SSM_P_SQRT_ST          */
#define SSM_P_MAX_T 15       /* float32 max_t (two-stage models)              */
/* Two-stage models: SSM_P_V = vh, SSM_P_Z = zh, SSM_P_VJ + 0/1 = vl1/vl2,
 * SSM_P_ZJ + 0/1 = zl1/zl2.                                                  */
#define SSM_P_VJ 16          /* race/LCA v_j, j < 16                          */
#define SSM_P_ZJ 32          /* race/LCA z_j; racing diffusion start range A  */
#define SSM_P_SSTJ 48        /* race/LCA float32(sqrt(dt)) * s_j              */

typedef struct ssm_em_problem {
  int32_t model;         /* SSM_MODEL_*                                       */
  int32_t boundary;      /* SSM_BOUNDARY_*                                    */
  int32_t n_particles;   /* 1 for DDM; 1..16 for RACE/LCA                     */
  int32_t n_sets;        /* number of parameter sets (trials)                 */
  int64_t n_total;       /* offsets[n_sets]: total trajectories               */
  int32_t n_steps;       /* boundary table length int(max_t/dt + 1)          */
  float dt;              /* delta_t as float32                                */
  int32_t smooth_unif;   /* 1: add the reference's uniform RT jitter          */
  int32_t out_layout;    /* 0: out[offsets[k] + n]; 1: out[n * n_sets + k]    */
  int32_t choice_bytes;  /* 1: int8 choices; 4: int32 (np.intc) choices       */
  int32_t set_stride;    /* floats per set record; 0 = SSM_SET_STRIDE. DDM,
                            full DDM, Ornstein, Levy and legacy DDM read slots
                            0..15 only and accept SSM_SET_STRIDE_SINGLE (16);
                            the other models need at least SSM_SET_STRIDE   */
  uint64_t seed;         /* Philox key                                        */
  int64_t replay_stride; /* 0: native Philox; >0: replay buffer row length    */
} ssm_em_problem;

typedef struct ssm_launch_config {
  int32_t block_size;    /* threads per block (multiple of 64), 0 = default   */
  int32_t blocks_per_cu; /* resident blocks per CU in the persistent grid     */
  int32_t refill_min;    /* idle lanes (1..64) before a wavefront refills; a
                            native Gaussian-diffusion job of at most 64 x CUs
                            (160 x CUs) trajectories runs 4 (16) walks per
                            wavefront instead and refills as soon as one ends */
  int32_t queue_block;   /* largest index range a wavefront claims from the
                            global queue per atomic (0 = 64, at most 65536)  */
  /* Optional visiting order of the sets (both NULL: index order). set_order
   * int32[n_sets] is a permutation of 0..n_sets-1; queue_offsets int64[n_sets+1]
   * the prefix sums of the set sizes taken in that order. Long-running sets
   * first shortens the tail of a launch; outputs do not depend on it.        */
  const int32_t* set_order;
  const int64_t* queue_offsets;
} ssm_launch_config;

/* ---- library ---------------------------------------------------------------*/
/* 3: ssm_em_problem.set_stride (compact 16-float records for the single-particle
 *    diffusions; 0 keeps SSM_SET_STRIDE, the layout of version 2).
 * 4: ssm_lan_assemble takes the negative-RT rows (n_neg, n_choices, neg_ll);
 *    ssm_copy_rows, ssm_em_trajectory, ssm_kde_eval_noalloc.
 * 5: ssm_mfma32_probe.
 * 6: ssm_em_records, ssm_uniform_offsets.                                     */
int32_t ssm_abi_version(void);
const char* ssm_last_error(void);
/* Number of CUs of the current device (0 and an error when no GPU).          */
int32_t ssm_device_cu_count(void);

/* ---- Euler–Maruyama first-passage simulation ------------------------------
 * Replaces cssm.ddm_flexbound (src/cssm/ddm_models.pyx:589-878, sequential
 * loop :696-744), cssm.race_model (src/cssm/race_models.pyx:356-506),
 * cssm.lca (src/cssm/race_models.pyx:757-920), cssm.full_ddm,
 * cssm.ornstein_uhlenbeck, cssm.levy_flexbound (src/cssm/levy_models.pyx:47-220),
 * cssm.racing_diffusion_model (src/cssm/race_models.pyx:925-1260) and
 * cssm.ddm_flexbound_par2 / ddm_flexbound_seq2 (src/cssm/parallel_models.pyx,
 * src/cssm/sequential_models.pyx), batched over ragged sets.
 *   set_params_dev   float32[n_sets][set_stride] (set_stride 0: SSM_SET_STRIDE)
 *   offsets_dev      int64[n_sets + 1], offsets[0] = 0, non-decreasing
 *   table_dev        float32[n_sets][n_steps] when boundary == TABLE, else NULL
 *   replay_normals_dev float32[n_total][replay_stride] (replay mode) or NULL;
 *                    draw d of trajectory i reads [i][d % replay_stride],
 *                    d = step * n_particles + particle (full DDM: [i][0] is
 *                    the drift draw and step d reads [i][(1 + d) % stride];
 *                    Levy: the alpha-stable variates themselves; racing
 *                    diffusion: draws up to the winner's of its last step)
 *   replay_uniforms_dev float32[n_total] (replay + smooth_unif) or NULL;
 *                    full DDM replay: float32[n_total][3] = (u_t, u_z, u_smooth);
 *                    racing diffusion replay (always given):
 *                    float32[n_total][n_particles + 1] = (start u_j..., u_smooth);
 *                    two-stage replay (always given): float32[n_total][4] =
 *                    (u_high, u_low_prior, u_smooth, u_low); their normals rows
 *                    hold the high walk's draws, then the low walk's
 *   rts_dev          float32[n_total]; two-stage par2: float32[3][n_total] =
 *                    (rt, rt_high, rt_low) (the reference's rts_high / rts_low)
 *   choices_dev      int8/int32[n_total]
 *   workspace_dev    uint64[SSM_EM_WORKSPACE_WORDS] scratch owned by the caller
 *                    ([0] work-queue head, [1] executed EM steps)
 * `launch` may be NULL (defaults). Two launches on `stream`: a workspace reset
 * kernel and the persistent simulation kernel.
 * The walks read their times from a table of t after k steps (k sequential float32
 * additions of delta_t, k = 0..n_steps) that the library builds once per (device,
 * delta_t, n_steps) and keeps for the life of the process (4 (n_steps + 1) bytes).
 * The first call of a (delta_t, n_steps) on a device allocates and uploads it, so
 * that call may not run inside a stream capture (SSM_EINVAL if it does;
 * ssm_em_graph_create builds the table before it captures).                */
#define SSM_EM_WORKSPACE_WORDS 4
int32_t ssm_em_simulate(const ssm_em_problem* prob, const ssm_launch_config* launch,
                        const float* set_params_dev, const int64_t* offsets_dev,
                        const float* table_dev, const float* replay_normals_dev,
                        const float* replay_uniforms_dev, float* rts_dev,
                        void* choices_dev, uint64_t* workspace_dev, void* stream);

/* Replaces the trajectory recording of the sequential loops (metadata['trajectory']:
 * ddm_models.pyx:714-716, :726-728 and the other single-walk diffusions; race_models.pyx:
 * 433-436, :449-452 (race), :841-844, :865-868 (lca), :1223-1226, :1250-1254 (racing
 * diffusion)): walks trajectory 0 of set 0 of the same problem (same draws, so the same walk
 * as in ssm_em_simulate) and writes its positions traj_dev[ix][j] (float32[traj_rows]
 * [n_particles]) for ix = 0 (the start) up to its last step; rows after it are left as the
 * caller filled them (the reference fills -999). Set 0 must hold at least one sample.
 * Two launches on `stream` (workspace reset, one wave). SSM_EUNSUPPORTED for the two-stage
 * models, whose reference trajectories also record both low walks.                    */
int32_t ssm_em_trajectory(const ssm_em_problem* prob, const float* set_params_dev, const int64_t* offsets_dev,
                          const float* table_dev, const float* replay_normals_dev,
                          const float* replay_uniforms_dev, float* traj_dev, int32_t traj_rows,
                          uint64_t* workspace_dev, void* stream);

/* Capture the two launches of ssm_em_simulate into a HIP graph executable.
 * The pointers are baked into the graph; replay with ssm_graph_launch. The
 * walks' time table (see ssm_em_simulate) is built before the capture.      */
int32_t ssm_em_graph_create(const ssm_em_problem* prob, const ssm_launch_config* launch,
                            const float* set_params_dev, const int64_t* offsets_dev,
                            const float* table_dev, const float* replay_normals_dev,
                            const float* replay_uniforms_dev, float* rts_dev,
                            void* choices_dev, uint64_t* workspace_dev, void** graph_out);
int32_t ssm_graph_launch(void* graph, void* stream);
int32_t ssm_graph_destroy(void* graph);

/* ---- LBA (closed form, no time stepping) ----------------------------------
 * Replaces cssm.lba_vanilla / lba_angle (src/cssm/lba_models.pyx:36-250).
 *   params_dev float32[n_sets][SSM_SET_STRIDE]: [SSM_P_A]=b threshold a,
 *     [SSM_P_Z]=start-point range z, [SSM_P_T0]=t, [SSM_P_DEADLINE]=deadline,
 *     [SSM_P_B1]=tan(theta) (angle) or 0, [SSM_P_VJ+j]=v_j, [SSM_P_SSTJ+j]=sd_j
 *   replay_dev NULL (Philox) or float32[n_total][2*nact]: per sample nact
 *     uniforms in [0,1) followed by nact standard normals.                  */
int32_t ssm_lba_simulate(int32_t nact, int32_t n_sets, int64_t n_total, int32_t out_layout,
                         int32_t choice_bytes, uint64_t seed, const float* params_dev,
                         const int64_t* offsets_dev, const float* replay_dev, float* rts_dev,
                         void* choices_dev, void* stream);

/* ---- RNG diagnostics --------------------------------------------------------
 * Philox4x32-10 (Salmon et al., SC'11) blocks for counters (c0 + i, c1, c2, c3),
 * key = seed, i < n: raw_dev uint32[n][4] (nullable) and their N(0, 1)
 * Box–Muller normals, normals_dev float32[n][4] (nullable). The simulators
 * draw the same normals; native Gaussian-diffusion walks draw them with the
 * radius scaled by sqrt(dt) s.                                               */
int32_t ssm_philox_fill(uint64_t seed, uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
                        int64_t n, uint32_t* raw_dev, float* normals_dev, void* stream);

/* ---- KDE likelihood labelling ------------------------------------------------
 * Per (set, choice) Gaussian KDE on log RT, C = n_choices <= 16 possible
 * choice values choice_values_dev[C]; group g = set * C + c.
 *
 * ssm_kde_fit replaces LogKDE.__init__ (ssms/support_utils/kde_class.py:48-87:
 * _attach_data_from_simulator :438-504, compute_bandwidths :89-126,
 * bandwidth_silverman :508-553, the KernelDensity fit :174-176). For every
 * set (rts/choices segment offsets[k]..offsets[k+1]) it writes log(rt) of the
 * non-omitted (rt != -999) draws of each choice, grouped by choice, into
 * samples_dev[total] and per group: kde_count (kept draws), choice_total (all
 * draws, p_c = choice_total / N), bandwidth (Silverman h; 0 = no base data)
 * and kde_start (absolute index into samples_dev).                           */
int32_t ssm_kde_fit(int32_t n_sets, int32_t n_choices, const int32_t* choice_values_dev, const float* rts_dev,
                    const void* choices_dev, int32_t choice_bytes, const int64_t* offsets_dev, float* samples_dev,
                    int32_t* kde_count_dev, int32_t* choice_total_dev, float* bandwidth_dev,
                    int64_t* kde_start_dev, void* stream);

/* ssm_kde_sample replaces LogKDE.kde_sample (kde_class.py:331-434) and the
 * uniform-row draws of ResampleMixtureStrategy.generate
 * (ssms/dataset_generators/strategies/resample_mixture_strategy.py:160-200).
 * Per set: n_kde KDE resamples (per-choice counts round(n_kde * p_c) half to
 * even with the reference's +-1 fix-up, or n_by_override_dev[n_sets*C]) and
 * n_unif points rt ~ U(unif_lo, max_t) with a uniform choice; m = n_kde +
 * n_unif <= 1024. Points are written grouped by choice: pts_rt_dev /
 * pts_choice_dev (choice index) [n_sets*m], grp_begin/grp_count [n_sets*C].   */
int32_t ssm_kde_sample(int32_t n_sets, int32_t n_choices, int32_t n_kde, int32_t n_unif, float max_t, float unif_lo,
                       uint64_t seed, const float* samples_dev, const int32_t* kde_count_dev,
                       const int32_t* choice_total_dev, const float* bandwidth_dev, const int64_t* kde_start_dev,
                       const int64_t* offsets_dev, const int32_t* n_by_override_dev, float* pts_rt_dev,
                       int8_t* pts_choice_dev, int64_t* grp_begin_dev, int32_t* grp_count_dev, void* stream);

/* ssm_kde_eval replaces LogKDE.kde_eval / __kde_eval_ (kde_class.py:188-329)
 * with sklearn KernelDensity.score_samples: for the grp_count[g] points at
 * pts_rt_dev[grp_begin[g]...] of group g writes out_ll_dev (same indexing):
 * log p_c + log mean_i N(log max(rt, eps); x_i, h) - log max(rt, eps), floored
 * at lb (also for rt <= 0); groups without base data get -log(N * max_t).
 * max_samples bounds every set's size (offsets[k+1] - offsets[k]);
 * workspace_dev: float32[n_points * ceil(max_samples / SSM_KDE_CHUNK)] where
 * n_points covers every point index of the group table.                    */
#define SSM_KDE_CHUNK 8192
int32_t ssm_kde_eval(int32_t n_sets, int32_t n_choices, const int64_t* grp_begin_dev, const int32_t* grp_count_dev,
                     const float* pts_rt_dev, const float* samples_dev, const int32_t* kde_count_dev,
                     const int32_t* choice_total_dev, const float* bandwidth_dev, const int64_t* kde_start_dev,
                     const int64_t* offsets_dev, float max_t, float lb, float eps, int64_t max_samples,
                     float* workspace_dev, float* out_ll_dev, void* stream);

/* ssm_kde_eval_noalloc has no counterpart in the reference. It is for callers
 * that cannot allocate device memory (code under HIP graph capture, serving
 * loops that evaluate likelihoods call after call, harnesses that forbid
 * hipMalloc). It computes exactly what ssm_kde_eval computes, with the same
 * kernels, but its partial sums live in a fixed device array of
 * SSM_KDE_EVAL_SCRATCH_FLOATS floats (64 MiB) that every GPU the library runs
 * on holds until the process ends. It allocates nothing.
 *   n_points  bounds every point index of the group table (the length of
 *             pts_rt_dev); n_points * ceil(max_samples / SSM_KDE_CHUNK) must
 *             not exceed SSM_KDE_EVAL_SCRATCH_FLOATS, else SSM_EINVAL.
 * All calls on one GPU share that GPU's array: calls that overlap on one GPU
 * (other streams, other host threads) race and must be serialised by the
 * caller. Every partial sum a call reads is written earlier in the same call,
 * so nothing needs clearing between calls.                                  */
#define SSM_KDE_EVAL_SCRATCH_FLOATS (16 * 1024 * 1024)
int32_t ssm_kde_eval_noalloc(int32_t n_sets, int32_t n_choices, const int64_t* grp_begin_dev,
                             const int32_t* grp_count_dev, const float* pts_rt_dev, const float* samples_dev,
                             const int32_t* kde_count_dev, const int32_t* choice_total_dev,
                             const float* bandwidth_dev, const int64_t* kde_start_dev, const int64_t* offsets_dev,
                             float max_t, float lb, float eps, int64_t max_samples, int64_t n_points,
                             float* out_ll_dev, void* stream);

/* ssm_lan_assemble replaces the row assembly and shuffle of
 * ResampleMixtureStrategy.generate (resample_mixture_strategy.py:129-214):
 * rows_dev[n_sets*(m + n_neg)][n_theta + 3] = (theta_dev[set][:], rt, choice value, ll).
 * Per set the m evaluated points (pts_rt / pts_choice / ll [n_sets*m]) come first, then
 * n_neg negative-RT rows (:202-209): rt ~ U(-1, 1e-4), choice uniform over the n_choices
 * values of choice_values_dev, ll = neg_ll (the reference's -66.77497). With shuffle != 0
 * the m + n_neg <= 1024 rows of each set are put in one Philox-keyed random order
 * (:211-212).                                                                 */
int32_t ssm_lan_assemble(int32_t n_sets, int32_t n_theta, int32_t m, int32_t n_neg, int32_t n_choices,
                         float neg_ll, int32_t shuffle, uint64_t seed, const float* theta_dev,
                         const float* pts_rt_dev, const int8_t* pts_choice_dev, const float* ll_dev,
                         const int32_t* choice_values_dev, float* rows_dev, void* stream);

/* ssm_em_records replaces the per-trial setup of the simulators' sequential loops (the
 * float32 quantities a trial forms once before its walk: ddm_models.pyx:696-717,
 * race_models.pyx:378-414) for many sets at once, building the set records on the device.
 * The host forms each used slot as one float32 column of n_sets values and uploads the
 * columns only: record k of records_dev (set_stride floats) gets slot slots[c] =
 * columns_dev[c * n_sets + k] for c < n_slots and 0 in every other slot. slots is a host
 * array of n_slots distinct indices below set_stride (set_stride <= SSM_SET_STRIDE).      */
int32_t ssm_em_records(int32_t n_sets, int32_t set_stride, int32_t n_slots, const int32_t* slots,
                       const float* columns_dev, float* records_dev, void* stream);

/* ssm_uniform_offsets writes offsets_dev[k] = k * n_samples for k <= n_sets: the segments
 * of Simulator's n_trials sets of n_samples draws each in the set-major outputs (the
 * reference indexes its [n_samples, n_trials] arrays directly, _utils.pyx:625-632).       */
int32_t ssm_uniform_offsets(int32_t n_sets, int64_t n_samples, int64_t* offsets_dev, void* stream);

/* ssm_copy_rows replaces the array assignments of the simulation filter's re-draw loop
 * (SimulationPipeline.generate_for_parameter_set, simulation_pipeline.py:103-115: a rejected
 * set's simulations are replaced by those of an accepted re-draw) and the dropping of sets
 * that never pass: row dst_idx[r] of dst_dev (r when dst_idx_dev is NULL) becomes row
 * src_idx[r] of src_dev (r when NULL), for r < n_rows, rows of row_bytes bytes.        */
int32_t ssm_copy_rows(int64_t n_rows, int64_t row_bytes, const void* src_dev, const int64_t* src_idx_dev,
                      void* dst_dev, const int64_t* dst_idx_dev, void* stream);

/* ssm_aux_labels replaces SimulationPipeline._compute_auxiliary_labels
 * (ssms/dataset_generators/pipelines/simulation_pipeline.py:258-351) as
 * integer counts per set k over its draws [offsets[k], offsets[k+1]):
 *   counts_dev[k][c]   draws with choice choice_values[c] (omitted rts included)
 *   nogo_dev[k]        draws with choice != max_choice or rt == -999
 *   hist128_dev[k][b][c], hist256_dev[k][b][c]
 *                      np.histogram counts of the non-omitted rts of choice c
 *                      over np.linspace(0, max_t, B + 1), B = 128 / 256
 * The caller forms choice_p = counts / n, nogo_p = nogo / n, etc.            */
int32_t ssm_aux_labels(int32_t n_sets, int32_t n_choices, const int32_t* choice_values_dev, int32_t max_choice,
                       const float* rts_dev, const void* choices_dev, int32_t choice_bytes,
                       const int64_t* offsets_dev, double max_t, int32_t* counts_dev, int32_t* nogo_dev,
                       int32_t* hist128_dev, int32_t* hist256_dev, void* stream);

/* ssm_filter_stats replaces the statistics of SimulationPipeline._is_valid_simulation
 * (ssms/dataset_generators/pipelines/simulation_pipeline.py:199-256): for each set k
 * and choice c, over the draws of [offsets[k], offsets[k+1]) with choice
 * choice_values[c] and rt != -999,
 *   stats_dev[k][c][0..4] = (mode, mean, std, mode count / n (0 when n < 5), n)
 * with the mode the smallest of the most frequent values (scipy.stats.mode) and
 * (-1, -1, 1, 0, 0) for a choice without draws. Mean and std are formed in double.
 * workspace_dev: at least ssm_filter_workspace_bytes(n_sets, n_total) bytes
 * (keys and rocPRIM segmented-sort storage); n_total < 2^32.                */
int32_t ssm_filter_workspace_bytes(int32_t n_sets, int64_t n_total, size_t* bytes);
int32_t ssm_filter_stats(int32_t n_sets, int32_t n_choices, const int32_t* choice_values_dev, const float* rts_dev,
                         const void* choices_dev, int32_t choice_bytes, const int64_t* offsets_dev, int64_t n_total,
                         void* workspace_dev, size_t workspace_bytes, double* stats_dev, void* stream);

/* ssm_theta_uniform replaces UniformParameterSampler._sample_parameter
 * (ssms/dataset_generators/parameter_samplers/uniform_sampler.py:24-51) for
 * independent bounds: theta_dev[k][j] = float32(lo[j] + (hi[j] - lo[j]) u) in
 * float64, u a 53-bit uniform from Philox4x32-10 keyed by seed (lo_dev, hi_dev:
 * float64[n_params], 1 <= n_params <= 64).                                    */
int32_t ssm_theta_uniform(int32_t n_sets, int32_t n_params, const double* lo_dev, const double* hi_dev,
                          uint64_t seed, float* theta_dev, void* stream);

/* ---- likelihood surrogate (no counterpart in the reference) ------------------
 * MLP 7 -> 100 -> 100 -> 120 -> 1 with tanh hidden layers on BF16 MFMA with
 * FP32 accumulation; inputs are the LAN rows' (5 theta, rt, choice)
 * (ssms/dataset_generators/strategies/resample_mixture_strategy.py:136-139).
 * params: FP32 master weights, flat nn.Linear layout W1[100][7], b1, W2, b2,
 * W3[120][100], b3, W4[1][120], b4 (ssm_mlp_sizes gives every length).
 * ssm_mlp_pack converts params to bf16 MFMA fragments + padded biases.
 * xnorm_dev (nullable) = shift[7], scale[7], rt_mode, t_index (float32[16]):
 *   the network sees (x' - shift) * scale where x' is x with input 5 (rt)
 *   replaced by log(max(u, 1e-4)), u = rt for rt_mode 1 and u = rt - x[t_index]
 *   for rt_mode 2 (unchanged for 0; t_index is clamped to 0..6); dx is
 *   returned w.r.t. the raw x (chain rule through u included).
 * ssm_mlp_forward: y[n] = net(x[n][7]).
 * ssm_mlp_backward: forward + backward-data for n rows; mode 0 takes the
 *   upstream gradient dL/dy in target_or_grad, mode 1 the Huber targets
 *   (loss scaled by `scale`); writes dx[n][8] (dL/dx, column 7 is padding),
 *   y, per-row loss, and, for ssm_mlp_wgrad, feature-major bf16 activations
 *   acts[act_rows][ld] and deltas[delta_rows][ld] (ld >= n, ld % 32 == 0,
 *   columns >= n zero). Any of these outputs may be NULL.
 * ssm_mlp_wgrad: grad[n_params] = dL/dparams from acts/deltas.
 * ssm_mlp_adam: Adam step `step` (>= 1) on params with moments m, v.
 * ssm_mlp_train_steps: n_steps Adam steps on the Huber loss, step i on rows
 *   [i*batch, (i+1)*batch) of x[n_steps*batch][7] / target[...], learning rate
 *   lr_host[i] (host array), bias corrections for steps step0+1 ...; per step
 *   the fused forward/backward, the weight gradient (summed over batch slices
 *   in a fixed order) and Adam fused with the repack of frags/bias are
 *   enqueued from the host without a synchronisation. frags_dev/bias_dev
 *   must hold ssm_mlp_pack(params) on entry (the steps rewrite only the
 *   entries of real parameters; padding stays zero). batch % 32 == 0;
 *   workspace_dev (16-byte aligned) holds at least
 *   ssm_mlp_train_workspace_bytes(batch) bytes (-1: invalid batch); loss_dev
 *   (nullable) receives the per-row Huber loss [n_steps*batch].             */
int32_t ssm_mlp_sizes(int32_t* n_params, int32_t* n_frag, int32_t* n_bias, int32_t* act_rows, int32_t* delta_rows);
int32_t ssm_mlp_pack(const float* params_dev, void* frags_dev, float* bias_dev, void* stream);
int32_t ssm_mlp_forward(const float* x_dev, int64_t n, const void* frags_dev, const float* bias_dev,
                        const float* xnorm_dev, float* y_dev, void* stream);
int32_t ssm_mlp_backward(const float* x_dev, const float* target_or_grad_dev, int64_t n, int32_t mode, float delta,
                         float scale, const void* frags_dev, const float* bias_dev, const float* xnorm_dev,
                         void* acts_dev, void* deltas_dev, int64_t ld, float* dx_dev, float* y_dev, float* loss_dev,
                         void* stream);
int32_t ssm_mlp_wgrad(const void* acts_dev, const void* deltas_dev, int64_t n, int64_t ld, float* grad_dev,
                      void* stream);
int32_t ssm_mlp_adam(float* params_dev, const float* grad_dev, float* m_dev, float* v_dev, int32_t step, float lr,
                     float beta1, float beta2, float eps, float weight_decay, void* stream);
int64_t ssm_mlp_train_workspace_bytes(int64_t batch);
int32_t ssm_mlp_train_steps(const float* x_dev, const float* target_dev, int64_t batch, int32_t n_steps,
                            int32_t step0, const float* lr_host, float beta1, float beta2, float eps,
                            float weight_decay, float delta, float* params_dev, float* m_dev, float* v_dev,
                            void* frags_dev, float* bias_dev, const float* xnorm_dev, void* workspace_dev,
                            int64_t workspace_bytes, float* loss_dev, void* stream);
/* One v_mfma_f32_16x16x32_bf16 with per-lane operands a[64][8], b[64][8]
 * (bf16) and result d[64][4]: pins the operand / result lane layout.        */
int32_t ssm_mfma_probe(const void* a_dev, const void* b_dev, float* d_dev, void* stream);
/* One v_mfma_f32_32x32x16_bf16 (the inference kernel's) with per-lane operands a[64][8],
 * b[64][8] (bf16) and result d[64][16]: pins that operand / result lane layout.      */
int32_t ssm_mfma32_probe(const void* a_dev, const void* b_dev, float* d_dev, void* stream);

/* ---- VALU roofline probes (no counterpart in the reference) ----------------
 * Branch-free loops at the occupancy of `n_threads` lanes (multiple of 256),
 * `iters` iterations each; out_dev float32[n_threads]. kind 0: the EM block
 * (Philox4x32-10 + 2 Box–Muller pairs + 4 angle-DDM steps); kind 1: the KDE
 * pair (sub, fma, exp2, add), 4 points x 4 samples; kinds 2..8: 8 chains of
 * one instruction (v_mad_u64_u32, v_xor_b32, v_fma_f32, v_log_f32, v_exp_f32,
 * v_sin_f32, v_sqrt_f32); kind 9: 8 chains of the surrogate's tanh
 * (exp2, add, rcp, fma). Time it with events on `stream`.                  */
int32_t ssm_valu_probe(int32_t kind, int64_t iters, int64_t n_threads, float* out_dev, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* SSMHIP_H */
