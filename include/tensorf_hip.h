/*
 * tensorf_hip.h -- C ABI of libtensorf_hip.so, the MI355X (gfx950) kernels of
 * the TensoRF render / train hot path.
 *
 * The reference (brentyi/tensorf-jax) has no FFI: its operator API is the
 * Python signatures of tensorf/tensor_vm.py, tensorf/render.py and
 * tensorf/networks.py. Each entry point below names the reference function
 * it replaces (file:line). A Python caller binds them with ctypes
 * (tensorf_amd/_lib.py); a jax.ffi custom call or a cgo/JNI stub would bind
 * the same symbols.
 *
 * Conventions
 *  - Every pointer is a device pointer owned by the caller (except the
 *    struct arguments themselves, which live in host memory).
 *  - Shapes are int64_t; float is IEEE binary32; indices are int32_t.
 *  - Every call is asynchronous on `stream` (a hipStream_t, NULL = the
 *    null stream) and returns a status: 0 = OK, negative = error. The
 *    message of the last error of the calling thread is tfh_last_error().
 *    No call allocates device memory: scratch comes from a caller-owned
 *    workspace whose size the *_workspace_size query returns.
 *  - VM factors are stored channel-last:
 *        plane[m][row][col][c] == matrix[m][c][row][col]   (3, G, G, C)
 *        line[m][v][c]         == vector[m][c][v]          (3, G, C)
 *    where matrix/vector are the reference layouts of TensorVMSingle
 *    (tensor_vm.py:129-133). tfh_vm_to_channel_last converts.
 */
#ifndef TENSORF_HIP_H
#define TENSORF_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TFH_ABI_VERSION 2

typedef void* tfh_stream_t;

enum tfh_status {
  TFH_OK = 0,
  TFH_ERR_INVALID_ARGUMENT = -1,
  TFH_ERR_UNSUPPORTED = -2,
  TFH_ERR_NO_DEVICE = -3,
  TFH_ERR_LAUNCH = -4,
  TFH_ERR_WORKSPACE = -5
};

/* render.RenderMode (render.py:18-23).
 * TFH_MODE_DIST_MEDIAN: the distance of the first sample whose 1 - E_i
 *   exceeds 0.5 (render.py:248-267), 0 when that is sample 0 (the xor of the
 *   padded mask never selects column 0) and +inf when no sample reaches it.
 * TFH_MODE_DIST_MEDIAN_AS_WRITTEN: render.py:248-267 evaluated as written
 *   under IEEE arithmetic: the 0/1 mask multiplies a distance row that ends in
 *   +inf, and 0 * inf = NaN, so a ray whose 1 - E_{N-1} exceeds 0.5 gets NaN
 *   and any other ray +inf. For parity checks against the reference. */
enum tfh_render_mode {
  TFH_MODE_RGB = 0,
  TFH_MODE_DIST_MEDIAN = 1,
  TFH_MODE_DIST_MEAN = 2,
  TFH_MODE_DIST_MEDIAN_AS_WRITTEN = 3
};

/* Arithmetic of the appearance MLP. FP32: FP32 MFMA. BF16: BF16 operands on
 * MFMA with FP32 accumulation. */
enum tfh_precision { TFH_PREC_FP32 = 0, TFH_PREC_BF16 = 1 };

/* Where the N density samples of a ray lie.
 * SCHEDULE: the reference's samplers -- stratified in the box segment
 *   (render.py:350-390) or, with scene_contraction, the contracted schedule
 *   base_ts / base_steps (render.py:122-173).
 * EDGES: per-ray interval edges e (R, N+1), non-decreasing along each ray,
 *   passed in the base_ts argument (base_steps unused); sample i lies at
 *   t_i = e_i + u_i (e_{i+1} - e_i) with u = jitter (R, N) (0.5 when jitter is
 *   NULL) and has step size e_{i+1} - e_i. This is how a proposal sampler
 *   (tfh_proposal_sample) hands its intervals to the field. No counterpart in
 *   the reference. */
enum tfh_sampling { TFH_SAMPLING_SCHEDULE = 0, TFH_SAMPLING_EDGES = 1 };

/* One VM-decomposed field: TensorVM (tensor_vm.py:14-19), channel-last. */
typedef struct {
  const float* plane; /* (3, G, G, C) */
  const float* line;  /* (3, G, C) */
  int64_t grid_dim;   /* G >= 2 */
  int64_t channels;   /* C per mode, 1..64 */
} tfh_vm;

typedef struct {
  float* plane; /* (3, G, G, C), accumulated into (+=) */
  float* line;  /* (3, G, C), accumulated into (+=) */
} tfh_vm_grad;

/* FeatureMlp parameters (networks.py:38-121). Kernels are (in, out)
 * row-major like Flax Dense kernels. */
typedef struct {
  const float* w0;    /* (feat_dim, squash_dim), no bias (networks.py:57-61) */
  const float* w1;    /* (encoded_dim, units) */
  const float* b1;    /* (units) */
  const float* w2;    /* (units, units) */
  const float* b2;    /* (units) */
  const float* w3;    /* (units, 3) */
  const float* b3;    /* (3) */
  const float* embed; /* (num_cameras, units) FiLM table (networks.py:103-111) or NULL */
  int64_t feat_dim;   /* 3 * appearance channels */
  int64_t squash_dim; /* 27 */
  int64_t units;      /* 128 (the only width the MFMA kernels are built for) */
  int64_t feature_n_freqs;
  int64_t viewdir_n_freqs;
  int64_t num_cameras; /* 0 = no camera embedding */
} tfh_mlp;

typedef struct {
  float *w0, *w1, *b1, *w2, *b2, *w3, *b3, *embed; /* accumulated into (+=) */
} tfh_mlp_grad;

/* render.RenderConfig (render.py:26-36) + LearnableParams.scene_contraction
 * (render.py:46) + the MLP arithmetic. */
typedef struct {
  float near_plane;
  float far_plane;
  int32_t mode;                       /* tfh_render_mode */
  int32_t density_samples_per_ray;    /* N, 2..1024 */
  int32_t appearance_samples_per_ray; /* k, 1..N */
  int32_t scene_contraction;          /* 0: bounded box, 1: L-inf contraction */
  int32_t mlp_precision;              /* tfh_precision */
  int32_t sampling;                   /* tfh_sampling */
  float density_shift;                /* sigma = softplus(sum + shift); the reference
                                         uses +10 (render.py:207), TensoRF's
                                         original code -10 */
} tfh_render_config;

/* ---- library ------------------------------------------------------------ */
int32_t tfh_abi_version(void);
const char* tfh_last_error(void);
/* Number of visible HIP devices (0 when none; never an error). */
int32_t tfh_device_count(int32_t* count);

/* ---- layout ------------------------------------------------------------- */
/* Reference layout (3,C,G,G)/(3,C,G) (tensor_vm.py:129-133) -> channel-last. */
int32_t tfh_vm_to_channel_last(const float* matrix, const float* vector, int64_t channels,
                               int64_t grid_dim, float* plane, float* line, tfh_stream_t stream);
/* Channel-last -> reference layout. */
int32_t tfh_vm_from_channel_last(const float* plane, const float* line, int64_t channels,
                                 int64_t grid_dim, float* matrix, float* vector,
                                 tfh_stream_t stream);

/* ---- TensorVM ----------------------------------------------------------- */
/* TensorVM.interpolate (tensor_vm.py:42-89, 140-167, 226-250).
 * ijk: (3, M) in [-1,1] (component stride M). out: (3C, M), channel m*C+c. */
int32_t tfh_vm_interpolate(const tfh_vm* vm, const float* ijk, int64_t num_points, float* out,
                           tfh_stream_t stream);
/* VJP of tfh_vm_interpolate w.r.t. the factors: grad_out (3C, M) is
 * scatter-added into grad (atomic FP32 adds, wave pre-reduced). */
int32_t tfh_vm_interpolate_bwd(const tfh_vm* vm, const float* ijk, int64_t num_points,
                               const float* grad_out, const tfh_vm_grad* grad,
                               tfh_stream_t stream);
/* TensorVM.resize (tensor_vm.py:91-100, 183-223): align-corners linear
 * upsampling to grid_dim_out >= G. Also used for the Adam moments
 * (training.py:257-275). out_plane (3,G',G',C), out_line (3,G',C). */
int32_t tfh_vm_resize(const tfh_vm* vm, int64_t grid_dim_out, float* out_plane, float* out_line,
                      tfh_stream_t stream);

/* ---- sampling and compositing ------------------------------------------ */
/* sample_points_along_ray_within_bbox vmapped over rays (render.py:177-186,
 * 350-434). jitter (N,) shared by all rays or NULL (no jitter).
 * points (3, R, N), ts (R, N), step_sizes (R, N); any output may be NULL. */
int32_t tfh_sample_bbox(const float* origins, const float* directions, int64_t num_rays,
                        const float* aabb, int64_t samples_per_ray, const float* jitter,
                        float* points, float* ts, float* step_sizes, tfh_stream_t stream);
/* Contracted sampler (render.py:122-173). base_ts/base_steps: (N,) schedule
 * (render.py:127-155, computed on the host in float64 and cast like the
 * reference); jitter (R, N). Outputs as tfh_sample_bbox. */
int32_t tfh_sample_contracted(const float* origins, const float* directions, int64_t num_rays,
                              int64_t samples_per_ray, const float* base_ts,
                              const float* base_steps, const float* jitter, float* points,
                              float* ts, float* step_sizes, tfh_stream_t stream);
/* compute_segment_probabilities (render.py:300-347): per-ray log-space
 * transmittance scan. sigmas, step_sizes, p_exits, p_terminates: (R, N). */
int32_t tfh_segment_probabilities(const float* sigmas, const float* step_sizes, int64_t num_rays,
                                  int64_t samples_per_ray, float* p_exits, float* p_terminates,
                                  tfh_stream_t stream);
/* VJP of compute_segment_probabilities w.r.t. the densities (the reference
 * differentiates it inside jax.value_and_grad, training.py:121-134):
 * with a_i = sigma_i delta_i,
 *   dL/da_i = gw_i E_i - sum_{j>i} gw_j w_j - sum_{j>=i} ge_j E_j,
 *   grad_sigmas_i = dL/da_i delta_i.
 * p_exits / p_terminates: the forward's outputs; either upstream gradient
 * (R, N) may be NULL, not both. sigmas is not read (may be NULL). */
int32_t tfh_segment_probabilities_bwd(const float* sigmas, const float* step_sizes,
                                      int64_t num_rays, int64_t samples_per_ray,
                                      const float* p_exits, const float* p_terminates,
                                      const float* grad_p_exits, const float* grad_p_terminates,
                                      float* grad_sigmas, tfh_stream_t stream);
/* Appearance sample selection (render.py:459-470): per ray the k indices of
 * the largest p_terminates (log p + gumbel when gumbel != NULL; gumbel is
 * (N,) shared by all rays), ties to the lower index, written ascending. */
int32_t tfh_topk_select(const float* p_terminates, int64_t num_rays, int64_t samples_per_ray,
                        int64_t k, const float* gumbel, int32_t* indices, tfh_stream_t stream);

/* ---- FeatureMlp --------------------------------------------------------- */
/* Workspace of tfh_feature_mlp_fwd (0 for FP32; BF16 packs the weights). */
size_t tfh_feature_mlp_fwd_workspace_size(const tfh_mlp* mlp, int32_t precision);
/* FeatureMlp.__call__ (networks.py:46-121): features (M, feat_dim),
 * viewdirs (M, 3), camera_indices (M) (may be NULL without embedding),
 * rgb (M, 3). */
int32_t tfh_feature_mlp_fwd(const tfh_mlp* mlp, const float* features, const float* viewdirs,
                            const uint32_t* camera_indices, int64_t num_rows, float* rgb,
                            int32_t precision, void* workspace, size_t workspace_bytes,
                            tfh_stream_t stream);
size_t tfh_feature_mlp_bwd_workspace_size(const tfh_mlp* mlp, int64_t num_rows,
                                          int32_t precision);
/* VJP of tfh_feature_mlp_fwd: grad_rgb (M,3) -> grad_features (M, feat_dim)
 * (may be NULL) and parameter gradients accumulated into grad. */
int32_t tfh_feature_mlp_bwd(const tfh_mlp* mlp, const float* features, const float* viewdirs,
                            const uint32_t* camera_indices, int64_t num_rows,
                            const float* grad_rgb, float* grad_features, const tfh_mlp_grad* grad,
                            int32_t precision, void* workspace, size_t workspace_bytes,
                            tfh_stream_t stream);

/* ---- render_rays -------------------------------------------------------- */
/* Workspace of tfh_render_fwd / tfh_render_bwd / tfh_train_step for R rays. */
size_t tfh_render_workspace_size(const tfh_render_config* cfg, const tfh_mlp* mlp,
                                 int64_t appearance_channels, int64_t num_rays);
/* Workspace that lets tfh_train_step keep the feature MLP's hidden
 * activations (1 KB per shaded sample, standard 48-channel shape, with or
 * without a camera embedding) between its forward and backward instead of
 * recomputing layers 1-3, plus 12 MB for copies of the factor-line gradients
 * that spread the scatters' atomics; with a tfh_render_workspace_size
 * workspace it recomputes the activations and adds straight into the
 * gradients. No counterpart in the reference (XLA keeps activations by
 * itself). */
size_t tfh_train_step_workspace_size(const tfh_render_config* cfg, const tfh_mlp* mlp,
                                     int64_t appearance_channels, int64_t num_rays);
/* render_rays (render.py:105-279): rays -> out (R,3) for RGB, (R) for the
 * distance modes. jitter: (N) bbox / (R,N) contraction, NULL = none.
 * base_ts/base_steps: contraction schedule (N) or NULL. gumbel: (N) or NULL.
 * The workspace keeps the residuals tfh_render_bwd consumes. */
int32_t tfh_render_fwd(const tfh_vm* density, const tfh_vm* appearance, const tfh_mlp* mlp,
                       const tfh_render_config* cfg, const float* aabb, const float* origins,
                       const float* directions, const uint32_t* camera_indices, int64_t num_rays,
                       const float* jitter, const float* base_ts, const float* base_steps,
                       const float* gumbel, float* out, void* workspace, size_t workspace_bytes,
                       tfh_stream_t stream);
/* VJP of tfh_render_fwd (RGB mode) for grad_out (R,3): accumulates the
 * density / appearance factor gradients and the MLP gradients. Must follow
 * tfh_render_fwd with the same arguments and workspace. */
int32_t tfh_render_bwd(const tfh_vm* density, const tfh_vm* appearance, const tfh_mlp* mlp,
                       const tfh_render_config* cfg, const float* aabb, const float* origins,
                       const float* directions, const uint32_t* camera_indices, int64_t num_rays,
                       const float* grad_out, const tfh_vm_grad* density_grad,
                       const tfh_vm_grad* appearance_grad, const tfh_mlp_grad* mlp_grad,
                       void* workspace, size_t workspace_bytes, tfh_stream_t stream);

/* tfh_render_fwd without a workspace argument, for callers that cannot
 * allocate device memory (code under HIP graph capture, serving loops,
 * harnesses that forbid hipMalloc). No counterpart in the reference. The
 * forward's residuals live in a device array of TFH_RENDER_NOALLOC_SCRATCH_BYTES
 * inside the library, one per device, held from the first call on that device
 * until the process ends; the call allocates nothing. A call whose residuals
 * do not fit returns TFH_ERR_WORKSPACE before any device work: per ray they
 * take about 32 N + 592 k bytes in RGB mode with 48 appearance channels
 * (N=512, k=64: ~2,400 rays per call). Results equal tfh_render_fwd's. Calls
 * on one device share the array: calls that overlap on one device (several
 * streams or host threads) race. The backward (tfh_render_bwd) needs the
 * workspace of tfh_render_fwd and cannot follow this call. */
#define TFH_RENDER_NOALLOC_SCRATCH_BYTES (128ull << 20)
int32_t tfh_render_fwd_noalloc(const tfh_vm* density, const tfh_vm* appearance,
                               const tfh_mlp* mlp, const tfh_render_config* cfg,
                               const float* aabb, const float* origins, const float* directions,
                               const uint32_t* camera_indices, int64_t num_rays,
                               const float* jitter, const float* base_ts,
                               const float* base_steps, const float* gumbel, float* out,
                               tfh_stream_t stream);

/* ---- optimizer ---------------------------------------------------------- */
/* One Adam step on a flat group (training.py:184-205, 213-243):
 * optax.scale_by_adam(b1, b2, eps, eps_root=0) -> scale(-lr) -> x decay ->
 * apply. count = Adam step after the increment (>= 1). When zero_grad != 0
 * the gradient is reset to 0 after it is read. */
int32_t tfh_adam_step(float* param, float* grad, float* exp_avg, float* exp_avg_sq, int64_t n,
                      float lr, float decay, float b1, float b2, float eps, int64_t count,
                      int32_t zero_grad, tfh_stream_t stream);

/* Fused TrainState.training_step (training.py:101-205): render (RGB) ->
 * MSE against colors (R,3) -> backward -> Adam on every group with the
 * per-group learning rates (lr_tensor for the factors, lr_mlp for the MLP)
 * times decay. params/grads/exp_avg/exp_avg_sq are the flat parameter
 * buffers described by tfh_train_layout; grads must be zero on entry and
 * are left zero. mse_out: device float (mean squared error of this step). */
typedef struct {
  float* params;     /* flat: density plane, line, appearance plane, line, MLP */
  float* grads;
  float* exp_avg;
  float* exp_avg_sq;
  int64_t total;     /* number of floats */
  int64_t tensor_count; /* first tensor_count floats use lr_tensor, the rest lr_mlp */
} tfh_train_buffers;

int32_t tfh_train_step(const tfh_vm* density, const tfh_vm* appearance, const tfh_mlp* mlp,
                       const tfh_render_config* cfg, const float* aabb, const float* origins,
                       const float* directions, const uint32_t* camera_indices,
                       const float* colors, int64_t num_rays, const float* jitter,
                       const float* base_ts, const float* base_steps, const float* gumbel,
                       const tfh_vm_grad* density_grad, const tfh_vm_grad* appearance_grad,
                       const tfh_mlp_grad* mlp_grad, const tfh_train_buffers* buffers,
                       float lr_tensor, float lr_mlp, float decay, int64_t adam_count,
                       float* mse_out, void* workspace, size_t workspace_bytes,
                       tfh_stream_t stream);

/* ---- proposal sampler ------------------------------------------------------
 * A learned proposal density field that places the field's samples. The
 * reference stops at a comment (render.py:126: "An occupancy grid or proposal
 * network would really help us here!"); nothing below has a reference
 * counterpart. The field is a VM decomposition (typically 64^3, C = 8 per
 * mode) whose 3C features feed a small MLP:
 *   z = w3 . relu(W2^T relu(W1^T f + b1) + b2) + b3,  sigma = softplus(z)
 * (kernels (in, out) row-major like the FeatureMlp's, hidden width 64).
 * Sampling runs rounds: round 0 takes samples_round0 strata of the
 * reference's sampler (the box segment split evenly, or the contracted
 * schedule); each round evaluates the proposal density at its samples,
 * composites their weights w_j over its intervals (the reference's
 * compute_segment_probabilities, render.py:300-347) and places the next
 * round's interval edges at the quantiles m/n' of the piecewise-linear CDF
 * of (w_j^power + pad / n_j) (the edges of the round's span are kept). The last
 * round's edges (R, samples_out + 1) are the field's intervals
 * (TFH_SAMPLING_EDGES). Training (mip-NeRF 360's interval loss) takes the
 * field's weights w_i on its intervals as constants and, for every round,
 * bound_i = sum of the round's w_j over the intervals overlapping interval i,
 *   L = weight / R * sum_rays sum_i max(0, w_i - bound_i)^2 / (w_i + 1e-7),
 * differentiated w.r.t. the proposal parameters only. */
typedef struct {
  const float* w1; /* (feat_dim, 64) */
  const float* b1; /* (64) */
  const float* w2; /* (64, 64) */
  const float* b2; /* (64) */
  const float* w3; /* (64, 1) */
  const float* b3; /* (1) */
  int64_t feat_dim; /* 3 * proposal channels, 1..32 */
} tfh_prop_mlp;

typedef struct {
  float *w1, *b1, *w2, *b2, *w3, *b3; /* accumulated into (+=) */
} tfh_prop_mlp_grad;

typedef struct {
  int32_t samples_round0;   /* strata of round 0, 2..1024 */
  int32_t samples_round1;   /* samples of round 1 (0 = one round only), <= 1024 */
  int32_t samples_out;      /* intervals handed to the field, 1..1024 */
  int32_t scene_contraction; /* as tfh_render_config */
  float histogram_pad;      /* pad mass added to each round's weights before resampling */
  float loss_weight;        /* weight of the interval loss */
  float weight_power;       /* resample on w^power (0: evenly, 1: the weights; a
                               training schedule ramps it from 0 to 1) */
} tfh_proposal_config;

/* The proposal MLP alone: features (M, feat_dim) -> z (M) (may be NULL) and
 * sigma = softplus(z) (M). */
int32_t tfh_prop_mlp_fwd(const tfh_prop_mlp* mlp, const float* features, int64_t num_rows,
                         float* z, float* sigma, tfh_stream_t stream);
size_t tfh_prop_mlp_bwd_workspace_size(const tfh_prop_mlp* mlp, int64_t num_rows);
/* VJP of tfh_prop_mlp_fwd w.r.t. z: grad_z (M) -> grad_features (M,
 * feat_dim) (may be NULL) and parameter gradients accumulated into grad. */
int32_t tfh_prop_mlp_bwd(const tfh_prop_mlp* mlp, const float* features, int64_t num_rows,
                         const float* grad_z, float* grad_features,
                         const tfh_prop_mlp_grad* grad, void* workspace, size_t workspace_bytes,
                         tfh_stream_t stream);

/* Inverse-CDF resampling of one round: for each ray, edges (n+1) and
 * weights (n) -> edges_out (num_out+1) at the quantiles m/num_out of the
 * piecewise-linear CDF of (w_j^power + pad/n) over the edges; edges_out[0]
 * and edges_out[num_out] repeat the ray's first and last edge. */
int32_t tfh_proposal_resample(const float* edges, const float* weights, int64_t num_rays,
                              int64_t num_intervals, int64_t num_out, float pad, float power,
                              float* edges_out, tfh_stream_t stream);
/* Interval loss per ray (see above) of field intervals edges (R, n+1) with
 * weights (R, n) against proposal intervals prop_edges (R, n_p+1) with
 * prop_weights (R, n_p), times scale: loss (R) and, when grad_prop_weights
 * != NULL, its gradient w.r.t. prop_weights (R, n_p). */
int32_t tfh_interval_loss(const float* edges, const float* weights, int64_t num_intervals,
                          const float* prop_edges, const float* prop_weights,
                          int64_t num_prop_intervals, int64_t num_rays, float scale, float* loss,
                          float* grad_prop_weights, tfh_stream_t stream);

/* Workspace of tfh_proposal_sample / tfh_proposal_loss_bwd for R rays. */
size_t tfh_proposal_workspace_size(const tfh_proposal_config* pc, const tfh_vm* field,
                                   int64_t num_rays);
/* Proposal rounds for R rays. jitter: (R, samples_round0 + samples_round1)
 * fractions in [0,1) placing each round's samples in their intervals (NULL:
 * midpoints). base_ts / base_steps: the contracted schedule of
 * samples_round0 entries when scene_contraction (render.py:127-155), else
 * NULL. edges_out: (R, samples_out + 1). The workspace keeps each round's
 * intervals, features and weights for tfh_proposal_loss_bwd. */
int32_t tfh_proposal_sample(const tfh_vm* field, const tfh_prop_mlp* mlp,
                            const tfh_proposal_config* pc, const float* aabb,
                            const float* origins, const float* directions, int64_t num_rays,
                            const float* jitter, const float* base_ts, const float* base_steps,
                            float* edges_out, void* workspace, size_t workspace_bytes,
                            tfh_stream_t stream);
/* Device pointers, inside a workspace filled by tfh_proposal_sample, to round
 * `round`'s interval edges (R, n_round + 1) and proposal weights (R, n_round)
 * (for inspection, e.g. weight histograms). No kernel runs. */
int32_t tfh_proposal_round(const tfh_proposal_config* pc, const tfh_vm* field, int64_t num_rays,
                           int32_t round, const void* workspace, size_t workspace_bytes,
                           const float** edges, const float** weights);
/* Interval loss of every round against the field's weights field_w (R, n)
 * on edges (R, n + 1) (the edges_out of tfh_proposal_sample, n =
 * samples_out): adds the loss to *loss_out (device float) and accumulates
 * the proposal gradients. Must follow tfh_proposal_sample with the same
 * arguments and workspace. */
int32_t tfh_proposal_loss_bwd(const tfh_vm* field, const tfh_prop_mlp* mlp,
                              const tfh_proposal_config* pc, int64_t num_rays, const float* edges,
                              const float* field_w, const tfh_vm_grad* field_grad,
                              const tfh_prop_mlp_grad* mlp_grad, float* loss_out, void* workspace,
                              size_t workspace_bytes, tfh_stream_t stream);

/* Fused training step with the proposal sampler: proposal rounds -> render
 * (RGB, TFH_SAMPLING_EDGES on the proposal's intervals) -> MSE -> backward
 * -> interval loss backward -> Adam on both parameter sets. cfg must have
 * sampling = TFH_SAMPLING_EDGES and density_samples_per_ray =
 * pc->samples_out. jitter: NULL, or the proposal's (R, samples_round0 +
 * samples_round1) fractions followed by the field's (R, N) fractions;
 * base_ts / base_steps: the proposal's round-0 schedule (contraction only).
 * The field's buffers are as tfh_train_step's; prop_buffers holds the
 * proposal factors (first tensor_count floats, lr_tensor) then the proposal
 * MLP (lr_mlp). Both gradients must be zero on entry and are left zero.
 * mse_out and prop_loss_out: device floats, overwritten. */
size_t tfh_train_step_proposal_workspace_size(const tfh_render_config* cfg, const tfh_mlp* mlp,
                                              int64_t appearance_channels,
                                              const tfh_proposal_config* pc,
                                              const tfh_vm* field, int64_t num_rays);
int32_t tfh_train_step_proposal(const tfh_vm* density, const tfh_vm* appearance,
                                const tfh_mlp* mlp, const tfh_render_config* cfg,
                                const tfh_vm* field, const tfh_prop_mlp* prop_mlp,
                                const tfh_proposal_config* pc, const float* aabb,
                                const float* origins, const float* directions,
                                const uint32_t* camera_indices, const float* colors,
                                int64_t num_rays, const float* jitter, const float* base_ts,
                                const float* base_steps, const float* gumbel,
                                const tfh_vm_grad* density_grad,
                                const tfh_vm_grad* appearance_grad, const tfh_mlp_grad* mlp_grad,
                                const tfh_vm_grad* field_grad,
                                const tfh_prop_mlp_grad* prop_mlp_grad,
                                const tfh_train_buffers* buffers,
                                const tfh_train_buffers* prop_buffers, float lr_tensor,
                                float lr_mlp, float decay, int64_t adam_count, float* mse_out,
                                float* prop_loss_out, void* workspace, size_t workspace_bytes,
                                tfh_stream_t stream);

/* ---- XLA custom calls ------------------------------------------------------
 * render_rays's forward and VJP and the fused training step in XLA's original
 * GPU custom-call form (api_version 0), so that the reference's JAX loop,
 * which differentiates render.render_rays (render.py:105-279) inside
 * jax.value_and_grad (training.py:153-156), can call the kernels through
 * jax.ffi.register_ffi_target(name, capsule, platform="ROCM", api_version=0)
 * without XLA headers:
 *   void f(hipStream_t stream, void** buffers, const char* opaque,
 *          size_t opaque_len);
 * buffers holds the device pointers of the operands in operand order, then
 * of the results in result order. opaque is an ASCII descriptor (XLA's
 * backend_config) of space-separated key=value fields:
 *   tfhx=1 near far mode N k contraction precision sampling shift   (tfh_render_config)
 *   R dG dC aG aC squash units nf nv ncam                          (rays, factors, MLP)
 *   jitter schedule gumbel (0/1: operand present) ws (workspace bytes)
 *   offsets=o0,...,o11 total tcount                   (tfh_xla_train_step only)
 * floats printed with enough digits to round-trip float32. Factor tensors
 * are channel-last as everywhere in this header. An absent optional operand
 * (embedding, jitter, schedule, gumbel) still takes its slot; any buffer does.
 * The form has no status channel: on a bad descriptor or a failed call the
 * entry point fills its float results with NaN, prints the message to stderr
 * and records it as the calling thread's tfh_last_error().
 *
 * tfh_xla_render_fwd (tfh_render_fwd):
 *   operands 0 density plane, 1 density line, 2 appearance plane,
 *     3 appearance line, 4 w0, 5 w1, 6 b1, 7 w2, 8 b2, 9 w3, 10 b3, 11 embed,
 *     12 aabb (2,3), 13 origins (R,3), 14 directions (R,3),
 *     15 camera_indices (R) u32, 16 jitter, 17 base_ts, 18 base_steps, 19 gumbel
 *   results 20 out (R,3) or (R), 21 workspace u8[ws] (ws =
 *     tfh_render_workspace_size; the residuals the VJP reads)
 * tfh_xla_render_bwd (tfh_render_bwd, RGB mode):
 *   operands 0-15 as above, 16 grad_out (R,3), 17 the forward's workspace
 *   results 18 density plane grad, 19 density line grad, 20 appearance plane
 *     grad, 21 appearance line grad, 22-28 w0 b... grads in operand order
 *     4-10, 29 embed grad, 30 workspace u8[ws] (alias it to operand 17 with
 *     input_output_aliases; otherwise operand 17 is copied into it first)
 *   The gradients are written, not accumulated.
 * tfh_xla_train_step (tfh_train_step):
 *   operands 0 params, 1 exp_avg, 2 exp_avg_sq (flat f32[total], views at
 *     offsets: density plane, line, appearance plane, line, w0, w1, b1, w2,
 *     b2, w3, b3, embed; the first tcount floats use the factor rate),
 *     3 aabb, 4 origins, 5 directions, 6 camera_indices, 7 colors (R,3),
 *     8 jitter, 9 base_ts, 10 base_steps, 11 gumbel,
 *     12 hyper f32[4] = {-lr_tensor * decay, -lr_mlp * decay, 1 - 0.9^count,
 *     1 - 0.99^count} (the Adam step's scalars, computed by the caller)
 *   results 13 params, 14 exp_avg, 15 exp_avg_sq (alias them to operands
 *     0-2; otherwise the operands are copied first), 16 mse f32[1],
 *     17 gradient scratch f32[total], 18 workspace u8[ws] (ws =
 *     tfh_train_step_workspace_size) */
void tfh_xla_render_fwd(void* stream, void** buffers, const char* opaque, size_t opaque_len);
void tfh_xla_render_bwd(void* stream, void** buffers, const char* opaque, size_t opaque_len);
void tfh_xla_train_step(void* stream, void** buffers, const char* opaque, size_t opaque_len);

/* ---- synthetic scenes ----------------------------------------------------- */
/* Ground truth of a procedural scene (replaces the image datasets of data.py):
 * density = sum of isotropic Gaussian blobs, blobs[b] = {cx, cy, cz, sigma,
 * peak, r, g, b} (float64), colour = density-weighted albedo, white
 * background, emission-absorption midpoint quadrature in FP64 with
 * samples_per_ray samples over [t_near, t_far] (uniform in t, or with
 * contracted_far != 0 uniform in t up to t=1 and uniform in 1/t beyond).
 * rgb (R,3); opacity (R) and median_depth (R) may be NULL. median_depth is
 * the distance at which the transmittance crosses 1/2 (density constant
 * within a quadrature step), +inf when it never does. No counterpart in the
 * reference. */
int32_t tfh_blob_scene_render(const double* blobs, int32_t num_blobs, const float* origins,
                              const float* directions, int64_t num_rays, double t_near,
                              double t_far, int32_t samples_per_ray, int32_t contracted_far,
                              float* rgb, float* opacity, float* median_depth,
                              tfh_stream_t stream);

/* ---- stage timing --------------------------------------------------------- */
/* Per-stage timing with HIP events recorded on the launching stream around
 * every kernel the library launches. tfh_profile_enable(1) clears and starts
 * recording, (0) stops. tfh_profile_read synchronizes the recorded events
 * and returns, per stage name (at most `max` stages, names NUL-terminated in
 * 48-byte slots), the summed milliseconds and the launch count. No
 * counterpart in the reference (it has no profiling hooks). */
int32_t tfh_profile_enable(int32_t on);
int32_t tfh_profile_read(int32_t max, char* names, double* total_ms, int64_t* counts,
                         int32_t* num_stages);

/* ---- utilities ---------------------------------------------------------- */
/* Counter-based uniform [0,1) floats: out[i] = U(seed, offset + i). The
 * explicit-jitter replacement of jax.random.uniform (render.py:158-160,
 * 375-379). */
int32_t tfh_uniform(uint64_t seed, uint64_t offset, int64_t n, float* out, tfh_stream_t stream);

#ifdef __cplusplus
}
#endif
#endif /* TENSORF_HIP_H */
