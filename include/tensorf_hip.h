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

#define TFH_ABI_VERSION 1

typedef void* tfh_stream_t;

enum tfh_status {
  TFH_OK = 0,
  TFH_ERR_INVALID_ARGUMENT = -1,
  TFH_ERR_UNSUPPORTED = -2,
  TFH_ERR_NO_DEVICE = -3,
  TFH_ERR_LAUNCH = -4,
  TFH_ERR_WORKSPACE = -5
};

/* render.RenderMode (render.py:18-23) */
enum tfh_render_mode { TFH_MODE_RGB = 0, TFH_MODE_DIST_MEDIAN = 1, TFH_MODE_DIST_MEAN = 2 };

/* Arithmetic of the appearance MLP. FP32: FP32 MFMA. BF16: BF16 operands on
 * MFMA with FP32 accumulation. */
enum tfh_precision { TFH_PREC_FP32 = 0, TFH_PREC_BF16 = 1 };

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
  int32_t reserved;
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

/* ---- synthetic scenes ----------------------------------------------------- */
/* Ground truth of a procedural scene (replaces the image datasets of data.py):
 * density = sum of isotropic Gaussian blobs, blobs[b] = {cx, cy, cz, sigma,
 * peak, r, g, b} (float64), colour = density-weighted albedo, white
 * background, emission-absorption midpoint quadrature in FP64 with
 * samples_per_ray samples over [t_near, t_far] (uniform in t, or with
 * contracted_far != 0 uniform in t up to t=1 and uniform in 1/t beyond).
 * rgb (R,3); opacity (R) may be NULL. No counterpart in the reference. */
int32_t tfh_blob_scene_render(const double* blobs, int32_t num_blobs, const float* origins,
                              const float* directions, int64_t num_rays, double t_near,
                              double t_far, int32_t samples_per_ray, int32_t contracted_far,
                              float* rgb, float* opacity, tfh_stream_t stream);

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
