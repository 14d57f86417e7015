/*
 * fuzzy_ellipses_amd.h -- C ABI of the MI355X fuzzy-metaball rasterizer.
 *
 * Plain C: pointers, sizes and POD structs only.  Every entry point is asynchronous on the
 * caller's HIP stream, performs no host synchronisation on the render path and returns a
 * status code instead of exiting or throwing.
 *
 * Reference interfaces replaced (paths are relative to the reference repository):
 *   - XLA legacy custom-call target "rasterize_gaussians_fwd",
 *     void(cudaStream_t, void** buffers, const char* opaque, size_t)   rasterize_main.h:26-30,
 *     registered from ext.cpp:22-32, descriptor rasterize_main.h:13-19.
 *   - The math it is meant to reproduce: baseline/fm_render.py:11-126 (render_func_rays,
 *     render_func_quat).
 *   - Gradients: the reference has no backward kernel; callers used jax.grad of
 *     baseline/fm_render.py (shape_from_silhouette.ipynb, "optim_bwd").
 *
 * Conventions (same as baseline/fm_render.py):
 *   means       (N,3) float32 row-major, world coordinates.
 *   prec        (N,3,3) float32 row-major; only the upper triangle is read (triu), the
 *               density of Gaussian j at x is w_j - 0.5 |U_j^T (x - mu_j)|^2.
 *   weights_log (N,)  float32 log-weights w_j.
 *   camera      pinhole intrinsics + pose: the ray of pixel (x, y) is
 *               r = ((x-cx)/fx, (y-cy)/fy, 1) @ R, starting at t; pixels row-major y*W+x.
 *   outputs     depth (H,W), alpha (H,W) float32, exactly render_func_rays' outputs
 *               (no nan_to_num; the reference binding applies it afterwards).
 * All array pointers are device pointers; structs are read on the host during the call.
 * `stream` is a hipStream_t passed as void* so this header needs no HIP include
 * (NULL = the default stream).
 */
#ifndef FUZZY_ELLIPSES_AMD_H
#define FUZZY_ELLIPSES_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum fe_status {
  FE_OK = 0,
  FE_ERR_INVALID_ARGUMENT = 1,  /* bad sizes, null pointers, unsupported tile size */
  FE_ERR_WORKSPACE = 2,         /* workspace missing or smaller than fe_plan reported */
  FE_ERR_HIP = 3,               /* a HIP runtime call failed (see fe_last_error_string) */
  FE_ERR_NO_DEVICE = 4
} fe_status;

/* Pinhole camera.  R is row-major; its rows are the camera x, y, z axes in world
 * coordinates (world ray = camera ray @ R, baseline/fm_render.py:116). */
typedef struct fe_camera {
  float R[9];
  float t[3];
  float fx, fy, cx, cy;
  int32_t height, width;
} fe_camera;

/* One scene: N Gaussians in device memory. */
typedef struct fe_scene {
  int32_t n;
  const float* means;        /* (N,3)   */
  const float* prec;         /* (N,3,3) upper triangle read */
  const float* weights_log;  /* (N,)    */
  /* Optional per-Gaussian features (N, n_features) row-major (colour, descriptors, ...), or
   * NULL / 0.  Views of this scene with a feature image get, per pixel, the features weighted
   * by render_func_rays' normalised depth weights w (baseline/fm_render.py:96-109):
   * F = sum_j w_j f_j -- the reference's w contracted with f, without materialising w (N, P). */
  const float* features;
  int32_t n_features;
} fe_scene;

/* One view of a scene: camera, beta hyper-parameters and output images. */
typedef struct fe_view {
  int32_t scene;   /* index into the scene array */
  fe_camera cam;
  float beta2, beta3;  /* culling bounds need finite beta2 >= 0 and beta3 >= 0; a view with
                          other values is rendered with every pair listed (same results) */
  float* depth;    /* (H,W) device, written by fe_forward */
  float* alpha;    /* (H,W) device, written by fe_forward */
  const float* pose_device; /* optional device array R[9] (row-major) then t[3]; when set it
                               replaces cam.R / cam.t, read on the device (the reference
                               passed R and t as device buffers, rasterize_main.cu:60-61) */
  /* Optional ray list (NULL = the pinhole pixel grid).  When set, the view renders n_rays
   * rays whose camera-frame directions are (rays[2k], rays[2k+1], 1), rotated by the pose
   * (world ray = camera ray @ R) and starting at t -- the per-ray form of
   * baseline/fm_render.py:151-166 (render_func_idx_quattrans) for the rays of one pose.
   * cam.fx/fy/cx/cy/height/width are then ignored and depth/alpha hold n_rays values.
   * Consecutive groups of tile*tile rays share one culling cone, so rays ordered by image
   * position cull best; any order gives the same images to within the culling budgets. */
  const float* rays;
  int64_t n_rays;
  /* Optional feature image (H*W or n_rays, n_features) row-major, written by fe_forward when
   * the view's scene has features (NULL = not wanted).  fe_backward reads it back. */
  float* features_out;
} fe_view;

/* Gradient buffers of one scene (device, accumulated into: zero them first). */
typedef struct fe_scene_grad {
  float* d_means;        /* (N,3)   */
  float* d_prec;         /* (N,3,3) only the upper triangle is written (lower stays as given) */
  float* d_weights_log;  /* (N,)    */
  float* d_features;     /* (N, n_features) or NULL */
} fe_scene_grad;

/* Upstream gradients of one view (device, (H,W) each; any may be NULL = zero). */
typedef struct fe_view_cotangent {
  const float* d_depth;
  const float* d_alpha;
  const float* d_features;   /* (H*W or n_rays, n_features) cotangent of the feature image */
} fe_view_cotangent;

enum {
  FE_MODE_CULLED = 0,  /* two-phase error-bounded tile culling (default) */
  FE_MODE_DENSE = 1    /* every tile evaluates every Gaussian */
};

enum {
  /* The workspace already holds this call's descriptors, written by an earlier call with
   * the same scenes, views, options and buffers: skip the host-to-device upload.  With this
   * flag (and profiling off) a call only enqueues kernels, so it can be captured into a
   * hipGraph and replayed. */
  FE_FLAG_REUSE_DESCRIPTORS = 1
};

typedef struct fe_options {
  int32_t tile;        /* tile edge in pixels: 8, 16 or 32 (0 = 16) */
  int32_t batch;       /* LDS batch depth in Gaussians: 32, 64, 128 or 256 (0 = 64) */
  int32_t mode;        /* FE_MODE_* */
  int32_t save_for_backward; /* keep per-pixel aggregates for fe_backward */
  int32_t flags;       /* FE_FLAG_* */
  /* Culling error budgets (0 = default).  Pairs are dropped only where the dropped alpha
   * mass is provably at most alpha_eps (absolute, per pixel; default 2e-5) and the dropped
   * share of the depth softmax weight at most depth_eps (default 1e-6). */
  float alpha_eps;
  float depth_eps;
} fe_options;

/* Diagnostics written by fe_forward into the workspace and readable with fe_read_stats. */
typedef struct fe_stats {
  int64_t pairs_phase1;      /* (tile, Gaussian) pairs listed by the alpha bound */
  int64_t pairs_phase2;      /* pairs added by the depth bound */
  int64_t evals;             /* ray-Gaussian evaluations performed (pixels x listed) */
  int64_t evals_phase1;      /* of which in the phase-1 (or dense) blend */
  int64_t floor_tiles;       /* tiles completed densely because the 1e-20 weight floor could matter */
  int64_t floor_unresolved;  /* tiles where it still could after phase 2 (always 0; a self-check) */
} fe_stats;

/* ---- library ---- */
int32_t fe_version(void);
const char* fe_status_string(fe_status s);
const char* fe_last_error_string(void);

/* ---- planning: pure host arithmetic, no device access ----
 * Reports the workspace bytes fe_forward/fe_backward need for this batch.  The workspace
 * holds job descriptors, per-Gaussian records, tile masks and lists (sized for the worst
 * case, so no host read-back is ever needed) and, with save_for_backward, per-pixel
 * aggregates.  Replaces the per-call GpuMemoryPool of rasterizer_impl.cu:47-79. */
fe_status fe_plan(const fe_scene* scenes, int32_t n_scenes, const fe_view* views, int32_t n_views,
                  const fe_options* opt, size_t* workspace_bytes);

/* ---- forward ----
 * Renders every view into its depth/alpha buffers.  Replaces RasterizeGaussiansCUDAJAX
 * (rasterize_main.cu:29-114) and Rasterizer::forwardJAX (rasterizer_impl.cu:174-344). */
fe_status fe_forward(void* stream, const fe_scene* scenes, int32_t n_scenes,
                     const fe_view* views, int32_t n_views, const fe_options* opt,
                     void* workspace, size_t workspace_bytes);

/* ---- backward ----
 * Given the workspace of the matching fe_forward (save_for_backward=1, untouched since),
 * accumulates d(loss)/d(means, prec upper triangle, weights_log) for every scene.
 * No reference counterpart: the reference took gradients from jax.grad of
 * baseline/fm_render.py:11-109. */
fe_status fe_backward(void* stream, const fe_scene* scenes, int32_t n_scenes,
                      const fe_view* views, int32_t n_views, const fe_options* opt,
                      const fe_view_cotangent* cot, const fe_scene_grad* grads,
                      void* workspace, size_t workspace_bytes);

/* Copies the diagnostics of the last fe_forward on this workspace to host (synchronises
 * the stream; not for the hot path). */
fe_status fe_read_stats(void* stream, const void* workspace, size_t workspace_bytes,
                        fe_stats* out);

/* ---- kernel timing (diagnostics) ----
 * When enabled, every kernel launch of fe_forward/fe_backward is bracketed by hipEvents on
 * the caller's stream; fe_profile_read synchronises those events and reports the summed
 * durations and launch counts per kernel, then optionally resets them.
 * fe_profile_enable(0) turns timing off, fe_profile_enable(1) times every kernel, and
 * fe_profile_enable(FE_PROFILE_SELECT | bits) times only the kernels k with (1 << k) set
 * (each timed launch adds two event records to the stream). */
enum {
  FE_K_PREPROCESS = 0, FE_K_MASK1, FE_K_SCAN1, FE_K_WRITE1, FE_K_BLEND1, FE_K_MASK2, FE_K_SCAN2,
  FE_K_WRITE2, FE_K_BLEND2, FE_K_BLEND_DENSE, FE_K_BWD_BLEND, FE_K_BWD_FINALIZE, FE_K_COUNT
};
#define FE_PROFILE_SELECT 0x40000000
typedef struct fe_kernel_times {
  double ms[FE_K_COUNT];
  int64_t launches[FE_K_COUNT];
} fe_kernel_times;
fe_status fe_profile_enable(int32_t on);
fe_status fe_profile_read(fe_kernel_times* out, int32_t reset);

/* ---- single-view convenience: the drop-in for the reference custom call ----
 * Same arguments as render_func_quat with the rotation given as a matrix. */
fe_status fe_render_workspace_bytes(int32_t n, int32_t height, int32_t width, int32_t tile,
                                    size_t* bytes);
fe_status fe_render_fwd(void* stream, int32_t n, const float* means, const float* prec,
                        const float* weights_log, const fe_camera* cam, float beta2,
                        float beta3, float* depth, float* alpha, void* workspace,
                        size_t workspace_bytes);

/* ---- fe_render_fwd without a workspace: allocates nothing ----
 * No counterpart in the reference.  For callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc): the same render as
 * fe_render_fwd (same arguments but the workspace, bit-identical images), with its workspace in a
 * static device array of FE_NOALLOC_SCRATCH_BYTES that each GPU the library runs on holds until
 * the process ends (32 MiB: enough for N = 4096 Gaussians at 480x640; a view that needs more is
 * refused with FE_ERR_WORKSPACE, and fe_render_workspace_bytes tells the need).  The call makes no
 * hipMalloc/hipFree or host staging allocation: the view's descriptors reach the device as a
 * kernel argument.  Every call on one GPU shares that GPU's array, so calls that overlap on one
 * GPU (different streams, or host threads) race; calls in order on one stream do not.  Nothing
 * in the array needs clearing between calls: every call rewrites what it reads. */
#define FE_NOALLOC_SCRATCH_BYTES (32u << 20)
fe_status fe_render_fwd_noalloc(void* stream, int32_t n, const float* means, const float* prec,
                                const float* weights_log, const fe_camera* cam, float beta2,
                                float beta3, float* depth, float* alpha);

/* ---- the reference's XLA custom-call target, unchanged in shape ----
 * Replaces RasterizeGaussiansCUDAJAX (rasterize_main.h:26-30, rasterize_main.cu:29-114),
 * which ext.cpp:22-32 exports as the capsule "rasterize_gaussians_fwd" and
 * fuzzy_ellipses/__init__.py:22-23,111-166 registers and lowers.  Same signature, same
 * 11 device buffers in the same order, same 20-byte opaque descriptor
 * (rasterize_main.h:13-19, packed by ext.cpp:34-42 / bindings.h:22-37):
 *   buffers[0] means (N,3)   [1] log-weights (N,)   [2] scales (N,3)   [3] quaternions (N,4),
 *              scalar first
 *   [4] camera rays (H*W,3), camera frame, row-major pixels   [5] R (3,3)   [6] t (3,)
 *   [7] view matrix (4,4)   [8] projection matrix (4,4)   (7, 8 and tan_fov are accepted and
 *       unused: they fed the reference's culling heuristic, not the math)
 *   [9] depth (H,W) written   [10] alpha (H,W) written
 * The precision factor is computed on the device from scales and quaternions (U upper
 * triangular, U U^T = (R(q) diag(s^2) R(q)^T)^-1, R(q) = quat_to_rot of
 * baseline/util_render.py:51-76); beta2 = 21.4 and beta3 = 2.66 as the reference fixes them
 * (rasterizer_impl.cu:196-197).  A ray (x, y, z) is used as given; for z != 1 alpha is exact
 * and depth is NaN (t*, and so the depth softmax, depends on the ray's length, which the
 * camera-plane ray list does not carry).  N = 0 writes zeros, as the reference does.
 * Asynchronous on `stream`; scratch memory is library-owned per (device, stream) and grows
 * (with one stream synchronisation) only when a call needs more than any before.
 * rasterize_gaussians_fwd returns nothing, like the reference; on failure it fills both
 * outputs with NaN and leaves the reason in fe_last_error_string.
 * fe_rasterize_gaussians_fwd is the same call returning the status. */
typedef struct fe_rasterize_descriptor {
  int32_t image_height;
  int32_t image_width;
  int32_t P;            /* number of Gaussians */
  float tan_fovx;
  float tan_fovy;
} fe_rasterize_descriptor;

void rasterize_gaussians_fwd(void* stream, void** buffers, const char* opaque, size_t opaque_len);
fe_status fe_rasterize_gaussians_fwd(void* stream, void** buffers, const char* opaque, size_t opaque_len);

/* ---- shape-from-silhouette caller: objective + cotangents, Adam ----
 * The optimisation loop that calls fe_forward / fe_backward each step
 * (shape_from_silhouette.ipynb: "jax_objective" = mean BCE of alpha clipped to
 * [3e-8, 1 - 3e-8] against the target silhouette; optimizers.adam; DegradeLR schedule in
 * utils/sfs_utils.py:62-131, which stays on the host).  Here the objective of one view is
 *   sum_p BCE(clip(alpha), mask) / P + w_depth * sum_p mask |depth - target_depth| / max(sum mask, 1)
 * and the loss is the mean over views.  The cotangents written are d(loss)/d(alpha) (zero where
 * the clip is active) and d(loss)/d(depth) (sign(0) = 0), ready for fe_backward. */
typedef struct fe_sil_view {
  const float* alpha;         /* (P,) rendered alpha of the view */
  const float* depth;         /* (P,) rendered depth */
  const float* mask;          /* (P,) target silhouette, values in [0, 1] */
  const float* target_depth;  /* (P,) */
  float* d_alpha;             /* (P,) written */
  float* d_depth;             /* (P,) written */
  int64_t n_pixels;           /* P > 0 */
} fe_sil_view;

/* Workspace bytes of fe_silhouette_loss for n_views views (descriptors and partial sums). */
fe_status fe_silhouette_workspace_bytes(int32_t n_views, size_t* bytes);

/* Writes the cotangents of every view and the scalar loss (device float, overwritten).
 * The sums run in a fixed order: the loss is bitwise reproducible.  `flags` may hold
 * FE_FLAG_REUSE_DESCRIPTORS (same views as the previous call on this workspace: nothing is
 * uploaded, so the call can be captured into a hipGraph). */
fe_status fe_silhouette_loss(void* stream, const fe_sil_view* views, int32_t n_views, float w_depth,
                             float clip, float* loss, int32_t flags, void* workspace,
                             size_t workspace_bytes);

/* One Adam update of n float parameters in place (jax.example_libraries.optimizers.adam, as
 * used by the notebook): m = b1 m + (1-b1) g; v = b2 v + (1-b2) g^2;
 * p -= lr (m / (1-b1^step)) / (sqrt(v / (1-b2^step)) + eps).  step counts from 1. */
fe_status fe_adam_step(void* stream, int64_t n, float* param, const float* grad, float* m, float* v,
                       float lr, float beta1, float beta2, float eps, int32_t step);

/* ---- Gaussian-mixture log-likelihood of a point cloud ----
 * Replaces baseline/fm_render.py:231-259 (log_likelihood(params, points)): with the mixture
 * weights softmax(weights_log) and the precision factor's upper triangle U (U U^T = Sigma^-1),
 *   res[i][j] = ln N(points_j; means_i, Sigma_i) + ln pi_i,   point_lse[j] = logsumexp_i res[i][j],
 *   nll = -mean_j point_lse[j].
 * fe_gmm_forward writes point_lse (M,), nll (device scalar) and, when res is not NULL, the
 * full (N, M) matrix res that the reference returns.  fe_gmm_backward, given the workspace of
 * the matching forward and its point_lse, accumulates d_nll * d(nll)/d(means, prec upper
 * triangle, weights_log) into the gradient buffers (zero them first).  No reference
 * counterpart for the backward: callers used jax.grad.  Both are deterministic. */
fe_status fe_gmm_workspace_bytes(int32_t n, int64_t m, size_t* bytes);
fe_status fe_gmm_forward(void* stream, int32_t n, const float* means, const float* prec,
                         const float* weights_log, int64_t m, const float* points, float* point_lse,
                         float* res, float* nll, void* workspace, size_t workspace_bytes);
fe_status fe_gmm_backward(void* stream, int32_t n, const float* prec, const float* weights_log,
                          int64_t m, const float* points, const float* point_lse, float d_nll,
                          float* d_means, float* d_prec, float* d_weights_log, void* workspace,
                          size_t workspace_bytes);

/* ---- multi-pose ray batches: optical flow ----
 * Replaces the flow half of baseline/fm_render.py:170-224 (render_func_idx_quattrans_flow).
 * pixel_posei (n_rays, 4) holds (x, y, z, pose index) per ray, the camera-frame direction
 * being (x invF, y invF, z); depth (n_rays,) is the rendered depth of each ray (ray parameter of
 * the surface point, as render_func_rays returns it); rot (n_poses, 3, 3) row-major and
 * trans (n_poses, 3) are the poses (world ray = camera ray @ R, origin t).  Each ray's point
 * is projected into the next and the previous pose (indices clamped to [0, n_poses-1]) and
 * flow = (x, y) - projection is written to flow_next / flow_prev (n_rays, 2).  The pose index
 * is truncated toward zero and clamped.  Device pointers; asynchronous on the stream. */
fe_status fe_ray_flow(void* stream, int64_t n_rays, const float* pixel_posei, float inv_f,
                      int32_t n_poses, const float* rot, const float* trans, const float* depth,
                      float* flow_next, float* flow_prev);

#ifdef __cplusplus
}
#endif

#endif /* FUZZY_ELLIPSES_AMD_H */
