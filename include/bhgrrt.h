/*
 * bhgrrt.h -- C ABI of the MI355X (gfx950) Kerr Stokes hot path.
 *
 * Plain pointers and sizes only.  Every entry point returns an int status
 * (BH_OK == 0, negative on error); bh_last_error() returns the calling
 * thread's message for the last failure.  Messages of the drop-in
 * micro-kernels are the reference's ValueError texts.
 *
 * Reference interfaces replaced (paths relative to the reference tree):
 *   blackhole_native (PyO3 module, native/core/src/lib.rs:368-376), called from
 *   python/blackhole_sim/native_kernels.py:252-276, 302-322, 359-407 and
 *   python/blackhole_sim/benchmark.py:135-151, 238-247.
 */
#ifndef BHGRRT_H
#define BHGRRT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define BH_OK            0
#define BH_ERR_VALUE    -1   /* invalid argument: the reference raises ValueError */
#define BH_ERR_DEVICE   -2   /* HIP runtime failure or no usable GPU */
#define BH_ERR_MEMORY   -3   /* device allocation failed */
#define BH_ERR_STATE    -4   /* bad handle / object state */

/* invalid_policy of native_kernels.py:16, 127-130, 392-396 */
#define BH_POLICY_NAN      0
#define BH_POLICY_INITIAL  1
#define BH_POLICY_ZERO     2

/* ---- module identity ----------------------------------------------------
 * bh_core_version replaces core_version() (lib.rs:19-22).
 * bh_detect_arch replaces detect_arch() (lib.rs:24-35): writes
 * "arch=<a>;os=<o>;family=<f>;pointer_width=<n>" into buf.
 */
const char *bh_core_version(void);
int bh_detect_arch(char *buf, size_t buflen);
const char *bh_last_error(void);
int bh_device_count(int *count);
int bh_device_name(int device, char *buf, size_t buflen, int *compute_units);

/* ---- drop-in FP64 micro-kernels (host buffers, synchronous) ------------- */

/* stokes_rk2_brick(coeffs, ds_cm, initial) -- lib.rs:227-279.
 * coeffs: n_coeffs = 11 * cells AoS values; initial: 4 or 4 * cells values;
 * out: 4 * cells values. */
int bh_stokes_rk2_brick(const double *coeffs, size_t n_coeffs, double ds_cm,
                        const double *initial, size_t n_initial, double *out);

/* sample_brick_trilinear(coeffs, r_grid, theta_grid, phi_grid, points) -- lib.rs:281-306.
 * coeffs: AoS (nr, nt, np, 11); points: n_points_flat = 3 * n values;
 * out: 11 * n values, NaN rows outside the r/theta domain. */
int bh_sample_brick_trilinear(const double *coeffs, size_t n_coeffs,
                              const double *r_grid, size_t nr,
                              const double *theta_grid, size_t nt,
                              const double *phi_grid, size_t np,
                              const double *points, size_t n_points_flat,
                              double *out);

/* sample_and_step_stokes(coeffs, grids, points, ds_cm, initial) -- lib.rs:308-366,
 * with the wrapper's invalid_policy applied on the device
 * (native_kernels.py:373-397).  out: 4 * n values. */
int bh_sample_and_step_stokes(const double *coeffs, size_t n_coeffs,
                              const double *r_grid, size_t nr,
                              const double *theta_grid, size_t nt,
                              const double *phi_grid, size_t np,
                              const double *points, size_t n_points_flat,
                              double ds_cm, const double *initial, size_t n_initial,
                              int invalid_policy, double *out);

/* sample_brick_valid_mask(r_grid, theta_grid, points) -- native_kernels.py:189-210.
 * out: n bytes, 1 where r and theta lie inside the closed grid ranges. */
int bh_sample_brick_valid_mask(const double *r_grid, size_t nr,
                               const double *theta_grid, size_t nt,
                               const double *points, size_t n_points_flat,
                               uint8_t *out);

/* ---- device contexts ----------------------------------------------------
 * A context owns one HIP stream on one device, uploaded bricks, a small
 * ray-queue counter block and timing events.  Calls on one context are not
 * thread-safe; use one context per host thread.
 */
typedef struct bh_ctx bh_ctx;

int bh_ctx_create(int device, bh_ctx **out);
int bh_ctx_destroy(bh_ctx *ctx);
int bh_ctx_synchronize(bh_ctx *ctx);
/* average duration (ms) of the kernels launched by the last render/sampler call */
int bh_ctx_last_kernel_ms(bh_ctx *ctx, double *ms);

int bh_device_alloc(bh_ctx *ctx, size_t bytes, void **ptr);
int bh_device_free(bh_ctx *ctx, void *ptr);
int bh_copy_to_device(bh_ctx *ctx, void *dst, const void *src, size_t bytes);
int bh_copy_to_host(bh_ctx *ctx, void *dst, const void *src, size_t bytes);
/* device-to-device copy on the context stream (synchronous) */
int bh_copy_device(bh_ctx *ctx, void *dst, const void *src, size_t bytes);

/* Device-resident FP64 sampler over validated inputs (no host copies, no
 * allocation): same arithmetic as bh_sample_brick_trilinear.  Used for the
 * large-N sampler measurement. */
int bh_sample_brick_trilinear_device(bh_ctx *ctx, const double *d_coeffs,
                                     const double *d_r, int32_t nr,
                                     const double *d_theta, int32_t nt,
                                     const double *d_phi, int32_t np,
                                     const double *d_points, int64_t n,
                                     double *d_out);

/* ---- coefficient bricks (CoefficientBrickGrid, coefficient_bricks.py:28-75) */
#define BH_STORE_F32   0
#define BH_STORE_BF16  1
#define BH_STORE_F16   2

/* Device layouts of a brick:
 *  BH_LAYOUT_SOA     channel planes [11][nr][nt][np];
 *  BH_LAYOUT_CORNER  one record per trilinear cell [nr-1][nt-1][np] holding the
 *                    cell's 8 corners x 11 channels contiguously (phi seam cell
 *                    included; corner-major, channel-major for BH_STORE_BF16):
 *                    8x the bytes, one contiguous read per sample. */
#define BH_LAYOUT_SOA     0
#define BH_LAYOUT_CORNER  1

/* Upload an AoS float32 brick (nr, nt, np, 11) with float32 grids, stored on
 * the device in the requested type and layout.  The FP64 render needs
 * BH_STORE_F32 + BH_LAYOUT_SOA.  nr + nt + np <= 2048; a SoA brick stays below
 * 2 GiB, a corner-packed one has no limit beyond device memory. */
int bh_brick_upload(bh_ctx *ctx, const float *coeffs_aos,
                    const float *r_grid, int32_t nr,
                    const float *theta_grid, int32_t nt,
                    const float *phi_grid, int32_t np,
                    int32_t storage, int32_t layout, int32_t *brick_id);
int bh_brick_release(bh_ctx *ctx, int32_t brick_id);

/* ---- fused render (render_stokes_image_bricks, accelerated_renderer.py:100-158,
 *      and the GPU asset's kerr_stokes_render_kernel, native/rocm/kerr_stokes_kernel.hip:20) */
typedef struct bh_scene_desc {
    double spin;              /* a, |a| < 1 */
    double cam_r;             /* LocalCamera (kerr.py:155-185) */
    double cam_theta;
    double cam_phi;
    double fov_y_degrees;
    double roll_degrees;
    int32_t width;
    int32_t height;
    int32_t step_mode;        /* 0: fixed affine step; 1: h = step * (r - r_+) */
    int32_t max_steps;
    double step;
    double escape_radius;     /* AcceleratedRenderConfig.escape_radius */
    double affine_to_length_cm;
    double max_optical_depth; /* PolarizedTransferConfig.max_optical_depth */
    int32_t transfer;         /* BH_TRANSFER_RK2 or BH_TRANSFER_EXACT
                                 (PolarizedTransferConfig.use_exact_matrix_step,
                                 polarized_transfer.py:51-80) */
    int32_t integrator;       /* BH_INTEGRATOR_RK2 or BH_INTEGRATOR_RK4 */
    int32_t sampling;         /* BH_SAMPLING_NATIVE or BH_SAMPLING_NUMPY32 */
    int32_t reserved;         /* 0 */
} bh_scene_desc;

#define BH_TRANSFER_RK2    0   /* one RK2 step of dS/ds = j - K S per segment (stokes_step_rk2) */
#define BH_TRANSFER_EXACT  1   /* exact constant-coefficient step: exp of [[-K, j], [0, 0]] ds (stokes_step_exact) */

/* Geodesic step: RK2 midpoint (the GPU asset's rk2_geodesic_step, native/rocm/kerr_stokes_kernel.hip:13),
 * or classical RK4 as the reference tracer (rk4_step, kerr.py:306-313, called by
 * trace_kerr_null_geodesic from render_tile_stokes_bricks, accelerated_renderer.py:142). */
#define BH_INTEGRATOR_RK2  0
#define BH_INTEGRATOR_RK4  1
/* Brick sampling of the FP64 render: NATIVE brackets and interpolates in FP64 like the drop-in
 * sampler (native_kernels.py:164-186, lib.rs:171-208); NUMPY32 reproduces the reference renderer's
 * sample_coefficient_brick (accelerated_renderer.py:73-97) on a float32 brick as NumPy evaluates it,
 * with coordinates, brackets, weights and lerps in float32 arithmetic.  FP32 and mixed renders always
 * sample in their own FP32 form. */
#define BH_SAMPLING_NATIVE   0
#define BH_SAMPLING_NUMPY32  1

#define BH_PREC_FP64   0   /* full FP64 render (reference-grade) */
#define BH_PREC_FP32   1   /* plain FP32 geodesic + transfer */
#define BH_PREC_MIXED  2   /* FP32 with FP64 islands; BF16/FP16 bricks */

/* Per-pixel status byte: low 3 bits 0 max_steps, 1 captured, 2 escaped,
 * 3 invalid, 4 optically thick; bit 3 set when the steps could not follow
 * the geodesic through a pole (theta clamp fired, or |d phi| > 1 rad in one step). */

/* Launch shape of the persistent ray-queue kernels (0 fields = defaults). */
typedef struct bh_tuning {
    int32_t block_threads;    /* multiple of 64 */
    int32_t blocks_per_cu;    /* persistent grid = CUs * blocks_per_cu (0: occupancy) */
    int32_t steps_per_sync;   /* steps between wave refill checks */
    int32_t refill_min;       /* refill when at least this many lanes are idle (1..64) */
    int32_t queues;           /* ray queues, 1..8; workgroup b serves queue b % queues first */
    int32_t order;            /* 0: row-major tiles; 1: tiles nearest the critical curve first */
    int32_t coop_steps;       /* render kernels: a ray alone in its wave for this many steps asks an idle
                                 wave of its workgroup to take its geodesic and brick samples (0: never;
                                 -1 in set: keep) */
} bh_tuning;

int bh_ctx_set_tuning(bh_ctx *ctx, int32_t precision, const bh_tuning *t);

/* Work done by the last render launch on this context (waits for it): total RK2
 * steps, transfer segments that sampled inside the brick, and lane slots
 * (wave loop iterations x 64), so steps / lane_slots is the SIMD lane utilisation
 * after ray termination and compaction.  Any pointer may be NULL. */
int bh_ctx_last_work(bh_ctx *ctx, uint64_t *steps, uint64_t *samples, uint64_t *lane_slots);
int bh_ctx_get_tuning(bh_ctx *ctx, int32_t precision, bh_tuning *t);
/* Lone rays of the last render whose geodesic an idle wave took over, and hand-overs that timed out
 * (the lone wave then finished the ray alone). */
int bh_ctx_last_coop(bh_ctx *ctx, uint64_t *split_rays, uint64_t *timeouts);

/* Host outputs: iquv is H*W*4 doubles for BH_PREC_FP64, floats otherwise;
 * steps H*W int32; status H*W bytes (either may be NULL). */
int bh_render(bh_ctx *ctx, const bh_scene_desc *scene, int32_t brick_id,
              int32_t precision, void *iquv, int32_t *steps, uint8_t *status);
/* Device outputs, enqueued on the context stream (bh_ctx_synchronize to wait). */
int bh_render_device(bh_ctx *ctx, const bh_scene_desc *scene, int32_t brick_id,
                     int32_t precision, void *d_iquv, int32_t *d_steps, uint8_t *d_status);

/* No counterpart in the reference: bh_render_device for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  Same kernels and results as
 * bh_render_device, but the launch's ray-queue counters and its tile schedule
 * live in a static device array of the library (BH_NOALLOC_SCRATCH_BYTES per
 * GPU, held until the process ends) instead of memory the context allocates,
 * so the call makes no hipMalloc or hipFree.  Images up to BH_NOALLOC_MAX_TILES
 * 8x8 tiles (2048 x 2048 pixels); larger ones give BH_ERR_VALUE.  Every call on
 * a GPU shares that GPU's array (it keeps the last call's counters, which
 * bh_ctx_last_work reads): calls that overlap on one GPU, from two contexts or
 * threads, race.  The call waits for the upload of its tile schedule. */
#define BH_NOALLOC_MAX_TILES     65536
#define BH_NOALLOC_SCRATCH_BYTES (4 * (64 + 65536 + 8))
int bh_render_device_noalloc(bh_ctx *ctx, const bh_scene_desc *scene, int32_t brick_id, int32_t precision,
                             void *d_iquv, int32_t *d_steps, uint8_t *d_status);

/* ---- batched persistent launch --------------------------------------------
 * Many renders (own image size, spin, camera and brick) traced by one persistent
 * kernel launch.  All bricks of a batch must share storage and layout.  Render i
 * writes pixels [offsets[i], offsets[i+1]) of the batch outputs (offsets has n+1
 * entries, filled by bh_batch_create).  bh_batch_launch_graph captures the launch
 * (counter reset + kernel) in a hipGraph on first use and replays it afterwards. */
typedef struct bh_batch bh_batch;
typedef struct bh_batch_item {
    bh_scene_desc scene;
    int32_t brick_id;
    int32_t reserved;
} bh_batch_item;

int bh_batch_create(bh_ctx *ctx, const bh_batch_item *items, int32_t n, int32_t precision,
                    int64_t *offsets, bh_batch **out);
int bh_batch_launch(bh_batch *batch, void *d_iquv, int32_t *d_steps, uint8_t *d_status);
int bh_batch_launch_graph(bh_batch *batch, void *d_iquv, int32_t *d_steps, uint8_t *d_status);
int bh_batch_destroy(bh_batch *batch);

/* ---- learned surrogate: MLP on BF16 MFMA ------------------------------------
 * No counterpart in the reference (its renderers have no learned part); this is
 * the surrogate of the Kerr ray tracer.  x (n_in) -> [Linear width, SiLU] x depth
 * -> Linear n_out.  FP32 master parameters, BF16 copies feed the MFMA GEMMs, FP32
 * accumulation.  n_in and n_out are padded to multiples of 8 inside; the
 * parameter vector (get/set, gradients) uses the padded layout: per layer l,
 * W_l[fi][fo] row-major then b_l[fo], with fi/fo the padded fan-in/fan-out.
 * Padded rows and columns must be zero.  Device pointers are FP32. */
typedef struct bh_mlp bh_mlp;
int bh_mlp_create(bh_ctx *ctx, int32_t n_in, int32_t width, int32_t depth, int32_t n_out, bh_mlp **out);
int bh_mlp_destroy(bh_mlp *mlp);
int bh_mlp_shape(bh_mlp *mlp, int32_t *n_in_pad, int32_t *n_out_pad, int64_t *n_params);
int bh_mlp_set_params(bh_mlp *mlp, const float *host_params);   /* also resets Adam state */
int bh_mlp_get_params(bh_mlp *mlp, float *host_params);
/* Adam's first and second moments (n_params floats each, host) and its step count, for checkpoints:
 * bh_mlp_set_adam after bh_mlp_set_params resumes training where bh_mlp_get_adam left it. */
int bh_mlp_get_adam(bh_mlp *mlp, float *host_m, float *host_v, int64_t *steps);
int bh_mlp_set_adam(bh_mlp *mlp, const float *host_m, const float *host_v, int64_t steps);
/* Inference: d_x [rows][n_in] -> d_y [rows][n_out], or [n_out][rows] when transposed.
 * Transposed inference of width-256 networks (n_in <= 8, n_out <= 1024) runs one fused kernel
 * (activations kept in LDS) unless bh_mlp_set_fused(mlp, 0) selects the per-layer GEMMs. */
int bh_mlp_forward(bh_mlp *mlp, const float *d_x, int64_t rows, float *d_y, int32_t transposed);
int bh_mlp_set_fused(bh_mlp *mlp, int32_t fused);
/* Loss = sum_{i,o} w_io (y_io - t_io)^2 / rows with d_t [rows][n_out]; w_io =
 * d_w[o] * d_row_w[i] for o < masked_cols and d_w[o] otherwise (d_row_w may be
 * NULL: all ones).  bh_mlp_gradient returns dLoss/dparams (n_params floats, may
 * be NULL) without an update; bh_mlp_train_step takes one Adam step (beta1 0.9,
 * beta2 0.999, eps 1e-8). */
int bh_mlp_gradient(bh_mlp *mlp, const float *d_x, const float *d_t, const float *d_w, const float *d_row_w,
                    int32_t masked_cols, int32_t rows, float *grad_host, float *loss_host);
int bh_mlp_train_step(bh_mlp *mlp, const float *d_x, const float *d_t, const float *d_w, const float *d_row_w,
                      int32_t masked_cols, int32_t rows, float lr, float *loss_host);
/* d[i][c] = d[i][c] * mul[c] + add[c] for a device [rows][cols] FP32 matrix (host mul/add). */
int bh_columns_affine(bh_ctx *ctx, float *d, int64_t rows, int32_t cols, const float *mul, const float *add);

/* ---- geodesic knots (training data of the surrogate) -----------------------
 * For each ray (alpha, beta, a, cos i) — image-plane offsets in M of a camera at
 * radius cam_r, inclination acos(cos i), phi 0, no roll: the direction is
 * (-1, -beta / cam_r, alpha / cam_r) in the ZAMO frame, as a pixel whose NDC
 * gives ndc_x * tan_x = alpha / cam_r — the FP64 RK2 geodesic of bh_render is
 * traced and the states from the start of the first step whose midpoint radius
 * lies in [r_lo, r_hi] are recorded as n_knots knots (r, theta, phi, g), where
 * g = E alpha_lapse / (E - omega L) is the redshift seen by a ZAMO; after the
 * ray ends (or a pole clamp fires) the last state repeats.  Output rows
 * (n_knots * 4 + 3 floats): knots, then captured (0/1), flag (1 when the brick
 * part needs more than n_knots states, a pole clamp fired, the ray failed or hit
 * max_steps) and the number of recorded states; d_steps (optional) gets the
 * total RK2 step count and d_weight (optional) 1 for unflagged rays, else 0. */
typedef struct bh_knot_desc {
    double cam_r;
    double step;             /* RK2 step, or factor of (r - r_+) when step_mode = 1 */
    double escape_radius;
    double r_lo, r_hi;       /* brick radial range */
    int32_t step_mode;
    int32_t max_steps;
    int32_t n_knots;         /* 2..512 */
    int32_t reserved;
} bh_knot_desc;
int bh_geodesic_knots(bh_ctx *ctx, const bh_knot_desc *desc, const float *d_params, int64_t n,
                      float *d_rows, int32_t *d_steps, float *d_weight);

/* ---- surrogate render --------------------------------------------------------
 * One render of `scene` from MLP outputs instead of RK2 geodesics.  The network
 * input of pixel (px, py) is x = ((alpha, beta, a, cos i) - in_add) * in_mul with
 * alpha = cam_r * ndc_x * tan_x, beta = cam_r * ndc_y * tan_y; its outputs,
 * denormalised as y * out_mul + out_add, are rows as written by
 * bh_geodesic_knots.  The Stokes transfer runs over the predicted knots with the
 * midpoint / Euclidean-ds rule of the RK2 transfer.  Pixels whose Bardeen image
 * coordinates lie within crit_distance of the critical curve, whose flag output
 * exceeds flag_max, or whose captured output lies in (cap_lo, cap_hi) are traced
 * by the FP32 RK2 kernel instead (bh_render, BH_PREC_FP32 or BH_PREC_MIXED per
 * fallback_precision).  Status bit 4 marks surrogate pixels; their step count is
 * the number of knot segments integrated.  d_work (device, bh_surrogate_work_bytes)
 * holds the network input and output and the fallback list.  The number of
 * fallback rays of the last call is returned by bh_surrogate_last_fallback
 * (waits for the call). */
typedef struct bh_surrogate_desc {
    float in_mul[4], in_add[4];
    const float *out_mul, *out_add;   /* host, n_knots * 4 + 3 values */
    int32_t n_knots;
    int32_t fallback_precision;
    float crit_distance;
    float flag_max, cap_lo, cap_hi;
} bh_surrogate_desc;
int bh_surrogate_render_device(bh_ctx *ctx, const bh_scene_desc *scene, int32_t brick_id, bh_mlp *mlp,
                               const bh_surrogate_desc *sd, float *d_work, size_t work_bytes,
                               void *d_iquv, int32_t *d_steps, uint8_t *d_status);
/* The same render from knot rows the caller supplies ([n_out][ld] FP32, ld >=
 * width * height, denormalised by out_mul / out_add) instead of the network:
 * with the rows of bh_geodesic_knots it reproduces the RK2 transfer of every
 * unflagged ray whose brick part fits in n_knots states. */
int bh_knot_render_device(bh_ctx *ctx, const bh_scene_desc *scene, int32_t brick_id, const float *d_rows,
                          int64_t ld, const bh_surrogate_desc *sd, float *d_work, size_t work_bytes,
                          void *d_iquv, int32_t *d_steps, uint8_t *d_status);
int bh_surrogate_last_fallback(bh_ctx *ctx, int64_t *n_fallback);
/* Bytes of d_work bh_surrogate_render_device needs for `rows` pixels. */
int bh_surrogate_work_bytes(int64_t rows, int32_t n_knots, size_t *bytes);

/* ---- batched surrogate renders -----------------------------------------------
 * No counterpart in the reference (its renderers have no learned part).  Many renders
 * (own image size, spin, camera, brick; RK2 transfer and integrator) through the
 * surrogate in one sequence: the network inputs of every render's rays, one fused
 * inference over all of them with FP16 knots, the transfer over knots of every
 * render's tiles in one launch, and every render's fallback rays (the rule of
 * bh_surrogate_render_device) pooled into one persistent batched RK2 launch at
 * sd->fallback_precision.  Critical curves are computed at creation, so a launch
 * needs no host synchronisation; with graph != 0 the sequence is captured in a
 * hipGraph on first use and replayed afterwards.  Outputs are laid out as in
 * bh_batch (offsets, n + 1 entries); status bit 4 marks surrogate pixels.  All
 * bricks must share storage and layout; the network must be the fused width-256
 * one (4 inputs, 4 * n_knots + 3 outputs). */
typedef struct bh_surrogate_batch bh_surrogate_batch;
int bh_surrogate_batch_create(bh_ctx *ctx, const bh_batch_item *items, int32_t n, bh_mlp *mlp,
                              const bh_surrogate_desc *sd, int64_t *offsets, bh_surrogate_batch **out);
int bh_surrogate_batch_launch(bh_surrogate_batch *batch, void *d_iquv, int32_t *d_steps, uint8_t *d_status,
                              int32_t graph);
/* Fallback rays of the last launch, per render (n values, may be NULL) and in all (waits for the launch). */
int bh_surrogate_batch_fallback(bh_surrogate_batch *batch, int64_t *per_render, int64_t *total);
int bh_surrogate_batch_destroy(bh_surrogate_batch *batch);

/* ---- coefficient-brick precompute (precompute_coefficient_bricks,
 * coefficient_bricks.py:118-152, with HybridSynchrotronCoefficients,
 * synchrotron.py:231-325) -------------------------------------------------------
 * Cell frames as coefficient_bricks.py:110-116: n_e = rho rho_cgs_per_code /
 * max(mu_e m_p, 1e-300), B = |b_con[1:4]| b_gauss_per_code, EVPA =
 * atan2(b_con[3], b_con[2]), theta_e floored at 1e-8, pitch angle to the line of
 * sight 90 deg.  Thermal (Maxwell-Juttner) plus power-law emission over the
 * reference's Lorentz-factor grids and Gauss-Legendre pitch nodes, Kirchhoff /
 * power-law absorption, Faraday terms, EVPA rotation; then the cast to the brick
 * dtype and nan_to_num (NaN -> 0, +inf -> the dtype's max, -inf -> 0). */
#define BH_STORE_F64   3   /* float64 brick values (precompute output only) */

typedef struct bh_synchrotron_desc {
    double nu_hz;                                  /* observing frequency */
    double rho_cgs_per_code, b_gauss_per_code;     /* PhysicalScaling (calibration.py:28-56) */
    double mean_molecular_weight_per_electron;
    /* ThermalSynchrotronCoefficients (synchrotron.py:231-238) */
    int32_t th_gamma_grid_size, th_pitch_nodes;    /* 192, 16 */
    double th_faraday_rotation_scale, th_faraday_conversion_scale;   /* 1, 1 */
    /* NonthermalPowerLawSynchrotronCoefficients (synchrotron.py:271-282) */
    double nt_p, nt_gamma_min, nt_gamma_max, nt_fraction;            /* 3.5, 10, 1e5, 0.02 */
    int32_t nt_gamma_grid_size, nt_pitch_nodes;                      /* 224, 16 */
    double nt_absorption_scale, nt_faraday_scale;                    /* 1, 0.15 */
} bh_synchrotron_desc;

/* Replaces precompute_coefficient_bricks (coefficient_bricks.py:118-152) over
 * n_cells cells given as flat field arrays (the caller takes the stride):
 * rho[n], theta_e[n], b_con[n][4] -> out[n][11] in out_dtype (BH_STORE_F32,
 * BH_STORE_F16 or BH_STORE_F64), channels in COEFF_NAMES order.  Grid sizes
 * 2..1024, pitch nodes 1..16.  Host buffers, synchronous. */
int bh_precompute_coefficient_bricks(const double *rho, const double *theta_e, const double *b_con,
                                     int64_t n_cells, const bh_synchrotron_desc *desc, int32_t out_dtype,
                                     void *out);
/* The same with device buffers on the context stream (kernel time in bh_ctx_last_kernel_ms). */
int bh_precompute_coefficient_bricks_device(bh_ctx *ctx, const double *d_rho, const double *d_theta_e,
                                            const double *d_b_con, int64_t n_cells,
                                            const bh_synchrotron_desc *desc, int32_t out_dtype, void *d_out);
/* K_nu(x), the modified Bessel function of the second kind, as scipy.special.kv
 * (synchrotron.py:148, 195), for finite nu >= 0 and finite x > 0 (trapezoidal
 * rule on K_nu(x) = int_0^inf exp(-x cosh t) cosh(nu t) dt).  Host only. */
int bh_bessel_k(double nu, double x, double *out);
/* Gauss-Legendre nodes (ascending) and weights on [-1, 1], n in [1, 64], as
 * numpy.polynomial.legendre.leggauss (synchrotron.py:219).  Host only. */
int bh_gauss_legendre(int32_t n, double *x, double *w);

#ifdef __cplusplus
}
#endif

#endif /* BHGRRT_H */
