/*
 * pmarch.h -- C ABI of the MI355X particle-fluid ray marcher (libpmarch.so).
 *
 * Replaces the reference's CPU marcher behind the same launch/poll protocol:
 *   class RayMarcher            src/app/AdvancedRenderer/RayMarcher.h:28-77
 *     Prepare(...)              src/app/AdvancedRenderer/RayMarcher.cpp:76-100
 *     Start()                   src/app/AdvancedRenderer/RayMarcher.cpp:102-112
 *     IsDone()                  src/app/AdvancedRenderer/RayMarcher.h:51 -> ThreadPool.cpp:33-36
 *     Exit()                    src/app/AdvancedRenderer/RayMarcher.cpp:71-74
 *   struct VisualizationSettings src/app/AdvancedRenderer/RayMarcher.h:12-26
 *   Frame::Frame / BuildSearch / ComputeAABB / BuildDensityGrid
 *                               src/app/Dataset.cpp:9-24, 49-164
 *   Dataset::GetNeighbors       src/app/Dataset.cpp:272-280 (CompactNSearch::find_neighbors)
 *
 * Conventions: every call returns an int status (PM_OK = 0, PM_PENDING = 1,
 * negative = error, message in pm_last_error); no exception crosses the ABI.
 * Outputs of pm_march_async are valid only after pm_query returns PM_OK or
 * pm_wait returns (this replaces the reference's racy IsDone counter poll).
 * One host thread per context.  Images are row-major, index = y * W + x;
 * positions and normals are RGBA32F (16 B per pixel, as BilateralBuffer::Color,
 * BilateralBuffer.cpp:18); a miss writes (0,0,0,0) to both.
 */
#ifndef PMARCH_H
#define PMARCH_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PM_OK          0
#define PM_PENDING     1
#define PM_ERR_ARG    (-1)
#define PM_ERR_HIP    (-2)
#define PM_ERR_FRAME  (-3)
#define PM_ERR_SIZE   (-4)
#define PM_ERR_NODEV  (-5)
#define PM_ERR_STATE  (-6)
#define PM_ERR_IO     (-7)

#define PM_MODE_CONFIG     0  /* AABB entry, fixed step, bisection, gradient normals   */
#define PM_MODE_REFERENCE  1  /* PerPixel_Isotropic, RayMarcher.cpp:256-344            */

#define PM_NORMAL_CDIFF    0  /* normalize(-grad rho) by central differences           */
#define PM_NORMAL_GRADW    1  /* normalize(sum gradW), Kernel.cpp:34-52 (with 1/|r|)   */

#define PM_SKIP_NONE       0
#define PM_SKIP_EMPTY      1  /* configured mode: skip samples whose 27 cells are empty */
#define PM_SKIP_DG_BOX     2  /* reference mode: DensityGrid, box flag                 */
#define PM_SKIP_DG_CONS    3  /* reference mode: DensityGrid, conservative flag        */
#define PM_SKIP_BOUND      4  /* configured mode: skip 8x8x8 sub-boxes of a cell whose
                                 density bound (sum of W at each particle's distance to
                                 the sub-box) stays below iso, a 4x4x4 box at once when
                                 all eight of its sub-boxes do; the image is that of the
                                 march without the skip.  With the anisotropic density:
                                 sub-boxes with no particle within h (the bound for
                                 iso PM_ZERO_ISO), where that density is exactly 0     */
#define PM_ZERO_ISO  7.88860905e-31f  /* 2^-100: a bound below it means no particle within h */

/* March settings.  The first four fields carry VisualizationSettings'
 * MaxSteps / StepSize / IsoDensity (RayMarcher.h:16-18), the anisotropic
 * fields its EnableAnisotropy / k_n / k_r / k_s / N_eps (RayMarcher.h:20-25);
 * the rest select the configured march.  Layout is fixed: 24 x 4 bytes. */
typedef struct pm_settings {
    int32_t mode;          /* PM_MODE_*                                              */
    int32_t max_steps;     /* reference: MaxSteps; configured: cap on march samples, 0 = none */
    float   step;          /* step length in world units (> 0)                        */
    float   iso;           /* iso density in kernel units                             */
    int32_t bisect_iters;  /* configured: bisection iterations on the crossing bracket */
    int32_t final_fp64;    /* configured: last bisection iteration evaluated in FP64  */
    float   normal_eps;    /* configured: central-difference offset (world units)     */
    int32_t normal_mode;   /* PM_NORMAL_*                                             */
    int32_t skip_mode;     /* PM_SKIP_*                                               */
    int32_t compensated;   /* 1 = Neumaier-compensated FP32 density sums              */
    int32_t learned;       /* 1 = learned step proposal (needs pm_mlp_set)            */
    int32_t schedule;      /* kernel schedule, same images and counters: 0 = over the
                              8x8 tiles a ray of which enters the AABB (reference
                              mode: with a drawn depth), the others answered as misses
                              at once -- the isotropic configured march with persistent
                              waves, ray refill, lane groups and per-wave finishing,
                              the anisotropic and reference marches one tile per wave
                              (the product); 1 = every tile, one ray per lane to its
                              end, direct sums (the anisotropic configured march:
                              persistent waves with ray refill) -- the schedules the
                              product is checked against                           */
    float   learned_alpha; /* learned: fraction of the predicted distance skipped (0 = 0.7) */
    int32_t learned_min_skip; /* learned: shorter proposals are not taken (0 = 3, >= 2) */
    int32_t learned_max_skip; /* learned: longest jump in steps (0 = 8)               */
    /* PerPixel_Anisotropic (RayMarcher.cpp:346-423): per sample, WPCA over the
       particles within ext_mult * h (RayMarcher.cpp:114-254), then the
       anisotropic kernel (Kernel.cpp:63-107) over those within h.  Either
       mode; not with the learned step or the batch.                          */
    int32_t anisotropic;   /* EnableAnisotropy                                        */
    float   k_n;           /* k_n: Sigma for N <= N_eps (> 0)                         */
    float   k_r;           /* k_r: eigenvalues clamped to max / k_r (> 0)             */
    float   k_s;           /* k_s: scale of the clamped eigenvalues (> 0)             */
    int32_t n_eps;         /* N_eps                                                   */
    float   ext_mult;      /* ParticleRadiusExt / ParticleRadius (Dataset.cpp:169-177),
                              in [1, 2] (the extended walk covers 2h); 0 = 2
                              (config.yml particleRadiusMultiplier)                   */
    int32_t instrument;    /* configured isotropic march, schedule 0: 1 = also count,
                              per density evaluation, the candidates of its own fine
                              stencil and of its 27 search cells (pm_stats.cand_stencil,
                              cand27; same images, slower)                          */
    float   learned_gate;  /* learned: a sample asks for a proposal only when its
                              density is below learned_gate x iso (0 = 0.25)       */
    int32_t bound_boxes;   /* PM_SKIP_BOUND: 0 = sub-boxes of 8x8x8 per cell (the
                              default; fewer march samples), 1 = 4x4x4 only (a build
                              about 5x cheaper, for a frame marched once, as in
                              streaming); the images are the same                   */
} pm_settings;

/* Per-launch counters, written by the march kernels themselves. */
typedef struct pm_stats {
    uint64_t rays;          /* primary rays (W*H per view)                                */
    uint64_t rays_in_aabb;  /* rays that enter the padded particle AABB                   */
    uint64_t hits;          /* rays with an iso crossing                                  */
    uint64_t samples;       /* density evaluations: march samples, bisection samples (FP32
                               and FP64) and central-difference samples                   */
    uint64_t cand;          /* candidates read by the launch's FP32 walks, each walk once
                               (a hit's six central differences share one walk)          */
    uint64_t march_samples; /* density evaluations of the march phase                     */
    uint64_t mlp_evals;     /* learned step: network evaluations                          */
    uint64_t march_cand;    /* candidates of the march samples' own fine stencils         */
    uint64_t cert_walks;    /* learned step: segment certificates walked                  */
    uint64_t cert_skipped;  /* learned step: lattice samples certified and not evaluated  */
    uint64_t cand_stencil;  /* instrumented launches: sum over every density evaluation of
                               the candidates of its own fine stencil                      */
    uint64_t cand27;        /* instrumented launches: sum over every density evaluation of
                               the particles in its 27 search cells (cells of h)           */
    uint64_t reserved;
    float    kernel_ms;     /* march kernels' time, HIP events on the context stream      */
    float    total_ms;      /* launch to completion incl. copies                          */
} pm_stats;

/* Frame description after pm_set_frame (the GPU grid of Frame::BuildSearch). */
typedef struct pm_frame_info {
    uint64_t n;             /* particles                                     */
    uint64_t ncells;        /* dense cells                                   */
    int32_t  origin[3];     /* cell key of dense cell (0,0,0)                */
    int32_t  dims[3];       /* dense grid dimensions                         */
    float    h;             /* support = cell width (ParticleRadius)         */
    float    pmin[3], pmax[3];  /* particle bounds                           */
    float    bmin[3], bmax[3];  /* padded AABB (Frame::ComputeAABB, Dataset.cpp:78-92) */
    float    build_ms;      /* grid build time on the GPU                    */
    /* march grid: cells of h/2, same key rule with 2/h, 4 cells of margin  */
    int32_t  forigin[3];
    int32_t  fdims[3];
    uint64_t fncells;
    float    bound_ms;      /* GPU time of the last build of the sub-box bound bits
                               (PM_SKIP_BOUND), -1 when none was built            */
    float    bound_iso;     /* the iso those bits are for                         */
} pm_frame_info;

typedef struct pm_ctx pm_ctx;

/* RayMarcher::RayMarcher (RayMarcher.cpp:66-69): opens device `device`. */
int  pm_create(int device, pm_ctx **out);
/* RayMarcher::Exit (RayMarcher.cpp:71-74). */
void pm_destroy(pm_ctx *ctx);
const char *pm_last_error(const pm_ctx *ctx);

/* Frame::Frame (Dataset.cpp:9-24): copies n float3 positions (the caller keeps
 * ownership) and builds the uniform grid, AABB and skip structures on the GPU. */
int pm_set_frame(pm_ctx *ctx, int32_t frame, const float *xyz, uint64_t n, float h);
/* Same, from positions already in device memory (frame streaming). */
int pm_set_frame_device(pm_ctx *ctx, int32_t frame, const void *d_xyz, uint64_t n, float h);
/* Frame streaming (the reference's autoplay re-renders the next frame each
 * second, AdvancedRenderer.cpp:283-290): pm_set_frame on a second stream,
 * allowed while a march of another frame is in flight, so the next frame's
 * upload and grid build run beside it; returns once the frame is built.
 * PM_ERR_STATE when the march in flight reads this frame (or is a batch). */
int pm_set_frame_stream(pm_ctx *ctx, int32_t frame, const float *xyz, uint64_t n, float h);
int pm_frame_info_get(pm_ctx *ctx, int32_t frame, pm_frame_info *out);

/* ---- dataset IO: Dataset::Dataset / ReadFile (Dataset.cpp:169-227, 292-306)
 * through Partio::readBGEO (vendor/partio/src/io/BGEO.cpp:187-300) ----
 * Positions (the first three floats of P) of a classic Houdini BGEO file,
 * version 5, plain or gzip-compressed.  With xyz == NULL only *n is set;
 * otherwise the n points are written (PM_ERR_SIZE when n > cap).  PM_ERR_IO
 * for an unreadable file, another format or version, a string attribute (as
 * Partio refuses it) or a truncated file.  Host code: needs no context or GPU. */
int pm_bgeo_read(const char *path, float *xyz, uint64_t cap, uint64_t *n);
/* One frame from a BGEO file: pm_bgeo_read, then pm_set_frame (or
 * pm_set_frame_stream while a march is in flight). */
int pm_load_bgeo(pm_ctx *ctx, int32_t frame, const char *path, float h);
/* The frame's grid: cell_start[ncells+1], sorted[4n] (x,y,z,0), index[n]
 * (input index of each stored particle).  Any pointer may be NULL. */
int pm_get_grid(pm_ctx *ctx, int32_t frame, uint32_t *cell_start, float *sorted, uint32_t *index);
/* The frame's march grid (cells of h/2): cell_start[fncells+1], sorted[4n]
 * (x,y,z,0) fine-cell-major, by input index inside a cell.  Either may be NULL. */
int pm_get_march_grid(pm_ctx *ctx, int32_t frame, uint32_t *cell_start, float *sorted);
/* The frame's skip structures: empty27[ncells] (1 where a cell and its 26
 * neighbours are empty) and the reference DensityGrid (Frame::BuildDensityGrid,
 * Dataset.cpp:94-164) over the padded AABB: dims, box flag, conservative flag
 * and per-node box counts.  Any pointer may be NULL; call with NULL arrays first
 * to read dg_dims. */
int pm_get_skip_grids(pm_ctx *ctx, int32_t frame, uint8_t *empty27, int32_t dg_dims[3], uint8_t *flag_box,
                      uint8_t *flag_cons, uint32_t *count_box);
/* The frame's sub-box bound bits for iso and bound_boxes `boxes` (PM_SKIP_BOUND):
 * bits[8 ncells], 512 bits per cell.  Sub-box (fx, fy, fz), f in 0..7 per axis,
 * is bit 8 l + (fx & 1) + 2 (fy & 1) + 4 (fz & 1) of the cell's words with
 * l = fx/2 + 4 (fy/2) + 16 (fz/2) (bit b in word b / 64); a set bit means the
 * density bound proves the sub-box below iso, a clear bit proves nothing
 * (boxes = 1: each byte 0xff or 0, the 4x4x4 boxes' bounds alone).  Builds
 * them if the frame has none for this iso and boxes. */
int pm_get_bound_grid(pm_ctx *ctx, int32_t frame, float iso, int32_t boxes, uint64_t *bits);

/* RayMarcher::Prepare + Start (RayMarcher.cpp:76-112).  inv_pv is
 * Camera3D::GetInvProjectionView (column-major 16 floats), cam is
 * CameraController3D::Position.  depth_in (W*H floats) is required in
 * PM_MODE_REFERENCE and ignored otherwise.  Output pointers are host memory;
 * any may be NULL, and with all three NULL the results stay on the device
 * (pm_download).  Returns immediately. */
int pm_march_async(pm_ctx *ctx, int32_t frame, const pm_settings *s, const float inv_pv[16],
                   const float cam[3], uint32_t W, uint32_t H, const float *depth_in,
                   float *positions_out, float *normals_out, float *t_out);
/* The same march with every image in device memory the caller owns -- e.g.
 * buffers the renderer imported from its Vulkan images (the RGBA32F
 * positions/normals of BilateralBuffer.cpp:14-24): the kernels write them
 * directly, no host copy.  d_positions and d_normals hold W*H float4 each,
 * d_t W*H floats or NULL; d_depth (reference mode) is device memory too.  They
 * must stay valid until pm_query returns PM_OK; pm_download then refuses. */
int pm_march_async_device(pm_ctx *ctx, int32_t frame, const pm_settings *s, const float inv_pv[16],
                          const float cam[3], uint32_t W, uint32_t H, const float *d_depth, void *d_positions,
                          void *d_normals, void *d_t);
/* RayMarcher::IsDone: PM_OK when the march and its copies are complete,
 * PM_PENDING while running, negative on error. */
int pm_query(pm_ctx *ctx);
/* Blocks until the pending march is complete. */
int pm_wait(pm_ctx *ctx);
/* Copies the last march's device results to host memory (any pointer NULL). */
int pm_download(pm_ctx *ctx, float *positions, float *normals, float *t);
int pm_get_stats(pm_ctx *ctx, pm_stats *out);

/* ---- batched views x frames in one persistent launch ----
 * The reference renders one view per RayMarcher::Start, farmed out pixel by
 * pixel to its ThreadPool (ThreadPool.cpp:5-55) and re-started once per
 * second (AdvancedRenderer.cpp:257-298); the batch form marches many views of
 * uneven size over many resident frames in one persistent kernel.  Waves take
 * runs of 8x8 tiles of one view from a single queue ordered largest view
 * first; every view's images equal those of pm_march_async for the same
 * frame, settings and camera. */
typedef struct pm_view {
    int32_t frame;      /* a frame set by pm_set_frame */
    uint32_t width, height;
    float inv_pv[16];   /* as in pm_march_async */
    float cam[3];
} pm_view;
/* Fixes the batch (configured march only): allocates the device images of
 * all views, uploads the tile queue and, with use_graph != 0, captures the
 * launch (counter fills + the persistent kernel) in a hipGraph. */
int pm_batch_set(pm_ctx *ctx, const pm_settings *s, const pm_view *views, uint32_t n_views, int use_graph);
/* Enqueues one run of the batch (graph replay or direct launch) and returns;
 * complete it with pm_query / pm_wait, which also fill pm_get_stats. */
int pm_batch_run(pm_ctx *ctx);
/* Copies one view's device images to host memory (any pointer NULL). */
int pm_batch_download(pm_ctx *ctx, uint32_t view, float *positions, float *normals, float *t);

/* Dataset::GetNeighbors (Dataset.cpp:272-280) as counts: for each query point,
 * the number of particles with |q - x_j|^2 < h^2 and of stencil candidates. */
int pm_neighbor_counts(pm_ctx *ctx, int32_t frame, const float *q, uint64_t nq,
                       uint32_t *counts, uint32_t *ncand);
/* Density sum_j W(x_j - q) (RayMarcher.cpp:321-325) at query points, FP32
 * (compensated or plain) and FP64. */
int pm_density(pm_ctx *ctx, int32_t frame, const float *q, uint64_t nq, int compensated,
               float *rho);
int pm_density64(pm_ctx *ctx, int32_t frame, const double *q, uint64_t nq, double *rho);
/* The anisotropic density at query points (RayMarcher.cpp:378-400) with the
 * anisotropic fields (and compensated) of s: FP32 and FP64. */
int pm_density_aniso(pm_ctx *ctx, int32_t frame, const pm_settings *s, const float *q, uint64_t nq, float *rho);
int pm_density_aniso64(pm_ctx *ctx, int32_t frame, const pm_settings *s, const double *q, uint64_t nq, double *rho);
/* WPCA at query points (RayMarcher.cpp:114-254): G (9 floats per point, row
 * major), det G and the number of extended neighbours (0: G and det are 0). */
int pm_aniso_matrix(pm_ctx *ctx, int32_t frame, const pm_settings *s, const float *q, uint64_t nq, float *G,
                    float *det, uint32_t *n_ext);

/* ---- screen-space depth smoothing: GaussRenderPass (GaussRenderPass.cpp:8-48,
 * 347-358; assets/shaders/advanced/gauss.frag) ----
 * pm_gauss_kernel: ComputeGaussKernel's table, K[j + i (N + 1)] =
 * exp(-r^2 / 2) (r = sqrt(i^2 + j^2), powf as the reference evaluates it)
 * divided by its sum over the (2N + 1)^2 window; host code.
 * pm_depth_smooth: out(x, y) = sum over i, j in [-N, N] (i horizontal, outer)
 * of K(|i|, |j|) depth(x + i, y + j), each term added with one fused
 * multiply-add (gauss.frag's `sum += k * t` may be fused by the shader
 * compiler), coordinates clamped to the image (the
 * depth texture's clamp-to-edge sampler, BilateralBuffer.cpp:246-248) at
 * texel centres (Spread = 1).  Host images, W*H floats, row-major; 0 <= N <= 31
 * (the reference's 32 x 32 table). */
int pm_gauss_kernel(int32_t N, float *K);
int pm_depth_smooth(pm_ctx *ctx, const float *depth, uint32_t W, uint32_t H, int32_t N, float *out);
/* pm_depth_smooth_device: the same pass on device memory (the sprite depth
 * image the application already holds on the GPU, as GaussRenderPass reads its
 * attachment): d_depth and d_out are W*H floats of the context's GPU, not
 * overlapping; the kernel is enqueued on `stream` (a hipStream_t, NULL for the
 * null stream) and the call returns without waiting.  The buffers must stay
 * valid, and the context alive, until that stream has run it. */
int pm_depth_smooth_device(pm_ctx *ctx, const float *d_depth, uint32_t W, uint32_t H, int32_t N, float *d_out,
                           void *stream);

/* ---- depth sprites: the reference march's ray starts (DepthRenderPass.cpp:
 * 243-250; sprite quads AdvancedRenderer.cpp:455-474; depth.vert, depth.frag;
 * depth test LESS, clear 1.0) ----
 * pm_depth_sprites draws every particle of `frame` as a camera-facing quad
 * p + R (+-x +-y) (x, y: the first two rows of View) and writes, per pixel, the
 * least depth.frag depth: at the pixel centre's ray through the quad's plane,
 * UV = ((P - p).x, -(P - p).y) / R, discarded when |UV|^2 > 1, else the
 * projected z of View P - R (0, 0, cos(pi/2 |UV|)).  Fragments whose plane
 * depth is outside [0, 1] are clipped, depths are clamped to [0, 1], and a
 * particle with a quad corner behind the camera is not drawn.  view and proj
 * are View and Projection (column-major, as the uniforms hold them), inv_pv and
 * cam the march's; depth is W*H floats (host), 1.0 where nothing is drawn:
 * the depth_in of a reference-mode march. */
int pm_depth_sprites(pm_ctx *ctx, int32_t frame, const float view[16], const float proj[16], const float inv_pv[16],
                     const float cam[3], float radius, uint32_t W, uint32_t H, float *depth);
/* pm_depth_sprites_device: the same image into d_depth, W*H floats of device
 * memory of the context's GPU, enqueued on the context's stream without a
 * wait: a pm_march_async_device on the same context that reads d_depth runs
 * after it (the reference pipeline with no host copy); pm_wait, or any call
 * that waits for the context, waits for it too. */
int pm_depth_sprites_device(pm_ctx *ctx, int32_t frame, const float view[16], const float proj[16],
                            const float inv_pv[16], const float cam[3], float radius, uint32_t W, uint32_t H,
                            float *d_depth);

/* ---- learned step proposal (no counterpart in the reference; it replaces the
 * role of the DensityGrid skip, Dataset.cpp:94-164, RayMarcher.cpp:284-304) ----
 * A 24 -> 64 -> 64 -> 1 MLP, BF16 weights with FP32 accumulation on MFMA,
 * predicts log2(1 + d/h) for d the distance along the ray to the first iso
 * crossing.  The configured march with learned = 1 keeps the exact skips; at a
 * sample below iso they do not pass, it may jump n lattice samples ahead, but
 * only when a bound of the density over the segment of the n - 1 samples it
 * passes (the sum of W at each particle's distance to the segment) proves them
 * below iso.  The first sample at or above iso -- and the image -- are
 * therefore the plain march's.
 * Parameters: PM_MLP_NPARAM floats, W1[64][32] (inputs 24..31 zero), b1[64],
 * W2[64][64], b2[64], w3[64], b3. */
#define PM_MLP_INPUTS 24
#define PM_MLP_PARAMS 6337
int pm_mlp_set(pm_ctx *ctx, const float *params);
int pm_mlp_get(pm_ctx *ctx, float *params);
/* Batch inference of n rows of 24 features (host arrays). */
int pm_mlp_forward(pm_ctx *ctx, const float *x, uint64_t n, float *y);
/* Loss sum_i w_i (y_i - t_i)^2 / n (w_i = w_over when y_i > t_i) and its
 * gradient for the loaded network, n a multiple of 64 (host arrays). */
int pm_mlp_grad(pm_ctx *ctx, const float *x, const float *target, uint64_t n, float w_over, float *loss,
                float *grad);
/* Device training set: clear (capacity rows), append traces of the plain
 * configured march of one view (rays with hash % ray_stride == 0, every
 * k_stride-th sample before the crossing, at most max_new_rows rows, 0 = up to
 * the capacity), read rows back (features, label log2(1 + d/h), hit flag, and
 * the (pixel, sample index) each row came from). */
int pm_dataset_clear(pm_ctx *ctx, uint64_t capacity);
int pm_trace_append(pm_ctx *ctx, int32_t frame, const pm_settings *s, const float inv_pv[16], const float cam[3],
                    uint32_t W, uint32_t H, uint32_t ray_stride, uint32_t k_stride, uint32_t seed,
                    uint64_t max_new_rows, uint64_t *rows_total);
int pm_dataset_get(pm_ctx *ctx, uint64_t offset, uint64_t n, float *feat, float *label, uint8_t *is_hit,
                   uint32_t *meta);
/* Adam on the device training set: epochs x (rows / batch) steps, batch a
 * multiple of 64; mean batch loss per epoch into loss_per_epoch (may be NULL). */
int pm_mlp_train(pm_ctx *ctx, int epochs, uint32_t batch, float lr, float w_over, uint64_t seed,
                 float *loss_per_epoch);

/* ---- a march that allocates nothing: no counterpart in the reference; for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * serving loops, harnesses that forbid hipMalloc) ----
 * pm_march_async_device's configured isotropic march (plain or learned), with
 * every buffer it needs beyond its inputs and outputs -- hit records, the tile
 * pre-pass's lists, counters -- in one static device array per GPU
 * (pm_noalloc_scratch_bytes(), about 34 MB, held until the process ends).
 * Limits: W * H <= PM_NOALLOC_MAX_PIXELS and W, H <= 16384 (PM_ERR_SIZE); the
 * frame must be set and, with PM_SKIP_BOUND, its bound bits for s->iso and
 * s->bound_boxes built (one pm_march_async_device or pm_get_bound_grid does
 * it; PM_ERR_STATE otherwise); the images are device memory of the context's GPU.  Nothing is
 * queued when a check fails.  Completes like pm_march_async_device (pm_query,
 * pm_wait, pm_get_stats).  Every call on one GPU shares that GPU's array:
 * marches that overlap on one GPU (two contexts, two threads) race. */
#define PM_NOALLOC_MAX_PIXELS (1u << 21)
int pm_march_device_noalloc(pm_ctx *ctx, int32_t frame, const pm_settings *s, const float inv_pv[16],
                            const float cam[3], uint32_t W, uint32_t H, void *d_positions, void *d_normals,
                            void *d_t);
uint64_t pm_noalloc_scratch_bytes(void);

/* Measurement utility, no counterpart in the reference: times `iters`
 * launches of a streaming copy of `bytes` bytes (read + write) on the context
 * stream and returns the achieved bandwidth, counting 2*bytes per launch. */
int pm_stream_copy_gbps(pm_ctx *ctx, uint64_t bytes, int iters, float *gbps);

#ifdef __cplusplus
}
#endif
#endif
