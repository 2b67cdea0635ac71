/*
 * pm_oracle.h -- CPU oracle for the particle-fluid iso-surface ray march.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: it is loaded by
 * tests/, by __graft_entry__.smoke() and by bench.py's cpu_baseline leg, and
 * never by the product path (the HIP library under bachelor-thesis_amd/).
 *
 * It restates, in plain C compiled with -ffp-contract=off, the reference's CPU
 * marcher (src/app/AdvancedRenderer/RayMarcher.cpp:51-62, 256-344), its SPH
 * kernel (src/app/Kernel.cpp:8-52), its frame structures (src/app/Dataset.cpp:
 * 26-47, 78-164, 272-280) and its work scheduler (src/app/ThreadPool.cpp:5-55),
 * plus the configured march (AABB entry, fixed step, bisection, central-
 * difference normals) that the GPU path implements.  Every float operation is
 * written out in the order the GPU kernels use, so FP32 results can be compared
 * bit for bit; the FP64 march is the ground truth errors are stated against.
 */
#ifndef PM_ORACLE_H
#define PM_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* pm_settings (include/pmarch.h): its first ten fields, then its anisotropic
 * fields, same names and meaning. */
typedef struct or_settings {
    int32_t mode;          /* 0 = configured march, 1 = reference march          */
    int32_t max_steps;     /* reference mode: samples that count (MaxSteps)       */
    float   step;          /* march step in world units                           */
    float   iso;           /* iso density, absolute kernel units                  */
    int32_t bisect_iters;  /* configured mode: bisection iterations              */
    int32_t final_fp64;    /* configured mode: last bisection iteration in FP64   */
    float   normal_eps;    /* configured mode: central-difference offset          */
    int32_t normal_mode;   /* 0 = central difference of rho, 1 = sum of gradW     */
    int32_t skip_mode;     /* 0 none; 1 empty-stencil skip (configured mode);
                              2 DensityGrid, box flag; 3 DensityGrid, conservative
                              flag (reference mode); 4 sub-box bound skip
                              (configured mode, needs or_bound_build)             */
    int32_t compensated;   /* 1 = Neumaier-compensated FP32 density sums          */
    /* anisotropic density (PerPixel_Anisotropic, RayMarcher.cpp:346-423) */
    int32_t anisotropic;   /* 1 = WPCA + AnisotropicKernel instead of CubicSpline  */
    float   k_n, k_r, k_s; /* VisualizationSettings k_n, k_r, k_s (RayMarcher.h:22-24) */
    int32_t n_eps;         /* VisualizationSettings N_eps                          */
    float   ext_mult;      /* ParticleRadiusExt / ParticleRadius (>= 1)            */
    int32_t bound_boxes;   /* skip_mode 4: 0 = 8x8x8 sub-boxes, 1 = 4x4x4 only     */
} or_settings;

enum { OR_MODE_CONFIG = 0, OR_MODE_REFERENCE = 1 };

typedef struct or_counters {
    uint64_t rays;          /* primary rays processed                      */
    uint64_t rays_in_aabb;  /* rays that intersect the padded AABB         */
    uint64_t hits;          /* rays that found an iso crossing             */
    uint64_t samples;       /* density evaluations of every kind           */
    uint64_t cand;          /* sum over samples of stencil candidates (the
                               six central-difference samples: their fused
                               walk's candidates once)                     */
    uint64_t march_samples; /* density evaluations of the march phase      */
    /* configured isotropic march over the dense grid only (0 otherwise):   */
    uint64_t march_cand;    /* march samples' own fine-stencil candidates  */
    uint64_t cand_stencil;  /* every density evaluation's own fine-stencil
                               candidates (the FP64 sample: those of its
                               FP32-rounded point)                         */
    uint64_t cand27;        /* every density evaluation's 27 search-cell
                               candidates                                  */
} or_counters;

/* Dense uniform grid, cell width h, cell keys as CompactNSearch computes them
 * (int(x/h), minus one for negative x), particles stored cell-major (x fastest)
 * and, inside a cell, by ascending input index. */
typedef struct or_grid {
    float    h, inv_h;
    float    pmin[3], pmax[3];      /* particle bounds                       */
    float    bmin[3], bmax[3];      /* bounds padded by h (Frame::ComputeAABB) */
    int32_t  origin[3];             /* key of dense cell (0,0,0)             */
    int32_t  dims[3];
    uint64_t ncells;
    uint64_t n;
    uint32_t *cell_start;           /* ncells + 1, exclusive scan of counts  */
    float    *sorted;               /* 4*n: x, y, z, unused                  */
    uint32_t *sorted_index;         /* input index of each stored particle   */
    /* reference DensityGrid over the padded AABB (Dataset.cpp:94-164)       */
    int32_t  dg_dims[3];
    uint8_t  *dg_flag_box;          /* any particle with |x_j-c|_inf < h      */
    uint8_t  *dg_flag_cons;         /* any particle with |x_j-c|_inf < 1.5h   */
    uint32_t *dg_count_box;
    /* 1 where the cell and its 26 neighbours hold no particle: every point of
     * the cell has zero density (empty-space skip of the configured march)   */
    uint8_t  *empty27;
    /* sub-box density bounds for one iso (or_bound_build): 8 words per cell,
     * bit 8 l + (fx&1) + 2 (fy&1) + 4 (fz&1) set when sub-box (fx,fy,fz) of the
     * cell's 8x8x8 stays below iso, l = fx/2 + 4 (fy/2) + 16 (fz/2); with
     * bound_boxes 1 every byte is 0xff or 0 (the 4x4x4 boxes' bounds alone)   */
    uint64_t *bound;
    float     bound_iso;
    int32_t   bound_boxes;
    /* march grid: cells of h/2 (the same key rule with inv_hf = 2/h), four
     * empty cells of margin on each side, particles fine-cell-major (x
     * fastest) and by input index inside a cell.  Every density sum walks the
     * 5x5x5 fine cells around the point's fine cell as x-rows (dz outer, dy),
     * dropping rows and row ends farther than h (or_fine_row).               */
    float     inv_hf;
    int32_t   forigin[3];
    int32_t   fdims[3];
    uint64_t  fncells;
    uint32_t *fcell_start;          /* fncells + 1                           */
    float    *fsorted;              /* 4*n                                   */
} or_grid;

/* ---- SPH cubic spline kernel (Kernel.cpp:8-52) ---- */
float  or_W_f(float r2, float h);
double or_W_d(double r2, double h);
void   or_gradW_f(const float r[3], float h, float out[3]);
float  or_sigma_f(float h);

/* ---- frame structures ---- */
int32_t or_key(float x, float inv_h);
int  or_grid_build(const float *xyz, uint64_t n, float h, or_grid *g);
void or_grid_free(or_grid *g);
uint64_t or_grid_struct_size(void);

/* Sub-box density bounds for iso (skip_mode 4); replaces any earlier bits. */
int  or_bound_build(or_grid *g, float iso, int32_t boxes);

/* ---- neighbour search ---- */
uint32_t or_neighbors_brute(const float *xyz, uint64_t n, float h, const float q[3],
                            uint32_t *out, uint32_t cap);
uint32_t or_neighbors_grid(const or_grid *g, const float q[3], uint32_t *out, uint32_t cap,
                           uint32_t *ncand);
uint32_t or_neighbors_hash(const float *xyz, uint64_t n, float h, const float q[3],
                           uint32_t *out, uint32_t cap);

/* ---- march-grid stencil ---- */
/* Fine-cell stencil centre of a point and its offsets inside the cell; 0 when
 * the 5x5x5 stencil would leave the grid (no particle within h). */
typedef struct or_fine { int32_t c[3]; float f[3]; } or_fine;
int or_fine_centre(const or_grid *g, const float q[3], or_fine *fs);
/* Row (dy, dz) of the stencil: 0 if no point of it lies within h; else the
 * stored range [b, e) of its fine cells within reach. */
int or_fine_row(const or_grid *g, const or_fine *fs, int dy, int dz, uint32_t *b, uint32_t *e);

/* Particles with |q - x_j|^2 < h^2 in the march grid's order (input indices
 * are not stored for the march grid, so positions are returned: 3 per hit). */
uint32_t or_neighbors_fine(const or_grid *g, const float q[3], float *out, uint32_t cap, uint32_t *ncand);

/* Candidates of the fused central-difference walk around q (box [q - eps,
 * q + eps]); 0 when its cells leave the grid.  out (optional): positions of
 * the candidates within (h + eps)(1 + 2^-10) of q, 3 floats each. */
int or_cdiff_union(const or_grid *g, const float q[3], float eps, uint32_t *ncand, float *out, uint32_t cap,
                   uint32_t *nout);

/* ---- density ---- */
float  or_density_f(const or_grid *g, const float q[3], int compensated, uint32_t *ncand);
double or_density_d(const or_grid *g, const double q[3]);
void   or_density_batch_f(const or_grid *g, const float *q, uint64_t nq, int compensated,
                          float *rho, uint32_t *ncand);
void   or_density_batch_d(const or_grid *g, const double *q, uint64_t nq, double *rho);

/* ---- anisotropic density (RayMarcher.cpp:114-254, 378-409; Kernel.cpp:55-125) ----
 * The extended neighbours (|x_j - q|^2 < (ext_mult h)^2, at most 8192) are
 * walked on the search grid (cells of h) over the cells within 2 cells of q,
 * rows dz outer, dy, x, by input index inside a cell. */
float  or_density_aniso_f(const or_grid *g, const or_settings *s, const float q[3], int compensated, uint32_t *ncand);
double or_density_aniso_d(const or_grid *g, const or_settings *s, const double q[3]);
void   or_gradW_aniso_f(const or_grid *g, const or_settings *s, const float q[3], float out[3]);
/* G (row major) and det G of WPCA at q; returns the number of extended
 * neighbours (0: G undefined). */
uint32_t or_aniso_G_f(const or_grid *g, const or_settings *s, const float q[3], float G[9], float *detG);
uint32_t or_aniso_G_d(const or_grid *g, const or_settings *s, const double q[3], double G[9], double *detG);
/* The weighted covariance C of WPCA at q (lower triangle c00, c10, c11, c20,
 * c21, c22); returns the number of extended neighbours.  G and det G from C
 * and that number (capped at 8192). */
uint32_t or_aniso_C_f(const or_grid *g, const or_settings *s, const float q[3], float c[6]);
uint32_t or_aniso_C_d(const or_grid *g, const or_settings *s, const double q[3], double c[6]);
void or_aniso_G_from_C_f(const or_grid *g, const or_settings *s, const float c[6], uint32_t n, float G[9],
                         float *detG);
/* Cyclic Jacobi eigen-decomposition of the symmetric 3x3 matrix given by its
 * lower triangle c = (c00, c10, c11, c20, c21, c22): eigenvalues lam[k],
 * eigenvector k = column k of V (row major). */
void or_eig3_f(const float c[6], float lam[3], float V[9]);
void or_eig3_d(const double c[6], double lam[3], double V[9]);
/* The extended neighbours of q in walk order (positions, 3 floats each). */
uint32_t or_neighbors_ext(const or_grid *g, const float q[3], float hx, float *out, uint32_t cap, uint32_t *ncand);

/* ---- screen-space depth smoothing (GaussRenderPass.cpp:8-48, gauss.frag) ---- */
int or_gauss_kernel(int32_t N, float *K);
int or_depth_smooth(const float *depth, uint32_t W, uint32_t H, int32_t N, float *out);

/* ---- depth sprites (DepthRenderPass, depth.vert/frag, AdvancedRenderer.cpp:455-474) ---- */
float or_cos_half_pi(float s);
int or_depth_sprites(const float *xyz, uint64_t n, const float view[16], const float proj[16], const float inv_pv[16],
                     const float cam[3], float R, uint32_t W, uint32_t H, float *depth);

/* ---- ray generation (RayMarcher.cpp:268-274) and AABB slab test ---- */
void or_ray(const float invpv[16], const float cam[3], uint32_t W, uint32_t H, uint32_t idx,
            float o[3], float d[3]);
int  or_slab(const float o[3], const float d[3], const float bmin[3], const float bmax[3], float tnf[2]);

/* ---- march ---- */
int or_march(const or_grid *g, const or_settings *s, const float invpv[16], const float cam[3],
             uint32_t W, uint32_t H, const float *depth, uint32_t row_step, int nthreads,
             float *pos, float *nrm, float *tout, or_counters *cnt);
/* Candidates of q's own trimmed fine stencil (no particle read) and the
 * particles in the 27 search cells around q's cell (0 when those leave the
 * grid): the per-sample counts of the GPU's instrumented launches. */
uint32_t or_stencil_cand(const or_grid *g, const float q[3]);
uint32_t or_cand27(const or_grid *g, const float q[3]);
int or_march_f64(const or_grid *g, const or_settings *s, const float invpv[16], const float cam[3],
                 uint32_t W, uint32_t H, uint32_t row_step, int nthreads,
                 double *pos, double *nrm, double *tout);

/* ---- reference-style CPU baseline: per-sample hashed 27-cell query ---- */
typedef struct or_hash or_hash;
or_hash *or_hash_build(const float *xyz, uint64_t n, float h);
void     or_hash_free(or_hash *hs);
int or_march_hashed(const or_hash *hs, const or_grid *g, const or_settings *s,
                    const float invpv[16], const float cam[3], uint32_t W, uint32_t H,
                    uint32_t row_step, int nthreads, float *pos, float *nrm, float *tout,
                    or_counters *cnt);

#ifdef __cplusplus
}
#endif
#endif
