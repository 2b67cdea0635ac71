// pm_internal.h -- host-side launchers shared by the kernels and the C ABI.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "pm_device.h"
#include "pm_mlp.h"
#include "pmarch.h"

namespace pm {

// Per-launch counters (u64 each), accumulated by the march kernels, in the
// order of pm_stats.
enum Counter {
    C_RAYS = 0, C_IN_AABB, C_HITS, C_SAMPLES, C_CAND, C_MARCH_SAMPLES, C_MLP, C_MARCH_CAND, C_CERT, C_CERT_SKIP,
    C_CAND_STENCIL, C_CAND27, C_COUNT
};
// The counters live in kCtrSlots copies, one per XCD (a workgroup adds to
// slot blockIdx % 8), kCtrStride u64 (one 128-byte line) apart: atomics of
// different XCDs never meet on one address.  The host sums the slots.
constexpr int kCtrSlots = 8, kCtrStride = 16, kCtrWords = kCtrSlots * kCtrStride;

// Coarse occupancy grid of the step-proposal features: 4x4x4 dense cells.
struct CoarseDev {
    const uint32_t *count;
    int32_t nx, ny, nz;
};

// Anisotropic density parameters (PerPixel_Anisotropic, RayMarcher.cpp:346-423):
// h = ParticleRadius, hx = ParticleRadiusExt, sig = 8/pi (Kernel.cpp:57-61),
// the clamp knobs k_n, k_r, k_s, N_eps (RayMarcher.h:22-25); FP32 and FP64.
struct AnParF {
    using T = float;
    float h, h2, hinv, hx, R2x, hxinv, sig, k_n, k_r, k_s;
    int32_t n_eps;
};
struct AnParD {
    using T = double;
    double h, h2, hinv, hx, R2x, hxinv, sig, k_n, k_r, k_s;
    int32_t n_eps;
};
struct AnisoDev {
    AnParF f;
    AnParD d;
};

struct MarchLaunch {
    GridDev g;
    pm_settings s;
    float M[16];
    float cam[3];
    uint32_t W, H;
    const float *depth;   // device, reference mode
    float4 *pos;
    float4 *nrm;
    float *t;             // may be null
    unsigned long long *counters;
    uint32_t *band_ctr;        // 8 per-XCD tile counters (persistent schedule)
    uint32_t persist_blocks;   // workgroups of the persistent schedule, 0 = default
    // learned step proposal
    CoarseDev cg;
    MlpDev net;
    float mlp_alpha;           // fraction of the predicted distance that is skipped
    int32_t mlp_min_skip;      // below this many steps the ray continues step by step
    int32_t mlp_max_skip;      // longest jump, in steps
    float mlp_gate_rho;        // a sample asks for a proposal only below this density
    // tiles with a ray entering the AABB, in row order: scratch of
    // 3 (T + 1) + scan_tmp_size(T + 1) u32; null = every tile
    uint32_t *tile_buf;
    float inv_step;            // 1 / s.step (bound skip)
    // per pixel: the hit handed from the march to k_finish (1 | fp64 << 1,
    // lo, hi, 0), zeros for a miss; written for every pixel of a marched tile
    // (the anisotropic march: its list of hits)
    uint4 *hit_rec;
    AnisoDev an;               // s.anisotropic
};

// Batched launch: many (view, frame) jobs in one persistent kernel; every job
// is a complete MarchLaunch in device memory (read through a wave-uniform
// pointer, so its fields stay in scalar registers and the scalar cache).
// A run of consecutive 8x8 tiles of one job; waves take chunks, so the job's
// parameters stay uniform across a wave.
struct BatchChunk {
    uint32_t job, tile0, ntiles, pad;
};
struct BatchLaunch {
    MarchLaunch base;          // settings, counters and network (each job repeats them)
    const MarchLaunch *jobs;
    const BatchChunk *chunks;
    uint32_t nchunks;
    uint32_t *chunk_ctr;
};
hipError_t launch_bound_grid(const GridDev &g, float iso, bool quarter_only, uint64_t *out, hipStream_t st);
hipError_t launch_march_batch(const BatchLaunch &B, uint32_t blocks, hipStream_t st);
hipError_t launch_zero_u32(uint32_t *p, uint32_t n, hipStream_t st);
hipError_t launch_zero_u64(unsigned long long *p, uint32_t n, hipStream_t st);
// counters[0..n) = 0 except counters[C_RAYS] = rays0 (rays answered without marching)
hipError_t launch_init_counters(unsigned long long *p, uint32_t n, unsigned long long rays0, hipStream_t st);
// Tile pre-pass pieces: flags (tile_buf[t] = 1 if a ray of tile t enters the
// AABB, t < T; tile_buf[T] = 0) and the misses of the inactive tiles.
hipError_t launch_tile_flags(const MarchLaunch &L, uint32_t T, hipStream_t st);
hipError_t launch_miss_tiles(const MarchLaunch &L, uint32_t T, hipStream_t st);

// Grid build: bounds of n float3 points (device) into out[6] (device).
hipError_t launch_bounds_reduce(const float *xyz, uint64_t n, float *partial, uint32_t npartial,
                                float *out, hipStream_t st);
uint32_t bounds_partials(uint64_t n);

struct GridBuildArgs {
    const float *xyz;      // device AoS float3
    uint64_t n;
    float inv_h;
    int32_t ox, oy, oz, nx, ny, nz;
    uint32_t ncells;
    uint32_t *cell_of;     // n
    uint32_t *counts;      // ncells + 1 (scan in place into cell_start)
    uint32_t *cursor;      // ncells
    uint32_t *scan_tmp;    // scan_tmp_size(ncells + 1)
    float4 *sorted;        // n
    uint32_t *cell_start;  // ncells + 1
    float4 *scattered;     // n: cell-major, arbitrary order inside a cell
};
size_t scan_tmp_size(uint32_t len);
// Exclusive scan of len u32 (out may not alias in); tmp of scan_tmp_size(len).
hipError_t launch_scan_u32(const uint32_t *in, uint32_t *out, uint32_t len, uint32_t *tmp, hipStream_t st);
hipError_t launch_grid_build(const GridBuildArgs &a, hipStream_t st);

// The tile pre-pass of the configured march (L.tile_buf set): flags of the 8x8
// tiles one of whose rays enters the padded AABB, their scan, the list of
// active tiles (order[T] = their count) and the misses of the other tiles.
hipError_t launch_tile_prepass(const MarchLaunch &L, hipStream_t st);
hipError_t launch_march(const MarchLaunch &L, hipStream_t st);
// The anisotropic march (configured or reference mode), one kernel.
hipError_t launch_march_aniso(const MarchLaunch &L, hipStream_t st);
hipError_t launch_density_aniso(const GridDev &g, const AnParF &a, const float *q, uint32_t nq, int compensated,
                                float *rho, hipStream_t st);
hipError_t launch_density_aniso64(const GridDev &g, const AnParD &a, const double *q, uint32_t nq, double *rho,
                                  hipStream_t st);
hipError_t launch_aniso_matrix(const GridDev &g, const AnParF &a, const float *q, uint32_t nq, float *G, float *det,
                               uint32_t *next, hipStream_t st);
hipError_t launch_density(const GridDev &g, const float *q, uint32_t nq, int compensated, float *rho,
                          hipStream_t st);
hipError_t launch_density64(const GridDev &g, const double *q, uint32_t nq, double *rho, hipStream_t st);
hipError_t launch_neighbor_counts(const GridDev &g, const float *q, uint32_t nq, uint32_t *counts,
                                  uint32_t *ncand, hipStream_t st);

hipError_t launch_stream_copy(const void *src, void *dst, uint64_t bytes, hipStream_t st);
// Gaussian depth smoothing (GaussRenderPass): K = the (N + 1)^2 table of pm_gauss_kernel.
// The depth-sprite pass: View, Projection (column-major), the march's inverse
// projection-view and camera for the pixel rays, the particle radius.
struct SpriteLaunch {
    float view[16], proj[16], inv_pv[16], cam[3];
    float radius;
    uint32_t W, H;
};
// big: the large-footprint list (1 + cap u32: count, then indices), or null to
// draw every sprite with its own thread.
hipError_t launch_depth_sprites(const SpriteLaunch &S, const float4 *pts, uint32_t n, uint32_t *depth, uint32_t *big,
                                uint32_t cap, hipStream_t st);
hipError_t launch_depth_smooth(const float *in, float *out, uint32_t W, uint32_t H, int N, const float *K,
                               hipStream_t st);
hipError_t launch_empty27(const uint32_t *cs, int32_t nx, int32_t ny, int32_t nz, uint8_t *out, hipStream_t st);
hipError_t launch_densitygrid(const GridDev &g, int32_t dnx, int32_t dny, int32_t dnz, uint8_t *flag_box, uint8_t *flag_cons,
                              uint32_t *count_box, hipStream_t st);

hipError_t launch_coarse_counts(const GridDev &g, const CoarseDev &cg, uint32_t *out, hipStream_t st);

struct TraceArgs {
    GridDev g;
    CoarseDev cg;
    float M[16];
    float cam[3];
    uint32_t W, H;
    float step, iso;
    const float *t_hit;       // plain-march t per pixel (-1 = miss)
    uint32_t ray_stride, k_stride, seed;
    uint32_t *count;          // rows written so far (device)
    uint32_t cap;
    float *feat;              // [cap][24]
    float *label;             // [cap]
    uint8_t *is_hit;          // [cap]
    uint32_t *meta;           // [cap][2]: pixel index, sample index k
};
hipError_t launch_trace(const TraceArgs &a, hipStream_t st);

struct TrainArgs {
    MlpDev net;
    const float *feat;
    const float *label;
    const uint32_t *idx;      // batch sample ids (NULL: 0..B-1)
    float inv_batch, w_over;
    float *partial;           // [nwaves][NPARAM + 1]
};
struct AdamArgs {
    const float *partial;
    uint32_t nparts;
    float *p, *m, *v;
    float lr, b1, b2, eps, bc1, bc2;
    int apply;
    float *grad_out;          // optional summed gradient
    float *loss_out;
};
hipError_t launch_mlp_train(const TrainArgs &a, uint32_t nwaves, hipStream_t st);
hipError_t launch_mlp_adam(const AdamArgs &a, hipStream_t st);
hipError_t launch_mlp_sync(const float *p, uint16_t *W1, uint16_t *W2, uint16_t *W2T, hipStream_t st);
hipError_t launch_mlp_forward(const MlpDev &net, const float *feat, uint32_t n, float *y, hipStream_t st);

// Scratch of pm_march_device_noalloc (pm_noalloc.hip): the static device
// array of the device (looked up once per device), and its layout for a W x H
// march (false beyond the limits).
struct NoAllocScratch {
    uint4 *hit_rec;
    uint32_t *tiles;
    unsigned long long *counters;
    uint32_t *band;
};
hipError_t noalloc_scratch_init(int device, void **out);
bool noalloc_layout(void *base, uint32_t W, uint32_t H, NoAllocScratch &s);
size_t noalloc_scratch_bytes();

// Host restatement of the CompactNSearch key, for the grid extent.
inline int32_t host_key(float x, float inv_h)
{
    float v = inv_h * x;
    return x >= 0.0f ? (int32_t)v : (int32_t)v - 1;
}

}  // namespace pm
