// pm_internal.h -- host-side launchers shared by the kernels and the C ABI.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "pm_device.h"
#include "pmarch.h"

namespace pm {

// Per-launch counters (u64 each), accumulated by the march kernel.
enum Counter { C_RAYS = 0, C_IN_AABB, C_HITS, C_SAMPLES, C_CAND, C_MARCH_SAMPLES, C_MLP, C_COUNT };

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
};

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
};
size_t scan_tmp_size(uint32_t len);
hipError_t launch_grid_build(const GridBuildArgs &a, hipStream_t st);

hipError_t launch_march(const MarchLaunch &L, hipStream_t st);
hipError_t launch_density(const GridDev &g, const float *q, uint32_t nq, int compensated, float *rho,
                          hipStream_t st);
hipError_t launch_density64(const GridDev &g, const double *q, uint32_t nq, double *rho, hipStream_t st);
hipError_t launch_neighbor_counts(const GridDev &g, const float *q, uint32_t nq, uint32_t *counts,
                                  uint32_t *ncand, hipStream_t st);

// Host restatement of the CompactNSearch key, for the grid extent.
inline int32_t host_key(float x, float inv_h)
{
    float v = inv_h * x;
    return x >= 0.0f ? (int32_t)v : (int32_t)v - 1;
}

}  // namespace pm
