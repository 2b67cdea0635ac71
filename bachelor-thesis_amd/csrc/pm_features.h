// pm_features.h -- inputs of the step-proposal network at a march sample.
//
//  f[0..7]   particles in the h-cells at x + (j+1) h d, j = 0..7   (/8, capped at 32)
//  f[8..15]  particles in the 4h-cells at x + 4 j h d, j = 0..7    (/256, capped at 1024)
//  f[16..18] ray direction d
//  f[19]     distance to the AABB exit / h                         (/64, capped at 256)
//  f[20]     rho(x) / iso                                          (capped at 4)
//  f[21]     |grad rho(x)| h / iso                                 (/4, capped at 16)
//  f[22]     distance marched since the AABB entry / h             (/64, capped at 256)
//  f[23]     particles in the 27 cells around x                    (/128, capped at 512)
#pragma once

#include "pm_internal.h"

namespace pm {

// ---------------------------------------------------------------- features --

// Particles in the h-cell containing p (0 outside the grid).
__device__ __forceinline__ float cell_count(const GridDev &g, float x, float y, float z)
{
    int32_t cx = pm_key(x, g.inv_h) - g.ox, cy = pm_key(y, g.inv_h) - g.oy, cz = pm_key(z, g.inv_h) - g.oz;
    if (cx < 0 || cx >= g.nx || cy < 0 || cy >= g.ny || cz < 0 || cz >= g.nz) return 0.0f;
    uint32_t c = (uint32_t)cx + (uint32_t)g.nx * ((uint32_t)cy + (uint32_t)g.ny * (uint32_t)cz);
    return (float)(g.cell_start[c + 1] - g.cell_start[c]);
}

// Particles in the 4h-cell containing p (coarse cell = 4x4x4 dense cells).
__device__ __forceinline__ float coarse_count(const GridDev &g, const CoarseDev &cg, float x, float y, float z)
{
    int32_t cx = pm_key(x, g.inv_h) - g.ox, cy = pm_key(y, g.inv_h) - g.oy, cz = pm_key(z, g.inv_h) - g.oz;
    if (cx < 0 || cx >= g.nx || cy < 0 || cy >= g.ny || cz < 0 || cz >= g.nz) return 0.0f;
    uint32_t c = (uint32_t)(cx >> 2) + (uint32_t)cg.nx * ((uint32_t)(cy >> 2) + (uint32_t)cg.ny * (uint32_t)(cz >> 2));
    return (float)cg.count[c];
}

// grad rho at q over the trimmed fine stencil (the particles within h, in the
// canonical order), plain FP32 sums, for the network's inputs only; adds the
// walk's candidates to cand.
__device__ __forceinline__ void grad_feat(const GridDev &g, float qx, float qy, float qz, float &gx, float &gy,
                                          float &gz, uint32_t &cand)
{
    FineWalk fw;
    gx = gy = gz = 0.0f;
    if (!pm_fine_begin(g, qx, qy, qz, fw)) return;
    const float dwk = g.sig * g.inv_h;
    _Pragma("unroll 1") for (int dz = -2; dz <= 2; dz++) {
        const float tz2 = pm_fine_tz2(fw, dz);
        _Pragma("unroll 1") for (int dy = -2; dy <= 2; dy++) {
            uint32_t b, e;
            pm_fine_row(g, fw, dy, dz, tz2, b, e);
            cand += e - b;
            for (uint32_t i = b; i < e; i++) {
                const float4 p = g.fsorted[i];
                const float rx = p.x - qx, ry = p.y - qy, rz = p.z - qz;
                const float r2 = (rx * rx + ry * ry) + rz * rz;
                if (r2 < g.h2 && r2 > 0.0f) {
                    const float r = sqrtf(r2);
                    const float q = r * g.inv_h;
                    float dw;
                    if (q >= 0.5f) {
                        const float qq = 1.0f - q;
                        dw = -dwk * 6.0f * qq * qq;
                    } else {
                        dw = dwk * 6.0f * (3.0f * q * q - 2.0f * q);
                    }
                    const float f = -dw / r;          // d/dx W(|x_j - x|) = -W'(r) r_vec / r
                    gx += f * rx; gy += f * ry; gz += f * rz;
                }
            }
        }
    }
}

__device__ __forceinline__ void mlp_features(const GridDev &g, const CoarseDev &cg, const float *o, const float *d, float t, float tn,
                             float tf, float rho, float gmag, float cand, float iso, float *f)
{
    const float px = o[0] + t * d[0], py = o[1] + t * d[1], pz = o[2] + t * d[2];
    for (int j = 0; j < 8; j++) {
        float s = (float)(j + 1) * g.h;
        f[j] = fminf(cell_count(g, px + s * d[0], py + s * d[1], pz + s * d[2]), 32.0f) * 0.125f;
    }
    for (int j = 0; j < 8; j++) {
        float s = (float)j * 4.0f * g.h;
        f[8 + j] = fminf(coarse_count(g, cg, px + s * d[0], py + s * d[1], pz + s * d[2]), 1024.0f) * (1.0f / 256.0f);
    }
    f[16] = d[0];
    f[17] = d[1];
    f[18] = d[2];
    f[19] = fminf((tf - t) * g.inv_h, 256.0f) * (1.0f / 64.0f);
    f[20] = fminf(rho / iso, 4.0f);
    f[21] = fminf(gmag * g.h / iso, 16.0f) * 0.25f;
    f[22] = fminf((t - tn) * g.inv_h, 256.0f) * (1.0f / 64.0f);
    f[23] = fminf(cand, 512.0f) * (1.0f / 128.0f);
}

}  // namespace pm
