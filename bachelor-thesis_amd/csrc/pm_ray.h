// pm_ray.h -- per-ray pieces shared by the march kernels (isotropic and
// anisotropic): ray set-up (RayMarcher.cpp:268-274), AABB slab test, the
// empty-space and bound skips, the reference DensityGrid skip
// (RayMarcher.cpp:284-304), counters and the XCD-aware tile mapping.
#pragma once

#include "pm_internal.h"

namespace pm {

// PH_MARCH: the next lattice sample; PH_BISECT: FP32 bisection steps (the
// anisotropic march and the one-ray-per-lane schedule); PH_PROPOSE: waiting for
// a step proposal of the learned march (taken by the whole wave at once).
enum Phase : int {
    PH_MARCH = 0, PH_BISECT = 1, PH_DONE = 5, PH_PROPOSE = 6
};

// glm mat4 * vec4 (type_mat4x4.inl): (m0 x + m1 y) + (m2 z + m3 w), column-major M.
__device__ __forceinline__ void pm_unproject(const float *M, float cx, float cy, float cz, float &px, float &py,
                                             float &pz)
{
    float w[4];
#pragma unroll
    for (int r = 0; r < 4; r++) {
        float a0 = M[0 * 4 + r] * cx, a1 = M[1 * 4 + r] * cy;
        float a2 = M[2 * 4 + r] * cz, a3 = M[3 * 4 + r] * 1.0f;
        w[r] = (a0 + a1) + (a2 + a3);
    }
    px = w[0] / w[3];
    py = w[1] / w[3];
    pz = w[2] / w[3];
}

// Slab test of the ray against [bmin, bmax] for t >= 0.
__device__ __forceinline__ bool pm_slab(const float o[3], const float d[3], const float *bmin, const float *bmax,
                                        float &tn, float &tf)
{
    float tnear = 0.0f, tfar = INFINITY;
    bool ok = true;
#pragma unroll
    for (int a = 0; a < 3; a++) {
        if (d[a] != 0.0f) {
            float inv = 1.0f / d[a];
            float t0 = (bmin[a] - o[a]) * inv, t1 = (bmax[a] - o[a]) * inv;
            float lo = t0 < t1 ? t0 : t1, hi = t0 < t1 ? t1 : t0;
            if (lo > tnear) tnear = lo;
            if (hi < tfar) tfar = hi;
        } else if (o[a] < bmin[a] || o[a] > bmax[a]) {
            ok = false;
        }
    }
    tn = tnear;
    tf = tfar;
    return ok && tnear <= tfar;
}

// sum_j gradW(x_j - q) (Kernel.cpp:34-52, with its extra 1/|r|; RayMarcher.cpp:331-336).
__device__ __forceinline__ void pm_sum_gradw(const GridDev &g, float qx, float qy, float qz, float &gx, float &gy,
                                             float &gz, uint32_t &cand)
{
    FineWalk fw;
    gx = gy = gz = 0.0f;
    if (!pm_fine_begin(g, qx, qy, qz, fw)) return;
    _Pragma("unroll 1") for (int dz = -2; dz <= 2; dz++)
        _Pragma("unroll 1") for (int dy = -2; dy <= 2; dy++) {
            uint32_t b, e;
            pm_fine_row(g, fw, dy, dz, pm_fine_tz2(fw, dz), b, e);
            cand += e - b;
            const int m = (int)(e - b);
            const float4 *const p0 = g.fsorted + b;
            for (int i = 0; i < m; i += 4) {
              float4 v[4];
#pragma unroll
              for (int u = 0; u < 4; u++) v[u] = p0[i + u < m ? i + u : i];
#pragma unroll
              for (int u = 0; u < 4; u++) {
                const float4 p = v[u];
                float r0 = p.x - qx, r1 = p.y - qy, r2c = p.z - qz;
                float rn = (r0 * r0 + r1 * r1) + r2c * r2c;
                if (i + u < m && rn < g.h2) {
                    float rl = sqrtf(rn);
                    float q = rl * g.inv_h;
                    float inv = 1.0f / sqrtf(rn);
                    float den = rl * g.h;
                    float q0 = (r0 * inv) / den, q1 = (r1 * inv) / den, q2 = (r2c * inv) / den;
                    float w0, w1, w2;
                    if (q >= 0.5f) {
                        float qq = 1.0f - q;
                        float f = 6.0f * qq * qq;
                        float ns = -g.sig;
                        w0 = (ns * q0) * f; w1 = (ns * q1) * f; w2 = (ns * q2) * f;
                    } else {
                        float f = 6.0f * (3.0f * q * q - 2.0f * q);
                        w0 = (g.sig * q0) * f; w1 = (g.sig * q1) * f; w2 = (g.sig * q2) * f;
                    }
                    gx += w0; gy += w1; gz += w2;
                }
              }
            }
        }
}

__device__ __forceinline__ void pm_normalize(float x, float y, float z, float &ox, float &oy, float &oz)
{
    float dd = (x * x + y * y) + z * z;
    float inv = 1.0f / sqrtf(dd);
    ox = x * inv; oy = y * inv; oz = z * inv;
}

__device__ __forceinline__ unsigned long long wave_sum(unsigned long long v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// This workgroup's counter slot.
__device__ __forceinline__ unsigned long long *ctr_slot(unsigned long long *ctr)
{
    return ctr + (blockIdx.x & (kCtrSlots - 1)) * kCtrStride;
}

// Wave totals into the workgroup's slot; zero totals add nothing (atomics on
// one line serialise, so each saved one counts).
__device__ __forceinline__ void flush_counters(unsigned long long *ctr, uint32_t rays, uint32_t in_aabb, uint32_t hits,
                                               uint32_t samples, unsigned long long cand, uint32_t msamples)
{
    unsigned long long r = wave_sum(rays), a = wave_sum(in_aabb), h = wave_sum(hits);
    unsigned long long s = wave_sum(samples), c = wave_sum(cand), m = wave_sum(msamples);
    if ((threadIdx.x & 63) == 0) {
        unsigned long long *p = ctr_slot(ctr);
        if (r) atomicAdd(&p[C_RAYS], r);
        if (a) atomicAdd(&p[C_IN_AABB], a);
        if (h) atomicAdd(&p[C_HITS], h);
        if (s) atomicAdd(&p[C_SAMPLES], s);
        if (c) atomicAdd(&p[C_CAND], c);
        if (m) atomicAdd(&p[C_MARCH_SAMPLES], m);
    }
}

// Per-lane counters of one launch (pm_stats): cand counts every candidate a
// walk reads; mcand those of the march samples' walks; cst and c27 (instrumented
// launches only) the candidates of every sample's own fine stencil and of its
// 27 search cells, however the sample was evaluated.
struct Counts {
    uint32_t rays = 0, in_aabb = 0, hits = 0, samples = 0, march = 0, mlp = 0, cert = 0, cskip = 0;
    unsigned long long cand = 0, mcand = 0, cst = 0, c27 = 0;
};

__device__ __forceinline__ void flush_counts(unsigned long long *ctr, const Counts &cn)
{
    const unsigned long long v[C_COUNT] = {
        wave_sum(cn.rays), wave_sum(cn.in_aabb), wave_sum(cn.hits), wave_sum(cn.samples), wave_sum(cn.cand),
        wave_sum(cn.march), wave_sum(cn.mlp), wave_sum(cn.mcand), wave_sum(cn.cert), wave_sum(cn.cskip),
        wave_sum(cn.cst), wave_sum(cn.c27) };
    if ((threadIdx.x & 63) == 0) {
        unsigned long long *p = ctr_slot(ctr);
#pragma unroll
        for (int i = 0; i < C_COUNT; i++)
            if (v[i]) atomicAdd(&p[i], v[i]);
    }
}

// XCD-aware tile of this wave; false if the wave has no tile.
__device__ __forceinline__ bool pm_tile(uint32_t W, uint32_t H, uint32_t &px, uint32_t &py)
{
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint32_t bid = blockIdx.x, nb = gridDim.x;
    uint32_t nfull = nb & ~7u;
    uint32_t rb = bid < nfull ? (bid & 7u) * (nfull >> 3) + (bid >> 3) : bid;
    uint32_t tile = rb * 4 + wave;
    uint32_t tiles_x = (W + 7) / 8, tiles_y = (H + 7) / 8;
    if (tile >= tiles_x * tiles_y) return false;
    px = (tile % tiles_x) * 8 + (lane & 7);
    py = (tile / tiles_x) * 8 + (lane >> 3);
    return true;
}

// Per-ray state of the march, carried across loop iterations.  The march of
// a ray ends with its hit record for k_finish, or with its miss.
struct RayState {
    float d0, d1, d2;     // direction (origin = camera)
    float tn, tf;         // AABB entry / exit
    float lo, hi;         // bisection bracket
    int k, it, phase;     // sample index, bisection step
    uint32_t idx;
    int lrn;              // learned march: 1 while the ray may ask for a proposal
    float rho;            // learned march: density of sample k (below iso) for the proposal
    float id0, id1, id2;  // 1 / direction (bound skip)
};



__device__ __forceinline__ void write_miss(const MarchLaunch &L, uint32_t idx)
{
    L.pos[idx] = make_float4(0.f, 0.f, 0.f, 0.f);
    L.nrm[idx] = make_float4(0.f, 0.f, 0.f, 0.f);
    if (L.t) L.t[idx] = -1.0f;
}

// A miss of the configured march: the outputs, and no hit record.
__device__ __forceinline__ void march_miss(const MarchLaunch &L, uint32_t idx)
{
    write_miss(L, idx);
    L.hit_rec[idx] = make_uint4(0u, 0u, 0u, 0u);
}

// Direction of the ray of pixel (px, py) (RayMarcher.cpp:268-274, with clip z = 1).
__device__ __forceinline__ void pm_ray_dir(const MarchLaunch &L, uint32_t px, uint32_t py, float d[3])
{
    float cx = (float)px * (2.0f / (float)L.W) - 1.0f;
    float cy = (float)py * (2.0f / (float)L.H) - 1.0f;
    float wx, wy, wz;
    pm_unproject(L.M, cx, cy, 1.0f, wx, wy, wz);
    pm_normalize(wx - L.cam[0], wy - L.cam[1], wz - L.cam[2], d[0], d[1], d[2]);
}

// Ray of pixel (px, py) and its AABB interval.  Leaves phase = PH_MARCH, or
// writes the miss and PH_DONE.
__device__ __forceinline__ void ray_setup(const MarchLaunch &L, RayState &r, uint32_t px, uint32_t py, Counts &cn)
{
    cn.rays++;
    r.idx = py * L.W + px;
    float d[3];
    pm_ray_dir(L, px, py, d);
    r.d0 = d[0]; r.d1 = d[1]; r.d2 = d[2];
    r.id0 = 1.0f / d[0]; r.id1 = 1.0f / d[1]; r.id2 = 1.0f / d[2];
    float o[3] = { L.cam[0], L.cam[1], L.cam[2] };
    if (pm_slab(o, d, L.g.bmin, L.g.bmax, r.tn, r.tf)) {
        r.phase = PH_MARCH;
        r.k = 0;
        r.lrn = L.s.learned;
        cn.in_aabb++;
    } else {
        march_miss(L, r.idx);
        r.phase = PH_DONE;
    }
}

// Empty-space skip: when the sample's cell has an empty 27-neighbourhood,
// every sample the ray places in that cell has zero density, so the march
// resumes at the last sample before the ray leaves the cell box
// [key h, (key+1) h].  The first-crossing sample -- and the image -- are those
// of the march without the skip.  Returns the next sample index.
__device__ __forceinline__ float pm_cell_exit(int32_t key, float h, float o, float d, float texit)
{
    if (d == 0.0f) return texit;
    float lo = (float)key * h, hi = (float)(key + 1) * h;
    float ta = ((d > 0.0f ? hi : lo) - o) / d;
    return ta < texit ? ta : texit;
}

__device__ __forceinline__ int pm_skip_empty(const GridDev &g, float o0, float o1, float o2, const RayState &r, float qx,
                                             float qy, float qz, float step)
{
    int32_t cx, cy, cz;
    if (!pm_stencil(g, qx, qy, qz, cx, cy, cz)) return r.k;
    uint32_t lin = (uint32_t)cx + (uint32_t)g.nx * ((uint32_t)cy + (uint32_t)g.ny * (uint32_t)cz);
    if (!g.empty27[lin]) return r.k;
    float texit = INFINITY;
    texit = pm_cell_exit(cx + g.ox, g.h, o0, r.d0, texit);
    texit = pm_cell_exit(cy + g.oy, g.h, o1, r.d1, texit);
    texit = pm_cell_exit(cz + g.oz, g.h, o2, r.d2, texit);
    float kf = floorf((texit - r.tn) / step);
    return kf > (float)r.k ? (int)kf : r.k + 1;
}

// Bound skip: when the sample's eighth box (8x8x8 per cell) has its bit set
// in g.bound, no point of the box reaches iso, so the march resumes at the
// first sample past the box exit; when all eight eighth boxes of its quarter
// box (4x4x4 per cell) are set, past the quarter box's exit.  Every sample it
// passes lies in a set box (within the widening the bound was computed for)
// and would have been below iso.  Returns the next sample index (r.k when
// there is nothing to skip).
__device__ __forceinline__ float pm_box_exit(float lo, float hi, float o, float d, float id, float texit)
{
    if (d == 0.0f) return texit;
    float ta = ((d > 0.0f ? hi : lo) - o) * id;
    return ta < texit ? ta : texit;
}

__device__ __forceinline__ int pm_skip_bound(const GridDev &g, float o0, float o1, float o2, const RayState &r, float qx,
                                             float qy, float qz, float inv_step)
{
    const float u0 = g.inv_h * qx, u1 = g.inv_h * qy, u2 = g.inv_h * qz;
    const int32_t k0 = qx >= 0.0f ? (int32_t)u0 : (int32_t)u0 - 1;
    const int32_t k1 = qy >= 0.0f ? (int32_t)u1 : (int32_t)u1 - 1;
    const int32_t k2 = qz >= 0.0f ? (int32_t)u2 : (int32_t)u2 - 1;
    const int32_t cx = k0 - g.ox, cy = k1 - g.oy, cz = k2 - g.oz;
    if (cx < 0 || cy < 0 || cz < 0 || cx >= g.nx || cy >= g.ny || cz >= g.nz) return r.k;
    const int f0 = min(7, (int)((u0 - (float)k0) * 8.0f));
    const int f1 = min(7, (int)((u1 - (float)k1) * 8.0f));
    const int f2 = min(7, (int)((u2 - (float)k2) * 8.0f));
    const uint32_t l = (uint32_t)((f0 >> 1) + 4 * (f1 >> 1) + 16 * (f2 >> 1));
    const uint64_t cell = (uint64_t)cx + (uint64_t)g.nx * ((uint64_t)cy + (uint64_t)g.ny * (uint64_t)cz);
    const uint32_t byte = (uint32_t)(g.bound[8 * cell + (l >> 3)] >> (8 * (l & 7u))) & 0xffu;
    if (byte == 0u) return r.k;
    float l0, l1, l2, span;
    if (byte == 0xffu) {   // the whole quarter box
        l0 = (float)k0 + 0.25f * (float)(f0 >> 1);
        l1 = (float)k1 + 0.25f * (float)(f1 >> 1);
        l2 = (float)k2 + 0.25f * (float)(f2 >> 1);
        span = 0.25f;
    } else {
        if (!((byte >> ((f0 & 1) + 2 * (f1 & 1) + 4 * (f2 & 1))) & 1u)) return r.k;
        l0 = (float)k0 + 0.125f * (float)f0;
        l1 = (float)k1 + 0.125f * (float)f1;
        l2 = (float)k2 + 0.125f * (float)f2;
        span = 0.125f;
    }
    float texit = INFINITY;
    texit = pm_box_exit(l0 * g.h, (l0 + span) * g.h, o0, r.d0, r.id0, texit);
    texit = pm_box_exit(l1 * g.h, (l1 + span) * g.h, o1, r.d1, r.id1, texit);
    texit = pm_box_exit(l2 * g.h, (l2 + span) * g.h, o2, r.d2, r.id2, texit);
    float kq = (texit - r.tn) * inv_step;
    kq -= 0x1p-8f + fabsf(kq) * 0x1p-18f;
    const float kf = floorf(kq);
    return kf >= (float)r.k ? (int)kf + 1 : r.k + 1;
}


// QueryDensityGrid (Dataset.cpp:26-47) followed by the Flag test: true when p
// lies in a DensityGrid node whose flag is clear (the node may be skipped), with
// the node's box in nmin/nmax.
__device__ __forceinline__ bool pm_dg_query(const GridDev &g, float x, float y, float z, float *nmin, float *nmax)
{
    const float p[3] = { x, y, z };
    const int32_t dims[3] = { g.dg_nx, g.dg_ny, g.dg_nz };
    const float inv = 1.0f / g.h;
    int32_t k[3];
#pragma unroll
    for (int a = 0; a < 3; a++) {
        k[a] = (int32_t)floorf((p[a] - g.bmin[a]) * inv);
        if (!(k[a] >= 0 && k[a] < dims[a])) return false;
    }
    uint32_t id = (uint32_t)k[0] + (uint32_t)dims[0] * ((uint32_t)k[1] + (uint32_t)dims[1] * (uint32_t)k[2]);
    if (g.dg_flag[id]) return false;
#pragma unroll
    for (int a = 0; a < 3; a++) {
        nmin[a] = g.bmin[a] + (float)k[a] * g.h;       // Dataset.cpp:132-133
        nmax[a] = nmin[a] + g.h;
    }
    return true;
}

__device__ __forceinline__ float pm_gmin(float x, float y) { return (y < x) ? y : x; }
__device__ __forceinline__ float pm_gmax(float x, float y) { return (x < y) ? y : x; }

// intersectAABB (RayMarcher.cpp:51-62): t in units of dir (the step vector).
__device__ __forceinline__ void pm_intersect_aabb(const float *o, const float *d, const float *bmin, const float *bmax,
                                                  float *out)
{
    float t1[3], t2[3];
#pragma unroll
    for (int a = 0; a < 3; a++) {
        float tmn = (bmin[a] - o[a]) / d[a], tmx = (bmax[a] - o[a]) / d[a];
        t1[a] = pm_gmin(tmn, tmx);
        t2[a] = pm_gmax(tmn, tmx);
    }
    float tnear = pm_gmax(pm_gmax(t1[0], t1[1]), t1[2]);
    float tfar = pm_gmin(pm_gmin(t2[0], t2[1]), t2[2]);
    float t = (tnear < tfar) ? tfar : tnear;
#pragma unroll
    for (int a = 0; a < 3; a++) out[a] = t >= 0.0f ? o[a] + d[a] * t : o[a];
}

}  // namespace pm
