// pm_march.hip -- the per-pixel iso-surface march on CDNA4.
//
// Replaces RayMarcher::PerPixel_Isotropic (RayMarcher.cpp:256-344) and the
// ThreadPool that drives it (ThreadPool.cpp:38-55).
//
// Work mapping: one wave64 per 8x8 pixel tile, one ray per lane, four waves
// per 256-thread workgroup.  Workgroups are dispatched round-robin over the 8
// XCDs; the block index is remapped so that each XCD receives a contiguous
// band of tiles and its L2 holds the particles of that band.
//
// Every lane runs a small state machine (march -> bisection -> FP64 bisection
// step -> normal samples) and each loop iteration performs exactly one density
// evaluation per active lane, whatever its phase, so the expensive 27-cell
// candidate loop stays converged across the wave while rays hit at different
// steps.
#include "pm_internal.h"

namespace pm {

enum Phase : int { PH_MARCH = 0, PH_BISECT = 1, PH_BISECT64 = 2, PH_NORMAL = 3, PH_GRADW = 4, PH_DONE = 5 };

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
    int32_t cx, cy, cz;
    gx = gy = gz = 0.0f;
    if (!pm_stencil(g, qx, qy, qz, cx, cy, cz)) return;
    for (int dz = -1; dz <= 1; dz++)
        for (int dy = -1; dy <= 1; dy++) {
            uint32_t base = (uint32_t)(cx - 1) + (uint32_t)g.nx * ((uint32_t)(cy + dy) + (uint32_t)g.ny * (uint32_t)(cz + dz));
            uint32_t b = g.cell_start[base], e = g.cell_start[base + 3];
            cand += e - b;
            for (uint32_t i = b; i < e; i++) {
                float4 p = g.sorted[i];
                float r0 = p.x - qx, r1 = p.y - qy, r2c = p.z - qz;
                float rn = (r0 * r0 + r1 * r1) + r2c * r2c;
                if (rn < g.h2) {
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

__device__ __forceinline__ void flush_counters(unsigned long long *ctr, uint32_t rays, uint32_t in_aabb, uint32_t hits,
                                               uint32_t samples, unsigned long long cand, uint32_t msamples)
{
    unsigned long long r = wave_sum(rays), a = wave_sum(in_aabb), h = wave_sum(hits);
    unsigned long long s = wave_sum(samples), c = wave_sum(cand), m = wave_sum(msamples);
    if ((threadIdx.x & 63) == 0) {
        atomicAdd(&ctr[C_RAYS], r);
        atomicAdd(&ctr[C_IN_AABB], a);
        atomicAdd(&ctr[C_HITS], h);
        atomicAdd(&ctr[C_SAMPLES], s);
        atomicAdd(&ctr[C_CAND], c);
        atomicAdd(&ctr[C_MARCH_SAMPLES], m);
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

__global__ __launch_bounds__(256) void k_march_config(MarchLaunch L)
{
    uint32_t px, py;
    if (!pm_tile(L.W, L.H, px, py)) return;   // wave-uniform
    const GridDev &g = L.g;
    const pm_settings &s = L.s;
    const bool active = px < L.W && py < L.H;
    const uint32_t idx = py * L.W + px;
    const bool comp = s.compensated != 0;

    uint32_t c_rays = 0, c_in = 0, c_hits = 0, c_samples = 0, c_ms = 0;
    unsigned long long c_cand = 0;

    float o[3] = { L.cam[0], L.cam[1], L.cam[2] }, d[3] = { 0.f, 0.f, 0.f };
    float tn = 0.0f, tf = 0.0f;
    int phase = PH_DONE;
    if (active) {
        c_rays = 1;
        float cx = (float)px * (2.0f / (float)L.W) - 1.0f;     // RayMarcher.cpp:268-270
        float cy = (float)py * (2.0f / (float)L.H) - 1.0f;
        float wx, wy, wz;
        pm_unproject(L.M, cx, cy, 1.0f, wx, wy, wz);
        pm_normalize(wx - o[0], wy - o[1], wz - o[2], d[0], d[1], d[2]);
        if (pm_slab(o, d, g.bmin, g.bmax, tn, tf)) {
            phase = PH_MARCH;
            c_in = 1;
        } else {
            L.pos[idx] = make_float4(0.f, 0.f, 0.f, 0.f);
            L.nrm[idx] = make_float4(0.f, 0.f, 0.f, 0.f);
            if (L.t) L.t[idx] = -1.0f;
        }
    }

    const int n32 = (s.final_fp64 && s.bisect_iters > 0) ? s.bisect_iters - 1 : s.bisect_iters;
    int k = 0, it = 0, ncd = 0;
    float lo = 0.f, hi = 0.f, th = 0.f;
    double tmd = 0.0;
    float xs0 = 0.f, xs1 = 0.f, xs2 = 0.f;
    float gv0 = 0.f, gv1 = 0.f, gv2 = 0.f, rplus = 0.f;

    while (phase != PH_DONE) {
        float qx = 0.f, qy = 0.f, qz = 0.f;
        if (phase == PH_MARCH) {
            float t = tn + (float)k * s.step;
            if (t > tf || (s.max_steps > 0 && k >= s.max_steps)) {
                L.pos[idx] = make_float4(0.f, 0.f, 0.f, 0.f);
                L.nrm[idx] = make_float4(0.f, 0.f, 0.f, 0.f);
                if (L.t) L.t[idx] = -1.0f;
                phase = PH_DONE;
                continue;
            }
            qx = o[0] + t * d[0]; qy = o[1] + t * d[1]; qz = o[2] + t * d[2];
        } else if (phase == PH_BISECT) {
            float tm = (lo + hi) * 0.5f;
            qx = o[0] + tm * d[0]; qy = o[1] + tm * d[1]; qz = o[2] + tm * d[2];
        } else if (phase == PH_NORMAL) {
            qx = xs0; qy = xs1; qz = xs2;
            int a = ncd >> 1;
            float e = s.normal_eps;
            if (a == 0) qx = (ncd & 1) ? xs0 - e : xs0 + e;
            else if (a == 1) qy = (ncd & 1) ? xs1 - e : xs1 + e;
            else qz = (ncd & 1) ? xs2 - e : xs2 + e;
        }

        if (phase == PH_BISECT64) {
            tmd = ((double)lo + (double)hi) * 0.5;
            double x = (double)o[0] + tmd * (double)d[0];
            double y = (double)o[1] + tmd * (double)d[1];
            double z = (double)o[2] + tmd * (double)d[2];
            double rho = pm_density64(g, x, y, z);
            c_samples++;
            double dlo = lo, dhi = hi;
            if (rho >= (double)s.iso) dhi = tmd; else dlo = tmd;
            double tt = (dlo + dhi) * 0.5;
            xs0 = (float)((double)o[0] + tt * (double)d[0]);
            xs1 = (float)((double)o[1] + tt * (double)d[1]);
            xs2 = (float)((double)o[2] + tt * (double)d[2]);
            th = (float)tt;
            phase = s.normal_mode == PM_NORMAL_GRADW ? PH_GRADW : PH_NORMAL;
            ncd = 0;
            continue;
        }
        if (phase == PH_GRADW) {
            float sx, sy, sz;
            uint32_t cand = 0;
            pm_sum_gradw(g, xs0, xs1, xs2, sx, sy, sz, cand);
            float nx, ny, nz;
            pm_normalize(sx, sy, sz, nx, ny, nz);
            L.pos[idx] = make_float4(xs0, xs1, xs2, 1.0f);
            L.nrm[idx] = make_float4(nx, ny, nz, 1.0f);
            if (L.t) L.t[idx] = th;
            c_hits++;
            phase = PH_DONE;
            continue;
        }

        uint32_t cand = 0;
        float rho = pm_density32(g, qx, qy, qz, comp, cand);
        c_samples++;
        c_cand += cand;

        if (phase == PH_MARCH) {
            c_ms++;
            if (rho >= s.iso) {
                if (k == 0) {
                    th = tn;
                    xs0 = o[0] + th * d[0]; xs1 = o[1] + th * d[1]; xs2 = o[2] + th * d[2];
                    phase = s.normal_mode == PM_NORMAL_GRADW ? PH_GRADW : PH_NORMAL;
                    ncd = 0;
                } else {
                    lo = tn + (float)(k - 1) * s.step;
                    hi = tn + (float)k * s.step;
                    it = 0;
                    if (n32 > 0) phase = PH_BISECT;
                    else if (s.final_fp64 && s.bisect_iters > 0) phase = PH_BISECT64;
                    else {
                        th = (lo + hi) * 0.5f;
                        xs0 = o[0] + th * d[0]; xs1 = o[1] + th * d[1]; xs2 = o[2] + th * d[2];
                        phase = s.normal_mode == PM_NORMAL_GRADW ? PH_GRADW : PH_NORMAL;
                        ncd = 0;
                    }
                }
            } else {
                k++;
            }
        } else if (phase == PH_BISECT) {
            float tm = (lo + hi) * 0.5f;
            if (rho >= s.iso) hi = tm; else lo = tm;
            it++;
            if (it >= n32) {
                if (s.final_fp64 && s.bisect_iters > 0) phase = PH_BISECT64;
                else {
                    th = (lo + hi) * 0.5f;
                    xs0 = o[0] + th * d[0]; xs1 = o[1] + th * d[1]; xs2 = o[2] + th * d[2];
                    phase = s.normal_mode == PM_NORMAL_GRADW ? PH_GRADW : PH_NORMAL;
                    ncd = 0;
                }
            }
        } else {  // PH_NORMAL
            if ((ncd & 1) == 0) {
                rplus = rho;
            } else {
                float gg = rplus - rho;
                int a = ncd >> 1;
                if (a == 0) gv0 = gg; else if (a == 1) gv1 = gg; else gv2 = gg;
            }
            ncd++;
            if (ncd == 6) {
                float dd = (gv0 * gv0 + gv1 * gv1) + gv2 * gv2;
                float nx = 0.f, ny = 0.f, nz = 0.f;
                if (dd > 0.0f) {
                    float inv = 1.0f / sqrtf(dd);
                    nx = -gv0 * inv; ny = -gv1 * inv; nz = -gv2 * inv;
                }
                L.pos[idx] = make_float4(xs0, xs1, xs2, 1.0f);
                L.nrm[idx] = make_float4(nx, ny, nz, 1.0f);
                if (L.t) L.t[idx] = th;
                c_hits++;
                phase = PH_DONE;
            }
        }
    }
    flush_counters(L.counters, c_rays, c_in, c_hits, c_samples, c_cand, c_ms);
}

// PerPixel_Isotropic (RayMarcher.cpp:256-344): start on the depth-buffer
// surface, fixed step vector, MaxSteps samples, first sample with rho >= iso,
// normal = normalize(sum gradW) (or central differences on request).
__global__ __launch_bounds__(256) void k_march_reference(MarchLaunch L)
{
    uint32_t px, py;
    if (!pm_tile(L.W, L.H, px, py)) return;
    const GridDev &g = L.g;
    const pm_settings &s = L.s;
    const bool active = px < L.W && py < L.H;
    const uint32_t idx = py * L.W + px;
    const bool comp = s.compensated != 0;
    uint32_t c_rays = 0, c_in = 0, c_hits = 0, c_samples = 0, c_ms = 0;
    unsigned long long c_cand = 0;
    if (active) {
        c_rays = 1;
        float z = L.depth[idx];
        float4 P = make_float4(0.f, 0.f, 0.f, 0.f), N = make_float4(0.f, 0.f, 0.f, 0.f);
        float tout = -1.0f;
        if (z != 1.0f) {
            c_in = 1;
            float cx = (float)px * (2.0f / (float)L.W) - 1.0f;
            float cy = (float)py * (2.0f / (float)L.H) - 1.0f;
            float p0, p1, p2, d0, d1, d2;
            pm_unproject(L.M, cx, cy, z, p0, p1, p2);
            pm_normalize(p0 - L.cam[0], p1 - L.cam[1], p2 - L.cam[2], d0, d1, d2);
            float st0 = s.step * d0, st1 = s.step * d1, st2 = s.step * d2;
            for (int i = 0; i < s.max_steps; i++) {
                p0 += st0; p1 += st1; p2 += st2;
                uint32_t cand = 0;
                float rho = pm_density32(g, p0, p1, p2, comp, cand);
                c_samples++; c_ms++; c_cand += cand;
                if (rho >= s.iso) {
                    float nx, ny, nz;
                    if (s.normal_mode == PM_NORMAL_GRADW) {
                        float sx, sy, sz;
                        uint32_t c2 = 0;
                        pm_sum_gradw(g, p0, p1, p2, sx, sy, sz, c2);
                        pm_normalize(sx, sy, sz, nx, ny, nz);
                    } else {
                        float gv[3];
                        float pp[3] = { p0, p1, p2 };
                        for (int a = 0; a < 3; a++) {
                            float xp[3] = { pp[0], pp[1], pp[2] }, xm[3] = { pp[0], pp[1], pp[2] };
                            xp[a] = pp[a] + s.normal_eps;
                            xm[a] = pp[a] - s.normal_eps;
                            uint32_t c3 = 0;
                            float rp = pm_density32(g, xp[0], xp[1], xp[2], comp, c3);
                            float rm = pm_density32(g, xm[0], xm[1], xm[2], comp, c3);
                            c_samples += 2; c_cand += c3;
                            gv[a] = rp - rm;
                        }
                        float dd = (gv[0] * gv[0] + gv[1] * gv[1]) + gv[2] * gv[2];
                        nx = ny = nz = 0.0f;
                        if (dd > 0.0f) {
                            float inv = 1.0f / sqrtf(dd);
                            nx = -gv[0] * inv; ny = -gv[1] * inv; nz = -gv[2] * inv;
                        }
                    }
                    P = make_float4(p0, p1, p2, 1.0f);
                    N = make_float4(nx, ny, nz, 1.0f);
                    float v0 = p0 - L.cam[0], v1 = p1 - L.cam[1], v2 = p2 - L.cam[2];
                    tout = sqrtf((v0 * v0 + v1 * v1) + v2 * v2);
                    c_hits = 1;
                    break;
                }
            }
        }
        L.pos[idx] = P;
        L.nrm[idx] = N;
        if (L.t) L.t[idx] = tout;
    }
    flush_counters(L.counters, c_rays, c_in, c_hits, c_samples, c_cand, c_ms);
}

hipError_t launch_march(const MarchLaunch &L, hipStream_t st)
{
    uint32_t tiles = ((L.W + 7) / 8) * ((L.H + 7) / 8);
    uint32_t blocks = (tiles + 3) / 4;
    if (L.s.mode == PM_MODE_REFERENCE)
        hipLaunchKernelGGL(k_march_reference, dim3(blocks), dim3(256), 0, st, L);
    else
        hipLaunchKernelGGL(k_march_config, dim3(blocks), dim3(256), 0, st, L);
    return hipGetLastError();
}

// ---- point queries used by the parity tests ----

__global__ __launch_bounds__(256) void k_density(GridDev g, const float *__restrict__ q, uint32_t nq, int comp,
                                                 float *__restrict__ rho)
{
    uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= nq) return;
    uint32_t cand = 0;
    rho[i] = pm_density32(g, q[3 * i], q[3 * i + 1], q[3 * i + 2], comp != 0, cand);
}

__global__ __launch_bounds__(256) void k_density64(GridDev g, const double *__restrict__ q, uint32_t nq,
                                                   double *__restrict__ rho)
{
    uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= nq) return;
    rho[i] = pm_density64(g, q[3 * i], q[3 * i + 1], q[3 * i + 2]);
}

__global__ __launch_bounds__(256) void k_neighbor_counts(GridDev g, const float *__restrict__ q, uint32_t nq,
                                                         uint32_t *__restrict__ counts, uint32_t *__restrict__ ncand)
{
    uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= nq) return;
    float qx = q[3 * i], qy = q[3 * i + 1], qz = q[3 * i + 2];
    int32_t cx, cy, cz;
    uint32_t cnt = 0, cand = 0;
    if (pm_stencil(g, qx, qy, qz, cx, cy, cz)) {
        for (int dz = -1; dz <= 1; dz++)
            for (int dy = -1; dy <= 1; dy++) {
                uint32_t base = (uint32_t)(cx - 1) + (uint32_t)g.nx * ((uint32_t)(cy + dy) + (uint32_t)g.ny * (uint32_t)(cz + dz));
                uint32_t b = g.cell_start[base], e = g.cell_start[base + 3];
                cand += e - b;
                for (uint32_t j = b; j < e; j++) {
                    float4 p = g.sorted[j];
                    float dx = qx - p.x, dy2 = qy - p.y, dz2 = qz - p.z;
                    float r2 = (dx * dx + dy2 * dy2) + dz2 * dz2;
                    if (r2 < g.h2) cnt++;
                }
            }
    }
    counts[i] = cnt;
    if (ncand) ncand[i] = cand;
}

hipError_t launch_density(const GridDev &g, const float *q, uint32_t nq, int compensated, float *rho, hipStream_t st)
{
    hipLaunchKernelGGL(k_density, dim3((nq + 255) / 256), dim3(256), 0, st, g, q, nq, compensated, rho);
    return hipGetLastError();
}

hipError_t launch_density64(const GridDev &g, const double *q, uint32_t nq, double *rho, hipStream_t st)
{
    hipLaunchKernelGGL(k_density64, dim3((nq + 255) / 256), dim3(256), 0, st, g, q, nq, rho);
    return hipGetLastError();
}

hipError_t launch_neighbor_counts(const GridDev &g, const float *q, uint32_t nq, uint32_t *counts, uint32_t *ncand,
                                  hipStream_t st)
{
    hipLaunchKernelGGL(k_neighbor_counts, dim3((nq + 255) / 256), dim3(256), 0, st, g, q, nq, counts, ncand);
    return hipGetLastError();
}

}  // namespace pm
