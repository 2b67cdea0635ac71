// pm_aniso.hip -- the anisotropic march on CDNA4: PerPixel_Anisotropic
// (RayMarcher.cpp:346-423), the reference's default mode.
//
// Per sample point q:
//   1. the extended neighbours |x_j - q|^2 < hx^2, hx = ParticleRadiusExt
//      (GetNeighborsExt, Dataset.cpp:282-290), at most 8192 (RayMarcher.cpp:14,
//      381), walked on the search grid (cells of h) over the cells within two
//      cells of q: rows dz outer, dy, x, by input index inside a cell;
//   2. WPCA (RayMarcher.cpp:114-254): weights 1 - (|r|/hx)^3 (CubicKernel,
//      Kernel.cpp:110-120), weighted mean and covariance C about it in one
//      pass over that walk (an_cov), a cyclic Jacobi eigen-decomposition of C,
//      the clamp of [YT13] eq. 15 (:232-242) and G = h^-1 R diag(1/Sigma) R^T
//      (:246-247);
//   3. rho = sum over the extended neighbours within h of
//      AnisotropicKernel::W(G, det G, r) (Kernel.cpp:63-83), walked over the
//      trimmed 5x5x5 march-grid stencil around q in the canonical order (the
//      extended walk up to the cap when the list was capped).
// One ray per lane, one 8x8 tile per wave, each lane to the end of its ray
// (march, bisection, FP64 step, normal, outputs); schedule 1 of the configured
// march runs persistent waves with ray refill and a finish pass over the hits
// instead (below).  All float arithmetic is in the order of the CPU oracle
// (oracle/pm_aniso.inc), -ffp-contract=off.
#include "pm_internal.h"
#include "pm_ray.h"

namespace pm {

static constexpr uint32_t kMaxNeighbors = 8192;       // RayMarcher.cpp:14
static constexpr float kLimExt = 4.00390625f;         // 2 cells: 4 (1 + 2^-10)
static constexpr int kJacobiSweeps = 5;

// ---- search-grid walk ----
struct SWalk {
    int32_t c[3];   // q's cell, relative to the grid origin
    float f[3];     // q's offset inside its cell, in cells
};

// False when every cell within R cells of q lies outside the grid.
__device__ __forceinline__ bool an_centre(const GridDev &g, const float q[3], int R, SWalk &w)
{
    const int32_t org[3] = { g.ox, g.oy, g.oz }, dim[3] = { g.nx, g.ny, g.nz };
    bool ok = true;
#pragma unroll
    for (int a = 0; a < 3; a++) {
        const float u = g.inv_h * q[a];
        const int32_t k = q[a] >= 0.0f ? (int32_t)u : (int32_t)u - 1;
        w.f[a] = u - (float)k;
        w.c[a] = k - org[a];
        if (w.c[a] < -R || w.c[a] > dim[a] - 1 + R) ok = false;
    }
    return ok;
}

// Row (dy, dz): false when it is dropped (farther than R cells, or outside the
// grid); else the stored range [b, e) of its kept cells, clamped to the grid.
__device__ __forceinline__ bool an_row(const GridDev &g, const SWalk &w, int R, float lim, int dy, int dz, uint32_t &b,
                                       uint32_t &e)
{
    const float ty = pm_fine_off(dy, w.f[1]), tz = pm_fine_off(dz, w.f[2]);
    const float r2 = ty * ty + tz * tz;
    if (!(r2 < lim)) return false;
    const int32_t y = w.c[1] + dy, z = w.c[2] + dz;
    if (y < 0 || y >= g.ny || z < 0 || z >= g.nz) return false;
    int lo = 0, hi = 0;
    for (int d = -R; d < 0; d++) {
        const float t = pm_fine_off(d, w.f[0]);
        if (t * t + r2 < lim) { lo = d; break; }
    }
    for (int d = R; d > 0; d--) {
        const float t = pm_fine_off(d, w.f[0]);
        if (t * t + r2 < lim) { hi = d; break; }
    }
    int32_t xl = w.c[0] + lo, xh = w.c[0] + hi;
    if (xl < 0) xl = 0;
    if (xh > g.nx - 1) xh = g.nx - 1;
    if (xl > xh) return false;
    const uint32_t row = (uint32_t)g.nx * ((uint32_t)y + (uint32_t)g.ny * (uint32_t)z);
    b = g.cell_start[row + (uint32_t)xl];
    e = g.cell_start[row + (uint32_t)xh + 1];
    return true;
}

// Row r of the (2R+1)^2 rows in walk order (dz = r / (2R+1) - R outer, dy
// inner): its stored range, empty when the row is dropped.
template <int R>
__device__ __forceinline__ void an_row_at(const GridDev &g, const SWalk &w, float lim, int r, uint32_t &b,
                                          uint32_t &e)
{
    const int dz = r / (2 * R + 1) - R, dy = r - (2 * R + 1) * (r / (2 * R + 1)) - R;
    if (!an_row(g, w, R, lim, dy, dz, b, e)) b = e = 0u;
}

// Every candidate of the walk, in order, through f(float4): the next row's
// bounds are loaded while a row is read, and a row's candidates are loaded 4
// at a time (the last group re-reads the row's first candidate in its unused
// slots), so each lane keeps several gathers in flight.  Returns the number
// of candidates.
template <int R, class F>
__device__ __forceinline__ uint32_t an_walk(const GridDev &g, const SWalk &w, float lim, F &&f)
{
    constexpr int NR = (2 * R + 1) * (2 * R + 1);
    uint32_t cnd = 0, nb, ne;
    an_row_at<R>(g, w, lim, 0, nb, ne);
    _Pragma("unroll 1") for (int r = 0; r < NR; r++) {
        const uint32_t b = nb, e = ne;
        if (r + 1 < NR) an_row_at<R>(g, w, lim, r + 1, nb, ne);
        cnd += e - b;
        for (uint32_t i = b; i < e; i += 4) {
            float4 v[4];
#pragma unroll
            for (int u = 0; u < 4; u++) v[u] = g.sorted[i + u < e ? i + u : i];
#pragma unroll
            for (int u = 0; u < 4; u++)
                if (i + u < e) f(v[u]);
        }
    }
    return cnd;
}

// Every candidate of the trimmed 5x5x5 march-grid stencil around q (the
// isotropic density's walk: the particles within h in the canonical order),
// loaded as an_walk loads its rows.  Returns the number of candidates; none
// when the stencil leaves the march grid (no particle lies within h of q).
template <class F>
__device__ __forceinline__ uint32_t an_fine_walk(const GridDev &g, const float qf[3], F &&f)
{
    FineWalk fw;
    if (!pm_fine_begin(g, qf[0], qf[1], qf[2], fw)) return 0;
    uint32_t cnd = 0, nb, ne;
    pm_fine_row_at(g, fw, 0, nb, ne);
    _Pragma("unroll 1") for (int r = 0; r < 25; r++) {
        const uint32_t b = nb, e = ne;
        if (r + 1 < 25) pm_fine_row_at(g, fw, r + 1, nb, ne);
        cnd += e - b;
        for (uint32_t i = b; i < e; i += 4) {
            float4 v[4];
#pragma unroll
            for (int u = 0; u < 4; u++) v[u] = g.fsorted[i + u < e ? i + u : i];
#pragma unroll
            for (int u = 0; u < 4; u++)
                if (i + u < e) f(v[u]);
        }
    }
    return cnd;
}

// ---- WPCA ----
template <class P>
__device__ __forceinline__ typename P::T an_wcubic(typename P::T r2, const P &a)
{
    using T = typename P::T;
    const T d = sqrt(r2);
    if (d >= a.hx) return (T)0;
    const T k = d * a.hxinv;
    return (T)1 - k * k * k;
}

// One Jacobi rotation zeroing A[p][q] (r the third index).
template <class T, int p, int q, int r>
__device__ __forceinline__ void an_jrot(T (&A)[3][3], T (&V)[3][3])
{
    const T apq = A[p][q];
    if (apq == (T)0) return;
    const T theta = (A[q][q] - A[p][p]) / ((T)2 * apq);
    T t = (T)1 / (fabs(theta) + sqrt(theta * theta + (T)1));
    if (theta < (T)0) t = -t;
    const T c = (T)1 / sqrt(t * t + (T)1);
    const T s = t * c;
    const T tp = t * apq;
    A[p][p] = A[p][p] - tp;
    A[q][q] = A[q][q] + tp;
    A[p][q] = A[q][p] = (T)0;
    const T arp = A[r][p], arq = A[r][q];
    const T nrp = c * arp - s * arq, nrq = s * arp + c * arq;
    A[r][p] = A[p][r] = nrp;
    A[r][q] = A[q][r] = nrq;
#pragma unroll
    for (int k = 0; k < 3; k++) {
        const T vkp = V[k][p], vkq = V[k][q];
        V[k][p] = c * vkp - s * vkq;
        V[k][q] = s * vkp + c * vkq;
    }
}

template <class T>
__device__ __forceinline__ T an_max(T x, T y) { return (x < y) ? y : x; }   // std::max

// C (lower triangle c00, c10, c11, c20, c21, c22) of N neighbours -> G (row
// major, symmetric) and det G.
template <class P>
__device__ __forceinline__ void an_gmat(const typename P::T c[6], uint32_t N, const P &a, typename P::T G[9],
                                        typename P::T &detG)
{
    using T = typename P::T;
    T A[3][3] = { { c[0], c[1], c[3] }, { c[1], c[2], c[4] }, { c[3], c[4], c[5] } };
    T V[3][3] = { { (T)1, (T)0, (T)0 }, { (T)0, (T)1, (T)0 }, { (T)0, (T)0, (T)1 } };
    _Pragma("unroll 1") for (int sw = 0; sw < kJacobiSweeps; sw++) {
        an_jrot<T, 0, 1, 2>(A, V);
        an_jrot<T, 0, 2, 1>(A, V);
        an_jrot<T, 1, 2, 0>(A, V);
    }
    T lam[3] = { A[0][0], A[1][1], A[2][2] };
    if (N <= (uint32_t)a.n_eps) {
        lam[0] = lam[1] = lam[2] = a.k_n;
    } else {
        const T mn = an_max(lam[0], an_max(lam[1], lam[2])) / a.k_r;
#pragma unroll
        for (int k = 0; k < 3; k++) lam[k] = an_max(lam[k], mn);
#pragma unroll
        for (int k = 0; k < 3; k++) lam[k] = lam[k] * a.k_s;
    }
    const T d[3] = { (T)1 / lam[0], (T)1 / lam[1], (T)1 / lam[2] };
#pragma unroll
    for (int r = 0; r < 3; r++)
#pragma unroll
        for (int s = r; s < 3; s++) {
            const T t0 = ((a.hinv * V[r][0]) * d[0]) * V[s][0];
            const T t1 = ((a.hinv * V[r][1]) * d[1]) * V[s][1];
            const T t2 = ((a.hinv * V[r][2]) * d[2]) * V[s][2];
            G[3 * r + s] = G[3 * s + r] = (t0 + t1) + t2;
        }
    // glm::determinant (mat3), m[c][r] = G[r][c]
    const T A0 = G[0] * (G[4] * G[8] - G[5] * G[7]);
    const T A1 = G[1] * (G[3] * G[8] - G[5] * G[6]);
    const T A2 = G[2] * (G[3] * G[7] - G[4] * G[6]);
    detG = (A0 - A1) + A2;
}

// WPCA's weighted mean and covariance (RayMarcher.cpp:138-169) in one walk.
// The weights come from r_j = x_j - q; the moments are taken about the first
// neighbour found, s: with u_j = x_j - s, W = sum w and mu = sum w u / W,
//     C = sum w u u^T / W - mu mu^T,
// which is the reference's second pass sum w (x - m)(x - m)^T / W about the
// weighted mean m = s + mu in exact arithmetic.  s lies among the neighbours,
// so u is of the size of their spread and the subtraction cancels little (a
// single neighbour gives C = 0 exactly, as the two-pass form does).  C (lower
// triangle) in c; returns the extended-neighbour count (before the cap) and
// adds the walk's candidates to cand.
template <class P>
__device__ __forceinline__ uint32_t an_cov(const GridDev &g, const P &a, const float qf[3], const typename P::T q[3],
                                           typename P::T c[6], uint32_t &cand)
{
    using T = typename P::T;
    SWalk w;
    if (!an_centre(g, qf, 2, w)) return 0;
    uint32_t nx = 0;
    T sx = (T)0, sy = (T)0, sz = (T)0;
    T ws = (T)0, s0 = (T)0, s1 = (T)0, s2 = (T)0;
    T c0 = (T)0, c1 = (T)0, c2 = (T)0, c3 = (T)0, c4 = (T)0, c5 = (T)0;
    cand += an_walk<2>(g, w, kLimExt, [&](const float4 p) {
        const T rx = (T)p.x - q[0], ry = (T)p.y - q[1], rz = (T)p.z - q[2];
        const T r2 = (rx * rx + ry * ry) + rz * rz;
        if (r2 < a.R2x) {
            if (nx == 0) { sx = (T)p.x; sy = (T)p.y; sz = (T)p.z; }
            if (nx < kMaxNeighbors) {
                const T wt = an_wcubic(r2, a);
                const T ux = (T)p.x - sx, uy = (T)p.y - sy, uz = (T)p.z - sz;
                const T w0 = wt * ux, w1 = wt * uy, w2 = wt * uz;
                ws += wt;
                s0 += w0; s1 += w1; s2 += w2;
                c0 += w0 * ux;
                c1 += w1 * ux;
                c2 += w1 * uy;
                c3 += w2 * ux;
                c4 += w2 * uy;
                c5 += w2 * uz;
            }
            nx++;
        }
    });
    if (nx == 0) return 0;
    const T inv = (T)1 / ws;
    const T m0 = s0 * inv, m1 = s1 * inv, m2 = s2 * inv;
    c[0] = c0 * inv - m0 * m0; c[1] = c1 * inv - m1 * m0; c[2] = c2 * inv - m1 * m1;
    c[3] = c3 * inv - m2 * m0; c[4] = c4 * inv - m2 * m1; c[5] = c5 * inv - m2 * m2;
    return nx;
}

// AnisotropicKernel::W (Kernel.cpp:63-83), r' = G r as glm's mat3 * vec3.
template <class P>
__device__ __forceinline__ typename P::T an_W(const typename P::T G[9], typename P::T detG, typename P::T rx,
                                              typename P::T ry, typename P::T rz, const P &a)
{
    using T = typename P::T;
    const T g0 = (G[0] * rx + G[1] * ry) + G[2] * rz;
    const T g1 = (G[3] * rx + G[4] * ry) + G[5] * rz;
    const T g2 = (G[6] * rx + G[7] * ry) + G[8] * rz;
    const T q2 = (g0 * g0 + g1 * g1) + g2 * g2;
    if (q2 >= a.h2) return (T)0;
    const T q = sqrt(q2) * a.hinv;
    const T sd = a.sig * detG;
    if (q >= (T)0.5) {
        const T qq = (T)1 - q;
        return sd * ((T)2 * qq * qq * qq);
    }
    return sd * ((T)6 * (q * q * q - q * q) + (T)1);
}

// AnisotropicKernel::gradW (Kernel.cpp:85-107): gradQ = normalize(r')/(|r'| h).
__device__ __forceinline__ void an_gradW(const float G[9], float detG, float rx, float ry, float rz, const AnParF &a,
                                         float &o0, float &o1, float &o2)
{
    const float g0 = (G[0] * rx + G[1] * ry) + G[2] * rz;
    const float g1 = (G[3] * rx + G[4] * ry) + G[5] * rz;
    const float g2 = (G[6] * rx + G[7] * ry) + G[8] * rz;
    const float rn = (g0 * g0 + g1 * g1) + g2 * g2;
    o0 = o1 = o2 = 0.0f;
    if (rn >= a.h2) return;
    const float rl = sqrtf(rn);
    const float q = rl * a.hinv;
    const float inv = 1.0f / sqrtf(rn);
    const float den = rl * a.h;
    const float q0 = (g0 * inv) / den, q1 = (g1 * inv) / den, q2 = (g2 * inv) / den;
    if (q >= 0.5f) {
        const float qq = 1.0f - q;
        const float f = 6.0f * qq * qq;
        const float ns = -a.sig * detG;
        o0 = (ns * q0) * f; o1 = (ns * q1) * f; o2 = (ns * q2) * f;
        return;
    }
    const float f = 6.0f * (3.0f * q * q - 2.0f * q);
    const float sd = a.sig * detG;
    o0 = (sd * q0) * f; o1 = (sd * q1) * f; o2 = (sd * q2) * f;
}

template <class T>
__device__ __forceinline__ void an_neu_add(T &s, T &c, T w)
{
    const T t = s + w;
    if (fabs(s) >= fabs(w)) c += (s - t) + w;
    else c += (w - t) + s;
    s = t;
}

// Density at q (compensated: Neumaier sum), or with grad != NULL the sum of
// gradW over the same neighbours (FP32).  qf is the FP32 point that picks the
// cells.  The extended neighbours within h are all the particles within h
// unless the list was capped: then they are taken from the extended walk up to
// the cap, else from the trimmed march-grid stencil in the canonical order
// (about 40 candidates per sample instead of the 27 search cells' 200).
template <class P>
__device__ __forceinline__ typename P::T an_sum(const GridDev &g, const P &a, const float qf[3],
                                                const typename P::T q[3], bool comp, uint32_t &cand,
                                                float *grad = nullptr)
{
    using T = typename P::T;
    T c[6], G[9], detG;
    if (grad) grad[0] = grad[1] = grad[2] = 0.0f;
    const uint32_t nx = an_cov(g, a, qf, q, c, cand);
    if (nx == 0) return (T)0;
    an_gmat(c, nx < kMaxNeighbors ? nx : kMaxNeighbors, a, G, detG);
    const bool capped = nx > kMaxNeighbors;
    SWalk w;
    T s = (T)0, cc = (T)0;
    uint32_t n = 0, cnd;
    auto term = [&](const float4 p) {
        const T rx = (T)p.x - q[0], ry = (T)p.y - q[1], rz = (T)p.z - q[2];
        const T r2 = (rx * rx + ry * ry) + rz * rz;
        if (capped) {
            if (!(r2 < a.R2x)) return;
            if (n++ >= kMaxNeighbors) return;
        }
        if (!(r2 < a.h2)) return;
        if constexpr (sizeof(T) == 4) {
            if (grad) {
                float o0, o1, o2;
                an_gradW(G, detG, rx, ry, rz, a, o0, o1, o2);
                grad[0] += o0; grad[1] += o1; grad[2] += o2;
                return;
            }
        }
        const T wv = an_W(G, detG, rx, ry, rz, a);
        if (comp) an_neu_add(s, cc, wv);
        else s += wv;
    };
    if (capped) {
        an_centre(g, qf, 2, w);
        cnd = an_walk<2>(g, w, kLimExt, term);
    } else {
        cnd = an_fine_walk(g, qf, term);
    }
    cand += cnd;
    return comp ? s + cc : s;
}

__device__ __forceinline__ float an_density32(const MarchLaunch &L, float x, float y, float z, bool comp,
                                              uint32_t &cand)
{
    const float q[3] = { x, y, z };
    return an_sum(L.g, L.an.f, q, q, comp, cand);
}

__device__ __forceinline__ double an_density64(const GridDev &g, const AnParD &a, double x, double y, double z)
{
    const double q[3] = { x, y, z };
    const float qf[3] = { (float)x, (float)y, (float)z };
    uint32_t cand = 0;
    return an_sum(g, a, qf, q, false, cand);
}

// Normal at the hit point p: normalize(sum gradW) (RayMarcher.cpp:402-409) or
// normalize(-grad rho) by central differences (six samples, each with its own
// G), counted as density evaluations.
__device__ __forceinline__ void an_normal(const MarchLaunch &L, const float p[3], Counts &cn, float &nx, float &ny,
                                          float &nz)
{
    const pm_settings &s = L.s;
    if (s.normal_mode == PM_NORMAL_GRADW) {
        float gs[3];
        uint32_t cand = 0;
        an_sum(L.g, L.an.f, p, p, false, cand, gs);
        const float dd = (gs[0] * gs[0] + gs[1] * gs[1]) + gs[2] * gs[2];
        const float inv = 1.0f / sqrtf(dd);
        nx = gs[0] * inv; ny = gs[1] * inv; nz = gs[2] * inv;
        return;
    }
    const bool comp = s.compensated != 0;
    float gv[3];
    _Pragma("unroll 1") for (int a = 0; a < 3; a++) {
        float xp[3] = { p[0], p[1], p[2] }, xm[3] = { p[0], p[1], p[2] };
        xp[a] = p[a] + s.normal_eps;
        xm[a] = p[a] - s.normal_eps;
        uint32_t cand = 0;
        const float rp = an_density32(L, xp[0], xp[1], xp[2], comp, cand);
        const float rm = an_density32(L, xm[0], xm[1], xm[2], comp, cand);
        cn.samples += 2;
        cn.cand += cand;
        gv[a] = rp - rm;
    }
    const float dd = (gv[0] * gv[0] + gv[1] * gv[1]) + gv[2] * gv[2];
    nx = ny = nz = 0.0f;
    if (dd > 0.0f) {
        const float inv = 1.0f / sqrtf(dd);
        nx = -gv[0] * inv; ny = -gv[1] * inv; nz = -gv[2] * inv;
    }
}

// The configured march of one pixel with the anisotropic density: AABB entry,
// fixed step, first sample at or above iso, bisection (last step in FP64),
// normal, outputs.
__device__ __forceinline__ void an_pixel_config(const MarchLaunch &L, uint32_t px, uint32_t py, Counts &cn)
{
    const pm_settings &s = L.s;
    const GridDev &g = L.g;
    const bool comp = s.compensated != 0;
    const uint32_t idx = py * L.W + px;
    cn.rays++;
    float d[3];
    pm_ray_dir(L, px, py, d);
    const float o[3] = { L.cam[0], L.cam[1], L.cam[2] };
    RayState r{};
    if (!pm_slab(o, d, g.bmin, g.bmax, r.tn, r.tf)) {
        write_miss(L, idx);
        return;
    }
    cn.in_aabb++;
    r.d0 = d[0]; r.d1 = d[1]; r.d2 = d[2];
    r.id0 = 1.0f / d[0]; r.id1 = 1.0f / d[1]; r.id2 = 1.0f / d[2];
    r.k = 0;
    const float step = s.step;
    int hitk = -1;
    for (;;) {
        if (s.max_steps > 0 && r.k >= s.max_steps) break;
        const float t = r.tn + (float)r.k * step;
        if (t > r.tf) break;
        const float x = o[0] + t * d[0], y = o[1] + t * d[1], z = o[2] + t * d[2];
        if (s.skip_mode == PM_SKIP_EMPTY || s.skip_mode == PM_SKIP_BOUND) {
            const int kn = s.skip_mode == PM_SKIP_EMPTY ? pm_skip_empty(g, o[0], o[1], o[2], r, x, y, z, step)
                                                        : pm_skip_bound(g, o[0], o[1], o[2], r, x, y, z, L.inv_step);
            if (kn != r.k) { r.k = kn; continue; }
        }
        uint32_t cand = 0;
        const float rho = an_density32(L, x, y, z, comp, cand);
        cn.samples++; cn.march++; cn.cand += cand;
        if (rho >= s.iso) { hitk = r.k; break; }
        r.k++;
    }
    if (hitk < 0) {
        write_miss(L, idx);
        return;
    }
    cn.hits++;
    float xs[3], th;
    if (hitk == 0) {
        th = r.tn;
        xs[0] = o[0] + th * d[0]; xs[1] = o[1] + th * d[1]; xs[2] = o[2] + th * d[2];
    } else {
        float lo = r.tn + (float)(hitk - 1) * step, hi = r.tn + (float)hitk * step;
        const bool fp64 = s.final_fp64 && s.bisect_iters > 0;
        const int n32 = fp64 ? s.bisect_iters - 1 : s.bisect_iters;
        for (int it = 0; it < n32; it++) {
            const float tm = (lo + hi) * 0.5f;
            uint32_t cand = 0;
            const float rho = an_density32(L, o[0] + tm * d[0], o[1] + tm * d[1], o[2] + tm * d[2], comp, cand);
            cn.samples++; cn.cand += cand;
            if (rho >= s.iso) hi = tm; else lo = tm;
        }
        if (fp64) {
            double dlo = lo, dhi = hi;
            const double tm = (dlo + dhi) * 0.5;
            const double rho = an_density64(g, L.an.d, (double)o[0] + tm * (double)d[0],
                                            (double)o[1] + tm * (double)d[1], (double)o[2] + tm * (double)d[2]);
            cn.samples++;
            if (rho >= (double)s.iso) dhi = tm; else dlo = tm;
            const double tt = (dlo + dhi) * 0.5;
            xs[0] = (float)((double)o[0] + tt * (double)d[0]);
            xs[1] = (float)((double)o[1] + tt * (double)d[1]);
            xs[2] = (float)((double)o[2] + tt * (double)d[2]);
            th = (float)tt;
        } else {
            th = (lo + hi) * 0.5f;
            xs[0] = o[0] + th * d[0]; xs[1] = o[1] + th * d[1]; xs[2] = o[2] + th * d[2];
        }
    }
    float nx, ny, nz;
    an_normal(L, xs, cn, nx, ny, nz);
    L.pos[idx] = make_float4(xs[0], xs[1], xs[2], 1.0f);
    L.nrm[idx] = make_float4(nx, ny, nz, 1.0f);
    if (L.t) L.t[idx] = th;
}

// PerPixel_Anisotropic itself: start on the depth-buffer surface, fixed step
// vector, MaxSteps samples (DensityGrid skips not counted), first sample with
// rho >= iso, normal = normalize(sum gradW) with that sample's G.
__device__ __forceinline__ void an_pixel_reference(const MarchLaunch &L, uint32_t px, uint32_t py, Counts &cn)
{
    const pm_settings &s = L.s;
    const GridDev &g = L.g;
    const bool comp = s.compensated != 0;
    const uint32_t idx = py * L.W + px;
    cn.rays++;
    const float z = L.depth[idx];
    float4 P = make_float4(0.f, 0.f, 0.f, 0.f), N = make_float4(0.f, 0.f, 0.f, 0.f);
    float tout = -1.0f;
    if (z != 1.0f) {                                                       // :354
        cn.in_aabb++;
        const float cx = (float)px * (2.0f / (float)L.W) - 1.0f;
        const float cy = (float)py * (2.0f / (float)L.H) - 1.0f;
        float p0, p1, p2, d0, d1, d2;
        pm_unproject(L.M, cx, cy, z, p0, p1, p2);                          // :356-362
        pm_normalize(p0 - L.cam[0], p1 - L.cam[1], p2 - L.cam[2], d0, d1, d2);
        const float st0 = s.step * d0, st1 = s.step * d1, st2 = s.step * d2;
        for (int i = 0; i < s.max_steps; i++) {
            p0 += st0; p1 += st1; p2 += st2;                               // :367
            if (s.skip_mode == PM_SKIP_DG_BOX || s.skip_mode == PM_SKIP_DG_CONS) {   // :370-376
                float nmin[3], nmax[3];
                while (pm_dg_query(g, p0, p1, p2, nmin, nmax)) {
                    const float pin[3] = { p0, p1, p2 }, stv[3] = { st0, st1, st2 };
                    float q[3];
                    pm_intersect_aabb(pin, stv, nmin, nmax, q);
                    p0 = q[0] + st0; p1 = q[1] + st1; p2 = q[2] + st2;
                }
            }
            uint32_t cand = 0;
            const float rho = an_density32(L, p0, p1, p2, comp, cand);
            cn.samples++; cn.march++; cn.cand += cand;
            if (rho >= s.iso) {                                            // :402
                const float p[3] = { p0, p1, p2 };
                float nx, ny, nz;
                an_normal(L, p, cn, nx, ny, nz);
                P = make_float4(p0, p1, p2, 1.0f);
                N = make_float4(nx, ny, nz, 1.0f);
                const float v0 = p0 - L.cam[0], v1 = p1 - L.cam[1], v2 = p2 - L.cam[2];
                tout = sqrtf((v0 * v0 + v1 * v1) + v2 * v2);
                cn.hits++;
                break;
            }
        }
    }
    L.pos[idx] = P;
    L.nrm[idx] = N;
    if (L.t) L.t[idx] = tout;
}

// Registers capped for five waves per SIMD (96 VGPRs, a few spills): the
// walks wait on gathers 37 % of the time, and the fifth wave hides more of it
// (14.7 ms against 15.9 ms at four waves on the 1080p scaled dam break).
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(5, 8))) void k_march_aniso(MarchLaunch L)
{
    uint32_t px, py;
    if (!pm_tile(L.W, L.H, px, py)) return;
    Counts cn;
    if (px < L.W && py < L.H) {
        if (L.s.mode == PM_MODE_REFERENCE) an_pixel_reference(L, px, py, cn);
        else an_pixel_config(L, px, py, cn);
    }
    flush_counters(L.counters, cn.rays, cn.in_aabb, cn.hits, cn.samples, cn.cand, cn.march);
}

// The march over the active tiles only (the tile pre-pass has written every
// pixel of the other tiles as a miss; in reference mode the active tiles are
// those with a drawn sprite depth), one tile per wave: every
// wave that starts has rays to march, and the heavy tiles are not dispatched
// between thousands of empty ones.  Workgroups go to the 8 XCDs round-robin
// (workgroup b to XCD b mod 8), so XCD x takes the x-th eighth of the active
// list: all XCDs share the work and neighbouring tiles share an L2.  The grid
// has a multiple of 8 workgroups and at least T / 32 per XCD.
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(5, 8))) void k_march_aniso_active(MarchLaunch L)
{
    const uint32_t T = ((L.W + 7) / 8) * ((L.H + 7) / 8);
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint32_t *order = L.tile_buf + 2 * (T + 1);
    const uint32_t count = order[T], chunk = (count + 7) / 8;
    const uint32_t local = (blockIdx.x >> 3) * 4 + wave, i = (blockIdx.x & 7u) * chunk + local;
    if (local >= chunk || i >= count) return;   // whole wave
    const uint32_t tile = order[i], tiles_x = (L.W + 7) / 8;
    const uint32_t px = (tile % tiles_x) * 8 + (lane & 7u), py = (tile / tiles_x) * 8 + (lane >> 3);
    Counts cn;
    if (px < L.W && py < L.H) {
        if (L.s.mode == PM_MODE_REFERENCE) an_pixel_reference(L, px, py, cn);
        else an_pixel_config(L, px, py, cn);
    }
    flush_counters(L.counters, cn.rays, cn.in_aabb, cn.hits, cn.samples, cn.cand, cn.march);
}

// ---- configured march: persistent waves, then a finish pass over the hits ----
//
// The march kernel keeps every lane on the same work: each loop iteration
// evaluates one FP32 density per busy lane (march sample or bisection step);
// a lane whose ray ends takes the next ray of the wave's tile, and the wave
// takes the next tile from one of 8 per-XCD band counters (stealing from the
// other bands at the end).  A ray's march ends at its hit: (1 | fp64 << 1,
// lo, hi, pixel) -- t itself, or the bracket of the FP64 bisection step -- is
// staged in LDS and appended to a hit list 64 entries at a time (one atomic
// per 64 hits).  k_finish_aniso then takes the list 64 hits per wave: the FP64
// step, the six central-difference samples (or sum gradW) and the outputs.

static constexpr uint32_t kListCtr = 16;   // band_ctr[16]: hit-list length

__device__ __forceinline__ uint32_t an_band_lo(uint32_t T, uint32_t b) { return (uint32_t)(((uint64_t)T * b) >> 3); }

__global__ __launch_bounds__(256) void k_march_aniso_persist(MarchLaunch L, uint32_t *__restrict__ band_ctr,
                                                             uint4 *__restrict__ list)
{
    __shared__ uint4 stage[4][128];
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const pm_settings &s = L.s;
    const GridDev &g = L.g;
    const bool comp = s.compensated != 0;
    const bool fp64 = s.final_fp64 && s.bisect_iters > 0;
    const int n32 = fp64 ? s.bisect_iters - 1 : s.bisect_iters;
    const float o0 = L.cam[0], o1 = L.cam[1], o2 = L.cam[2];
    const uint32_t tiles_x = (L.W + 7) / 8, tiles_y = (L.H + 7) / 8, T = tiles_x * tiles_y;
    const uint32_t xcd = blockIdx.x & 7u;
    const unsigned long long lt_mask = (1ull << lane) - 1ull;

    Counts cn;
    RayState r{};
    r.phase = PH_DONE;
    uint32_t tile = 0, steal = 0;
    int qpos = 64, qn = 0;
    bool exhausted = false;

    while (true) {
        // refill idle lanes from the wave's tile
        while (!exhausted) {
            const unsigned long long needy = __ballot(r.phase == PH_DONE);
            if (needy == 0) break;
            if (qpos >= 64) {
                uint32_t t = 0xffffffffu;
                if (lane == 0) {
                    while (steal < 8) {
                        const uint32_t b = (xcd + steal) & 7u;
                        const uint32_t lo = an_band_lo(T, b), hi = an_band_lo(T, b + 1);
                        const uint32_t v = atomicAdd(&band_ctr[b], 1u);
                        if (v < hi - lo) { t = lo + v; break; }
                        steal++;
                    }
                }
                t = __shfl(t, 0, 64);
                steal = __shfl(steal, 0, 64);
                if (t == 0xffffffffu) { exhausted = true; break; }
                tile = t;
                qpos = 0;
            }
            const int rank = __popcll(needy & lt_mask), take = __popcll(needy);
            if (r.phase == PH_DONE && qpos + rank < 64) {
                const uint32_t j = (uint32_t)(qpos + rank);
                const uint32_t px = (tile % tiles_x) * 8 + (j & 7u), py = (tile / tiles_x) * 8 + (j >> 3);
                if (px < L.W && py < L.H) {
                    cn.rays++;
                    r.idx = py * L.W + px;
                    float d[3];
                    pm_ray_dir(L, px, py, d);
                    r.d0 = d[0]; r.d1 = d[1]; r.d2 = d[2];
                    r.id0 = 1.0f / d[0]; r.id1 = 1.0f / d[1]; r.id2 = 1.0f / d[2];
                    const float o[3] = { o0, o1, o2 };
                    if (pm_slab(o, d, g.bmin, g.bmax, r.tn, r.tf)) {
                        r.phase = PH_MARCH;
                        r.k = 0;
                        cn.in_aabb++;
                    } else {
                        write_miss(L, r.idx);
                    }
                }
            }
            qpos = qpos + take < 64 ? qpos + take : 64;
        }
        if (__ballot(r.phase != PH_DONE) == 0) break;

        // this lane's next sample: march (after at most 8 empty-cell skips) or bisection
        bool want = false;
        float qx = 0.f, qy = 0.f, qz = 0.f;
        if (r.phase == PH_MARCH) {
            for (int hop = 0;; hop++) {
                const float t = r.tn + (float)r.k * s.step;
                if (t > r.tf || (s.max_steps > 0 && r.k >= s.max_steps)) {
                    write_miss(L, r.idx);
                    r.phase = PH_DONE;
                    break;
                }
                qx = o0 + t * r.d0; qy = o1 + t * r.d1; qz = o2 + t * r.d2;
                int kn;
                if (s.skip_mode == PM_SKIP_BOUND) kn = pm_skip_bound(g, o0, o1, o2, r, qx, qy, qz, L.inv_step);
                else if (s.skip_mode == PM_SKIP_EMPTY) kn = pm_skip_empty(g, o0, o1, o2, r, qx, qy, qz, s.step);
                else { want = true; break; }
                if (kn == r.k) { want = true; break; }
                r.k = kn;
                if (hop == 7) break;
            }
        } else if (r.phase == PH_BISECT) {
            const float tm = (r.lo + r.hi) * 0.5f;
            qx = o0 + tm * r.d0; qy = o1 + tm * r.d1; qz = o2 + tm * r.d2;
            want = true;
        }
        float rho = 0.0f;
        uint32_t cand = 0;
        if (want) rho = an_density32(L, qx, qy, qz, comp, cand);

        bool emit = false;
        uint4 rec = make_uint4(0u, 0u, 0u, 0u);
        if (want) {
            cn.samples++;
            cn.cand += cand;
            if (r.phase == PH_MARCH) {
                cn.march++;
                if (rho >= s.iso) {
                    cn.hits++;
                    if (r.k == 0) {
                        emit = true;
                        rec = make_uint4(1u, __float_as_uint(r.tn), __float_as_uint(r.tn), r.idx);
                    } else {
                        r.lo = r.tn + (float)(r.k - 1) * s.step;
                        r.hi = r.tn + (float)r.k * s.step;
                        r.it = 0;
                        if (n32 > 0) {
                            r.phase = PH_BISECT;
                        } else {
                            emit = true;
                            const float tt = (r.lo + r.hi) * 0.5f;
                            rec = fp64 ? make_uint4(3u, __float_as_uint(r.lo), __float_as_uint(r.hi), r.idx)
                                       : make_uint4(1u, __float_as_uint(tt), __float_as_uint(tt), r.idx);
                        }
                    }
                } else {
                    r.k++;
                }
            } else {
                const float tm = (r.lo + r.hi) * 0.5f;
                if (rho >= s.iso) r.hi = tm; else r.lo = tm;
                r.it++;
                if (r.it >= n32) {
                    emit = true;
                    const float tt = (r.lo + r.hi) * 0.5f;
                    rec = fp64 ? make_uint4(3u, __float_as_uint(r.lo), __float_as_uint(r.hi), r.idx)
                               : make_uint4(1u, __float_as_uint(tt), __float_as_uint(tt), r.idx);
                }
            }
            if (emit) r.phase = PH_DONE;
        }

        // stage the hits; append 64 of them to the list at a time
        const unsigned long long em = __ballot(emit);
        if (em) {
            if (emit) stage[wave][qn + __popcll(em & lt_mask)] = rec;
            qn += __popcll(em);
            if (qn >= 64) {
                wave_lds_sync();
                uint32_t base = 0;
                if (lane == 0) base = atomicAdd(&band_ctr[kListCtr], 64u);
                base = __shfl(base, 0, 64);
                list[base + lane] = stage[wave][lane];
                const int rem = qn - 64;
                uint4 v = make_uint4(0u, 0u, 0u, 0u);
                if ((int)lane < rem) v = stage[wave][64 + lane];
                wave_lds_sync();
                if ((int)lane < rem) stage[wave][lane] = v;
                qn = rem;
            }
        }
    }
    if (qn > 0) {
        wave_lds_sync();
        uint32_t base = 0;
        if (lane == 0) base = atomicAdd(&band_ctr[kListCtr], (uint32_t)qn);
        base = __shfl(base, 0, 64);
        if ((int)lane < qn) list[base + lane] = stage[wave][lane];
    }
    flush_counters(L.counters, cn.rays, cn.in_aabb, cn.hits, cn.samples, cn.cand, cn.march);
}

// One hit of the list: the FP64 bisection step (flag 2), the normal and the outputs.
__device__ __forceinline__ void an_finish_hit(const MarchLaunch &L, uint4 rec, Counts &cn)
{
    const pm_settings &s = L.s;
    const uint32_t idx = rec.w;
    const uint32_t px = idx % L.W, py = idx / L.W;
    const float lo = __uint_as_float(rec.y), hi = __uint_as_float(rec.z);
    float d[3];
    pm_ray_dir(L, px, py, d);
    const float o[3] = { L.cam[0], L.cam[1], L.cam[2] };
    float xs[3], th;
    if (rec.x & 2u) {
        double dlo = lo, dhi = hi;
        const double tm = (dlo + dhi) * 0.5;
        const double rho = an_density64(L.g, L.an.d, (double)o[0] + tm * (double)d[0],
                                        (double)o[1] + tm * (double)d[1], (double)o[2] + tm * (double)d[2]);
        cn.samples++;
        if (rho >= (double)s.iso) dhi = tm; else dlo = tm;
        const double tt = (dlo + dhi) * 0.5;
        xs[0] = (float)((double)o[0] + tt * (double)d[0]);
        xs[1] = (float)((double)o[1] + tt * (double)d[1]);
        xs[2] = (float)((double)o[2] + tt * (double)d[2]);
        th = (float)tt;
    } else {
        th = lo;
        xs[0] = o[0] + th * d[0]; xs[1] = o[1] + th * d[1]; xs[2] = o[2] + th * d[2];
    }
    float nx, ny, nz;
    an_normal(L, xs, cn, nx, ny, nz);
    L.pos[idx] = make_float4(xs[0], xs[1], xs[2], 1.0f);
    L.nrm[idx] = make_float4(nx, ny, nz, 1.0f);
    if (L.t) L.t[idx] = th;
}

__global__ __launch_bounds__(256) void k_finish_aniso(MarchLaunch L, const uint32_t *__restrict__ band_ctr,
                                                      const uint4 *__restrict__ list)
{
    const uint32_t n = band_ctr[kListCtr];
    Counts cn;
    for (uint32_t base = blockIdx.x * 256u + (threadIdx.x & ~63u); base < n; base += gridDim.x * 256u) {
        const uint32_t i = base + (threadIdx.x & 63u);
        if (i < n) an_finish_hit(L, list[i], cn);
    }
    flush_counters(L.counters, cn.rays, cn.in_aabb, cn.hits, cn.samples, cn.cand, cn.march);
}

__global__ void k_zero_band(uint32_t *p)
{
    if (threadIdx.x < 32) p[threadIdx.x] = 0u;
}

hipError_t launch_march_aniso(const MarchLaunch &L, hipStream_t st)
{
    const uint32_t tiles = ((L.W + 7) / 8) * ((L.H + 7) / 8);
    // one ray per lane to its end (neighbouring rays of a tile sample
    // neighbouring points, so the wave's gathers share cache lines): over the
    // active tiles after the tile pre-pass (schedule 0), or over all tiles (no
    // tile buffer); or (schedule 1, configured march) persistent waves with
    // ray refill and the hit-list finish
    if (L.s.schedule == 0 && L.tile_buf) {
        hipError_t e = launch_tile_prepass(L, st);
        if (e != hipSuccess) return e;
        // 8 ceil(ceil(T / 8) / 4) workgroups: enough waves for every XCD's share
        const uint32_t per_xcd = ((tiles + 7) / 8 + 3) / 4;
        hipLaunchKernelGGL(k_march_aniso_active, dim3(8 * per_xcd), dim3(256), 0, st, L);
        return hipGetLastError();
    }
    if (L.s.mode == PM_MODE_REFERENCE || L.s.schedule != 1 || !L.hit_rec) {
        hipLaunchKernelGGL(k_march_aniso, dim3((tiles + 3) / 4), dim3(256), 0, st, L);
        return hipGetLastError();
    }
    hipLaunchKernelGGL(k_zero_band, dim3(1), dim3(64), 0, st, L.band_ctr);
    uint32_t blocks = 1024, need = (tiles + 3) / 4;
    if (blocks > need) blocks = (need + 7) & ~7u;
    hipLaunchKernelGGL(k_march_aniso_persist, dim3(blocks), dim3(256), 0, st, L, L.band_ctr, L.hit_rec);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_finish_aniso, dim3(2048), dim3(256), 0, st, L, (const uint32_t *)L.band_ctr,
                       (const uint4 *)L.hit_rec);
    return hipGetLastError();
}

// ---- point queries used by the parity tests ----

__global__ __launch_bounds__(256) void k_density_aniso(GridDev g, AnParF a, const float *__restrict__ q, uint32_t nq,
                                                       int comp, float *__restrict__ rho)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nq) return;
    const float p[3] = { q[3 * i], q[3 * i + 1], q[3 * i + 2] };
    uint32_t cand = 0;
    rho[i] = an_sum(g, a, p, p, comp != 0, cand);
}

__global__ __launch_bounds__(256) void k_density_aniso64(GridDev g, AnParD a, const double *__restrict__ q,
                                                         uint32_t nq, double *__restrict__ rho)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nq) return;
    rho[i] = an_density64(g, a, q[3 * i], q[3 * i + 1], q[3 * i + 2]);
}

__global__ __launch_bounds__(256) void k_aniso_matrix(GridDev g, AnParF a, const float *__restrict__ q, uint32_t nq,
                                                      float *__restrict__ G, float *__restrict__ det,
                                                      uint32_t *__restrict__ next)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nq) return;
    const float p[3] = { q[3 * i], q[3 * i + 1], q[3 * i + 2] };
    float c[6], Gm[9], dt = 0.0f;
    uint32_t cand = 0;
    const uint32_t nx = an_cov(g, a, p, p, c, cand);
    if (nx != 0) an_gmat(c, nx < kMaxNeighbors ? nx : kMaxNeighbors, a, Gm, dt);
    else
        for (int k = 0; k < 9; k++) Gm[k] = 0.0f;
    for (int k = 0; k < 9; k++) G[9 * i + k] = Gm[k];
    det[i] = dt;
    next[i] = nx;
}

hipError_t launch_density_aniso(const GridDev &g, const AnParF &a, const float *q, uint32_t nq, int compensated,
                                float *rho, hipStream_t st)
{
    hipLaunchKernelGGL(k_density_aniso, dim3((nq + 255) / 256), dim3(256), 0, st, g, a, q, nq, compensated, rho);
    return hipGetLastError();
}

hipError_t launch_density_aniso64(const GridDev &g, const AnParD &a, const double *q, uint32_t nq, double *rho,
                                  hipStream_t st)
{
    hipLaunchKernelGGL(k_density_aniso64, dim3((nq + 255) / 256), dim3(256), 0, st, g, a, q, nq, rho);
    return hipGetLastError();
}

hipError_t launch_aniso_matrix(const GridDev &g, const AnParF &a, const float *q, uint32_t nq, float *G, float *det,
                               uint32_t *next, hipStream_t st)
{
    hipLaunchKernelGGL(k_aniso_matrix, dim3((nq + 255) / 256), dim3(256), 0, st, g, a, q, nq, G, det, next);
    return hipGetLastError();
}

}  // namespace pm
