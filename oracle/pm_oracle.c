/*
 * pm_oracle.c -- CPU oracle for the particle-fluid iso-surface ray march.
 *
 * TEST INFRASTRUCTURE ONLY: used by tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py as the checker; the product path never links it.
 *
 * Build: gcc -O2 -std=c11 -ffp-contract=off -fPIC -shared -pthread (oracle/Makefile).
 * -ffp-contract=off keeps every a*b+c as two rounded operations, which is what
 * the HIP kernels do as well, so FP32 results can match bit for bit.
 *
 * Parity pins (all in tests/): closed-form kernel values and the perfect-lattice
 * density sum (A5), brute-force neighbour lists (A4), closed-form ray geometry
 * (A1/A2), closed-form single-particle roots (A6), radial/planar normals (A7),
 * and the FP64 march as ground truth.  CompactNSearch's own return order and
 * find_neighbors_box are not available anywhere here: parity against those two
 * is unpinned, and the oracle fixes its own canonical order (below).
 */
#define _GNU_SOURCE
#include "pm_oracle.h"

#include <math.h>
#include <pthread.h>
#include <unistd.h>
#include <stdatomic.h>
#include <stdlib.h>
#include <string.h>

#define OR_PI_F 3.14159265358979323846f
#define OR_PI_D 3.14159265358979323846
#define OR_ZERO_ISO 0x1p-100f   /* bound below it: no particle within h of the sub-box */

/* ------------------------------------------------------------------------- */
/* SPH cubic spline kernel -- Kernel.cpp:8-14 (constants), :16-32 (W),        */
/* :34-52 (gradW).  sigma = 8 / (pi h^3), evaluated as ((pi*h)*h)*h.          */
/* ------------------------------------------------------------------------- */

float or_sigma_f(float h) { return 8.0f / (OR_PI_F * h * h * h); }

static inline float W_f(float r2, float h2, float hinv, float sig)
{
    if (r2 >= h2) return 0.0f;                  /* Kernel.cpp:20-21 */
    float q = sqrtf(r2) * hinv;                 /* :23 */
    if (q >= 0.5f) {                            /* :25-29 */
        float qq = 1.0f - q;
        return sig * (2.0f * qq * qq * qq);
    }
    return sig * (6.0f * (q * q * q - q * q) + 1.0f);   /* :31 */
}

static inline double W_d(double r2, double h2, double hinv, double sig)
{
    if (r2 >= h2) return 0.0;
    double q = sqrt(r2) * hinv;
    if (q >= 0.5) {
        double qq = 1.0 - q;
        return sig * (2.0 * qq * qq * qq);
    }
    return sig * (6.0 * (q * q * q - q * q) + 1.0);
}

float or_W_f(float r2, float h) { return W_f(r2, h * h, 1.0f / h, or_sigma_f(h)); }

double or_W_d(double r2, double h)
{
    return W_d(r2, h * h, 1.0 / h, 8.0 / (OR_PI_D * h * h * h));
}

/* gradW with the reference's extra 1/|r| (Kernel.cpp:43): gradQ = normalize(r)/(|r| h). */
static inline void gradW_f(const float r[3], float h, float h2, float hinv, float sig, float out[3])
{
    float rn = (r[0] * r[0] + r[1] * r[1]) + r[2] * r[2];
    if (rn >= h2) { out[0] = out[1] = out[2] = 0.0f; return; }
    float rl = sqrtf(rn);
    float q = rl * hinv;
    float inv = 1.0f / sqrtf(rn);               /* glm::normalize = v * inversesqrt(dot) */
    float den = rl * h;
    float gq[3] = { (r[0] * inv) / den, (r[1] * inv) / den, (r[2] * inv) / den };
    if (q >= 0.5f) {
        float qq = 1.0f - q;
        float f = 6.0f * qq * qq;
        float ns = -sig;
        for (int a = 0; a < 3; a++) out[a] = (ns * gq[a]) * f;
        return;
    }
    float f = 6.0f * (3.0f * q * q - 2.0f * q);
    for (int a = 0; a < 3; a++) out[a] = (sig * gq[a]) * f;
}

void or_gradW_f(const float r[3], float h, float out[3])
{
    gradW_f(r, h, h * h, 1.0f / h, or_sigma_f(h), out);
}

/* ------------------------------------------------------------------------- */
/* Cell keys.  CompactNSearch (vendor/CompactNSearch, absent here; call sites */
/* Dataset.cpp:54-74, 277) keys a point by int(x / r), minus one for x < 0,   */
/* with r the search radius = ParticleRadius = h.                             */
/* ------------------------------------------------------------------------- */

int32_t or_key(float x, float inv_h)
{
    float v = inv_h * x;
    return x >= 0.0f ? (int32_t)v : (int32_t)v - 1;
}

uint64_t or_grid_struct_size(void) { return sizeof(or_grid); }

static inline uint64_t cell_lin(const or_grid *g, int32_t x, int32_t y, int32_t z)
{
    return (uint64_t)x + (uint64_t)g->dims[0] * ((uint64_t)y + (uint64_t)g->dims[1] * (uint64_t)z);
}

/* Dense-cell coordinates of the stencil centre; 0 if the 27-cell stencil leaves
 * the grid, in which case no particle can lie within h of q. */
static inline int stencil_centre(const or_grid *g, float x, float y, float z, int32_t c[3])
{
    c[0] = or_key(x, g->inv_h) - g->origin[0];
    c[1] = or_key(y, g->inv_h) - g->origin[1];
    c[2] = or_key(z, g->inv_h) - g->origin[2];
    for (int a = 0; a < 3; a++)
        if (c[a] < 1 || c[a] > g->dims[a] - 2) return 0;
    return 1;
}

static void densitygrid_build(or_grid *g);
static void empty27_build(or_grid *g);

/* Stable counting sort of the particles into cells of key rule inv (cell
 * width 1/inv): histogram, exclusive scan, scatter in input order.  pad empty
 * cells of margin on each side of the particles' key range. */
static int sort_cells(const float *xyz, uint64_t n, float inv, const float pmin[3], const float pmax[3], int pad,
                      int32_t origin[3], int32_t dims[3], uint64_t *ncells, uint32_t **cell_start, float **sorted,
                      uint32_t **sorted_index)
{
    uint64_t nc = 1;
    for (int a = 0; a < 3; a++) {
        int32_t kmin = or_key(pmin[a], inv), kmax = or_key(pmax[a], inv);
        origin[a] = kmin - pad;
        dims[a] = kmax - kmin + 2 * pad + 1;
        nc *= (uint64_t)dims[a];
    }
    if (nc > (1ull << 31)) return -2;
    *ncells = nc;
    uint32_t *cs = (uint32_t *)calloc(nc + 1, sizeof(uint32_t));
    float *srt = (float *)malloc(4 * n * sizeof(float));
    uint32_t *sidx = sorted_index ? (uint32_t *)malloc(n * sizeof(uint32_t)) : NULL;
    uint32_t *cell = (uint32_t *)malloc(n * sizeof(uint32_t));
    uint32_t *cursor = (uint32_t *)malloc(nc * sizeof(uint32_t));
    *cell_start = cs;
    *sorted = srt;
    if (sorted_index) *sorted_index = sidx;
    if (!cs || !srt || (sorted_index && !sidx) || !cell || !cursor) { free(cell); free(cursor); return -3; }
    for (uint64_t i = 0; i < n; i++) {
        int32_t c[3];
        for (int a = 0; a < 3; a++) c[a] = or_key(xyz[3 * i + a], inv) - origin[a];
        cell[i] = (uint32_t)((uint64_t)c[0] + (uint64_t)dims[0] * ((uint64_t)c[1] + (uint64_t)dims[1] * (uint64_t)c[2]));
        cs[cell[i] + 1]++;
    }
    for (uint64_t c = 0; c < nc; c++) cs[c + 1] += cs[c];
    memcpy(cursor, cs, nc * sizeof(uint32_t));
    for (uint64_t i = 0; i < n; i++) {
        uint32_t slot = cursor[cell[i]]++;
        srt[4 * slot + 0] = xyz[3 * i + 0];
        srt[4 * slot + 1] = xyz[3 * i + 1];
        srt[4 * slot + 2] = xyz[3 * i + 2];
        srt[4 * slot + 3] = 0.0f;
        if (sidx) sidx[slot] = (uint32_t)i;
    }
    free(cursor);
    free(cell);
    return 0;
}

int or_grid_build(const float *xyz, uint64_t n, float h, or_grid *g)
{
    memset(g, 0, sizeof(*g));
    if (n == 0 || !(h > 0.0f)) return -1;
    g->h = h;
    g->inv_h = 1.0f / h;
    g->n = n;
    /* Frame::ComputeAABB (Dataset.cpp:78-92): min/max, then pad by ParticleRadius. */
    for (int a = 0; a < 3; a++) { g->pmin[a] = xyz[a]; g->pmax[a] = xyz[a]; }
    for (uint64_t i = 0; i < n; i++)
        for (int a = 0; a < 3; a++) {
            float v = xyz[3 * i + a];
            if (v < g->pmin[a]) g->pmin[a] = v;
            if (v > g->pmax[a]) g->pmax[a] = v;
        }
    float pad = 1.0f * h;
    for (int a = 0; a < 3; a++) { g->bmin[a] = g->pmin[a] - pad; g->bmax[a] = g->pmax[a] + pad; }

    /* Keys are monotone in x, so the key range follows from the bounds.  Two
     * empty cells of margin on each side keep every stencil of a point inside
     * the padded AABB within the grid (four fine cells for the march grid). */
    int rc = sort_cells(xyz, n, g->inv_h, g->pmin, g->pmax, 2, g->origin, g->dims, &g->ncells, &g->cell_start,
                        &g->sorted, &g->sorted_index);
    if (rc) { or_grid_free(g); return rc; }
    g->inv_hf = 2.0f * g->inv_h;
    rc = sort_cells(xyz, n, g->inv_hf, g->pmin, g->pmax, 4, g->forigin, g->fdims, &g->fncells, &g->fcell_start,
                    &g->fsorted, NULL);
    if (rc) { or_grid_free(g); return rc; }
    densitygrid_build(g);
    empty27_build(g);
    return 0;
}

void or_grid_free(or_grid *g)
{
    free(g->cell_start); free(g->sorted); free(g->sorted_index);
    free(g->dg_flag_box); free(g->dg_flag_cons); free(g->dg_count_box); free(g->empty27); free(g->bound);
    free(g->fcell_start); free(g->fsorted);
    memset(g, 0, sizeof(*g));
}

/* ------------------------------------------------------------------------- */
/* Neighbour search                                                           */
/* Canonical order: the 27 cells around key(q) row by row (dz outer, dy, then */
/* the three x-cells of a row), each cell in ascending input index.           */
/* The distance test is |q - x_j|^2 < h^2 with ((dx dx + dy dy) + dz dz).      */
/* ------------------------------------------------------------------------- */

uint32_t or_neighbors_brute(const float *xyz, uint64_t n, float h, const float q[3],
                            uint32_t *out, uint32_t cap)
{
    float h2 = h * h;
    uint32_t cnt = 0;
    for (uint64_t j = 0; j < n; j++) {
        float dx = q[0] - xyz[3 * j], dy = q[1] - xyz[3 * j + 1], dz = q[2] - xyz[3 * j + 2];
        float r2 = (dx * dx + dy * dy) + dz * dz;
        if (r2 < h2) { if (cnt < cap) out[cnt] = (uint32_t)j; cnt++; }
    }
    return cnt;
}

uint32_t or_neighbors_grid(const or_grid *g, const float q[3], uint32_t *out, uint32_t cap,
                           uint32_t *ncand)
{
    int32_t c[3];
    uint32_t cnt = 0, cand = 0;
    float h2 = g->h * g->h;
    if (stencil_centre(g, q[0], q[1], q[2], c)) {
        for (int dz = -1; dz <= 1; dz++)
            for (int dy = -1; dy <= 1; dy++) {
                uint64_t base = cell_lin(g, c[0] - 1, c[1] + dy, c[2] + dz);
                uint32_t b = g->cell_start[base], e = g->cell_start[base + 3];
                cand += e - b;
                for (uint32_t s = b; s < e; s++) {
                    const float *p = g->sorted + 4 * (uint64_t)s;
                    float dx = q[0] - p[0], dyy = q[1] - p[1], dzz = q[2] - p[2];
                    float r2 = (dx * dx + dyy * dyy) + dzz * dzz;
                    if (r2 < h2) { if (cnt < cap) out[cnt] = g->sorted_index[s]; cnt++; }
                }
            }
    }
    if (ncand) *ncand = cand;
    return cnt;
}

/* Hashed cell table: the CompactNSearch-style structure (cell key -> list of
 * point indices, 27 lookups per query) used by the reference-style baseline. */
typedef struct { int32_t k[3]; uint32_t start, count; } hash_slot;
struct or_hash {
    float h, inv_h;
    const float *xyz;
    uint64_t n;
    uint32_t *idx;        /* point indices grouped by cell, ascending inside a cell */
    hash_slot *table;     /* open addressing, count == 0 marks an empty slot        */
    uint64_t mask;
};

static inline uint64_t hash3(int32_t x, int32_t y, int32_t z)
{
    /* the usual large-prime spatial hash */
    return ((uint64_t)(uint32_t)x * 73856093ull) ^ ((uint64_t)(uint32_t)y * 19349663ull) ^
           ((uint64_t)(uint32_t)z * 83492791ull);
}

static const hash_slot *hash_find(const or_hash *hs, int32_t x, int32_t y, int32_t z)
{
    uint64_t s = hash3(x, y, z) & hs->mask;
    for (;;) {
        const hash_slot *e = &hs->table[s];
        if (e->count == 0) return NULL;
        if (e->k[0] == x && e->k[1] == y && e->k[2] == z) return e;
        s = (s + 1) & hs->mask;
    }
}

typedef struct { int32_t k[3]; uint32_t i; } keyed;
static int keyed_cmp(const void *a, const void *b)
{
    const keyed *x = (const keyed *)a, *y = (const keyed *)b;
    for (int d = 2; d >= 0; d--) {
        if (x->k[d] != y->k[d]) return x->k[d] < y->k[d] ? -1 : 1;
    }
    return x->i < y->i ? -1 : (x->i > y->i);
}

or_hash *or_hash_build(const float *xyz, uint64_t n, float h)
{
    or_hash *hs = (or_hash *)calloc(1, sizeof(or_hash));
    hs->h = h; hs->inv_h = 1.0f / h; hs->xyz = xyz; hs->n = n;
    keyed *kv = (keyed *)malloc(n * sizeof(keyed));
    for (uint64_t i = 0; i < n; i++) {
        for (int a = 0; a < 3; a++) kv[i].k[a] = or_key(xyz[3 * i + a], hs->inv_h);
        kv[i].i = (uint32_t)i;
    }
    qsort(kv, n, sizeof(keyed), keyed_cmp);
    uint64_t cap = 16;
    while (cap < 2 * n) cap <<= 1;
    hs->mask = cap - 1;
    hs->table = (hash_slot *)calloc(cap, sizeof(hash_slot));
    hs->idx = (uint32_t *)malloc(n * sizeof(uint32_t));
    uint64_t i = 0;
    while (i < n) {
        uint64_t j = i;
        while (j < n && kv[j].k[0] == kv[i].k[0] && kv[j].k[1] == kv[i].k[1] && kv[j].k[2] == kv[i].k[2]) {
            hs->idx[j] = kv[j].i;
            j++;
        }
        uint64_t s = hash3(kv[i].k[0], kv[i].k[1], kv[i].k[2]) & hs->mask;
        while (hs->table[s].count) s = (s + 1) & hs->mask;
        memcpy(hs->table[s].k, kv[i].k, sizeof(kv[i].k));
        hs->table[s].start = (uint32_t)i;
        hs->table[s].count = (uint32_t)(j - i);
        i = j;
    }
    free(kv);
    return hs;
}

void or_hash_free(or_hash *hs)
{
    if (!hs) return;
    free(hs->idx); free(hs->table); free(hs);
}

/* Dataset::GetNeighbors (Dataset.cpp:272-280) equivalent on the hashed table. */
static uint32_t hash_neighbors(const or_hash *hs, const float q[3], uint32_t *out, uint32_t cap,
                               uint32_t *ncand)
{
    int32_t k[3];
    for (int a = 0; a < 3; a++) k[a] = or_key(q[a], hs->inv_h);
    float h2 = hs->h * hs->h;
    uint32_t cnt = 0, cand = 0;
    for (int dz = -1; dz <= 1; dz++)
        for (int dy = -1; dy <= 1; dy++)
            for (int dx = -1; dx <= 1; dx++) {
                const hash_slot *e = hash_find(hs, k[0] + dx, k[1] + dy, k[2] + dz);
                if (!e) continue;
                cand += e->count;
                for (uint32_t s = e->start; s < e->start + e->count; s++) {
                    uint32_t j = hs->idx[s];
                    const float *p = hs->xyz + 3 * (uint64_t)j;
                    float ddx = q[0] - p[0], ddy = q[1] - p[1], ddz = q[2] - p[2];
                    float r2 = (ddx * ddx + ddy * ddy) + ddz * ddz;
                    if (r2 < h2) { if (cnt < cap) out[cnt] = j; cnt++; }
                }
            }
    if (ncand) *ncand = cand;
    return cnt;
}

uint32_t or_neighbors_hash(const float *xyz, uint64_t n, float h, const float q[3],
                           uint32_t *out, uint32_t cap)
{
    or_hash *hs = or_hash_build(xyz, n, h);
    uint32_t c = hash_neighbors(hs, q, out, cap, NULL);
    or_hash_free(hs);
    return c;
}

/* ------------------------------------------------------------------------- */
/* Density: rho(x) = sum_j W(x_j - x) (RayMarcher.cpp:315-325), in canonical */
/* order.  Compensated form: Neumaier's variant of Kahan summation.           */
/* ------------------------------------------------------------------------- */

/* ---- march-grid stencil -------------------------------------------------- */
/* In fine-cell units u = inv_hf x, the point sits at offset f in [0,1] inside
 * its cell (f = 1 only for CompactNSearch's negative-integer key).  The
 * distance from the point to the cell at offset d along one axis is
 * d - f (d > 0), f - (d + 1) (d < 0), 0 (d = 0).  A row or row end is kept
 * when its squared distance is below 4 (1 + 2^-10) (h = 2 fine cells): the
 * margin only keeps cells, so every particle within h is in a kept cell.    */

#define OR_FINE_LIM 4.00390625f

static inline float fine_off(int d, float f)
{
    return d > 0 ? (float)d - f : (d < 0 ? f - (float)(d + 1) : 0.0f);
}

static inline uint64_t fine_lin(const or_grid *g, int32_t x, int32_t y, int32_t z)
{
    return (uint64_t)x + (uint64_t)g->fdims[0] * ((uint64_t)y + (uint64_t)g->fdims[1] * (uint64_t)z);
}

int or_fine_centre(const or_grid *g, const float q[3], or_fine *fs)
{
    int ok = 1;
    for (int a = 0; a < 3; a++) {
        float u = g->inv_hf * q[a];
        int32_t k = q[a] >= 0.0f ? (int32_t)u : (int32_t)u - 1;
        fs->f[a] = u - (float)k;
        fs->c[a] = k - g->forigin[a];
        if (fs->c[a] < 2 || fs->c[a] > g->fdims[a] - 3) ok = 0;
    }
    return ok;
}

int or_fine_row(const or_grid *g, const or_fine *fs, int dy, int dz, uint32_t *b, uint32_t *e)
{
    float ty = fine_off(dy, fs->f[1]), tz = fine_off(dz, fs->f[2]);
    float r2 = ty * ty + tz * tz;
    if (!(r2 < OR_FINE_LIM)) return 0;
    float lm2 = fine_off(-2, fs->f[0]), lm1 = fine_off(-1, fs->f[0]);
    float lp1 = fine_off(1, fs->f[0]), lp2 = fine_off(2, fs->f[0]);
    int lo = (lm2 * lm2 + r2 < OR_FINE_LIM) ? -2 : ((lm1 * lm1 + r2 < OR_FINE_LIM) ? -1 : 0);
    int hi = (lp2 * lp2 + r2 < OR_FINE_LIM) ? 2 : ((lp1 * lp1 + r2 < OR_FINE_LIM) ? 1 : 0);
    int32_t y = fs->c[1] + dy, z = fs->c[2] + dz;
    *b = g->fcell_start[fine_lin(g, fs->c[0] + lo, y, z)];
    *e = g->fcell_start[fine_lin(g, fs->c[0] + hi, y, z) + 1];
    return 1;
}

/* Fused central differences (the GPU takes the six samples q +- eps e_a in
 * one walk): the march-grid cells near the box [q - eps, q + eps], per axis
 * from the fine cell of the box's low end to that of its high end, +-2 cells;
 * rows and row ends kept by their distance to the box (fine units) below
 * OR_FINE_LIM.  The box is never farther from a cell than any of the six
 * samples, so these cells include every cell their own stencils keep.  The
 * walk adds the candidates of its kept rows once; returns 0 when the cells
 * leave the grid (then the samples are taken one at a time).  With out != NULL
 * the positions of the candidates within (h + eps)(1 + 2^-10) of q are listed
 * (3 floats each, up to cap; the count is returned in *nout).               */
static inline float box_off(float c, float ulo, float uhi, float kl, float kh)
{
    return c > kh ? c - uhi : (c < kl ? ulo - (c + 1.0f) : 0.0f);
}

int or_cdiff_union(const or_grid *g, const float q[3], float eps, uint32_t *ncand, float *out, uint32_t cap,
                   uint32_t *nout)
{
    float ulo[3], uhi[3], kl[3], kh[3];
    int32_t ikl[3], ikh[3];
    for (int a = 0; a < 3; a++) {
        float lo = q[a] - eps, hi = q[a] + eps;
        ulo[a] = g->inv_hf * lo;
        uhi[a] = g->inv_hf * hi;
        ikl[a] = lo >= 0.0f ? (int32_t)ulo[a] : (int32_t)ulo[a] - 1;
        ikh[a] = hi >= 0.0f ? (int32_t)uhi[a] : (int32_t)uhi[a] - 1;
        kl[a] = (float)ikl[a];
        kh[a] = (float)ikh[a];
        if (ikl[a] - 2 - g->forigin[a] < 0 || ikh[a] + 2 - g->forigin[a] > g->fdims[a] - 1) return 0;
    }
    float t;
    t = box_off(kl[0] - 2.0f, ulo[0], uhi[0], kl[0], kh[0]); const float xm2 = t * t;
    t = box_off(kl[0] - 1.0f, ulo[0], uhi[0], kl[0], kh[0]); const float xm1 = t * t;
    t = box_off(kh[0] + 1.0f, ulo[0], uhi[0], kl[0], kh[0]); const float xp1 = t * t;
    t = box_off(kh[0] + 2.0f, ulo[0], uhi[0], kl[0], kh[0]); const float xp2 = t * t;
    const float he = (g->h + eps) * 1.0009765625f, R2 = he * he;
    uint32_t cand = 0, n = 0;
    for (int32_t cz = ikl[2] - 2; cz <= ikh[2] + 2; cz++) {
        for (int32_t cy = ikl[1] - 2; cy <= ikh[1] + 2; cy++) {
            float ty = box_off((float)cy, ulo[1], uhi[1], kl[1], kh[1]);
            float tz = box_off((float)cz, ulo[2], uhi[2], kl[2], kh[2]);
            float r2 = ty * ty + tz * tz;
            if (!(r2 < OR_FINE_LIM)) continue;
            int32_t lo = (xm2 + r2 < OR_FINE_LIM) ? ikl[0] - 2 : ((xm1 + r2 < OR_FINE_LIM) ? ikl[0] - 1 : ikl[0]);
            int32_t hi = (xp2 + r2 < OR_FINE_LIM) ? ikh[0] + 2 : ((xp1 + r2 < OR_FINE_LIM) ? ikh[0] + 1 : ikh[0]);
            int32_t y = cy - g->forigin[1], z = cz - g->forigin[2];
            uint32_t b = g->fcell_start[fine_lin(g, lo - g->forigin[0], y, z)];
            uint32_t e = g->fcell_start[fine_lin(g, hi - g->forigin[0], y, z) + 1];
            cand += e - b;
            if (!out) continue;
            for (uint32_t i = b; i < e; i++) {
                const float *p = g->fsorted + 4 * (uint64_t)i;
                float rx = p[0] - q[0], ry = p[1] - q[1], rz = p[2] - q[2];
                if (!((rx * rx + ry * ry) + rz * rz < R2)) continue;
                if (n < cap) { out[3 * n] = p[0]; out[3 * n + 1] = p[1]; out[3 * n + 2] = p[2]; }
                n++;
            }
        }
    }
    *ncand = cand;
    if (nout) *nout = n;
    return 1;
}

uint32_t or_neighbors_fine(const or_grid *g, const float q[3], float *out, uint32_t cap, uint32_t *ncand)
{
    or_fine fs;
    float h2 = g->h * g->h;
    uint32_t cnt = 0, cand = 0;
    if (or_fine_centre(g, q, &fs)) {
        for (int dz = -2; dz <= 2; dz++)
            for (int dy = -2; dy <= 2; dy++) {
                uint32_t b, e;
                if (!or_fine_row(g, &fs, dy, dz, &b, &e)) continue;
                cand += e - b;
                for (uint32_t s = b; s < e; s++) {
                    const float *p = g->fsorted + 4 * (uint64_t)s;
                    float rx = p[0] - q[0], ry = p[1] - q[1], rz = p[2] - q[2];
                    float r2 = (rx * rx + ry * ry) + rz * rz;
                    if (r2 < h2) {
                        if (cnt < cap) { out[3 * cnt] = p[0]; out[3 * cnt + 1] = p[1]; out[3 * cnt + 2] = p[2]; }
                        cnt++;
                    }
                }
            }
    }
    if (ncand) *ncand = cand;
    return cnt;
}

typedef struct { float s, c; } neu_f;
static inline void neu_add(neu_f *a, float w)
{
    float t = a->s + w;
    if (fabsf(a->s) >= fabsf(w)) a->c += (a->s - t) + w;
    else a->c += (w - t) + a->s;
    a->s = t;
}

float or_density_f(const or_grid *g, const float q[3], int compensated, uint32_t *ncand)
{
    or_fine fs;
    float h2 = g->h * g->h, hinv = 1.0f / g->h, sig = or_sigma_f(g->h);
    neu_f acc = { 0.0f, 0.0f };
    float plain = 0.0f;
    uint32_t cand = 0;
    if (or_fine_centre(g, q, &fs)) {
        for (int dz = -2; dz <= 2; dz++)
            for (int dy = -2; dy <= 2; dy++) {
                uint32_t b, e;
                if (!or_fine_row(g, &fs, dy, dz, &b, &e)) continue;
                cand += e - b;
                for (uint32_t s = b; s < e; s++) {
                    const float *p = g->fsorted + 4 * (uint64_t)s;
                    float rx = p[0] - q[0], ry = p[1] - q[1], rz = p[2] - q[2];
                    float r2 = (rx * rx + ry * ry) + rz * rz;
                    if (r2 < h2) {
                        float w = W_f(r2, h2, hinv, sig);
                        if (compensated) neu_add(&acc, w);
                        else plain += w;
                    }
                }
            }
    }
    if (ncand) *ncand = cand;
    return compensated ? acc.s + acc.c : plain;
}

double or_density_d(const or_grid *g, const double q[3])
{
    double h = (double)g->h;
    double h2 = h * h, hinv = 1.0 / h, sig = 8.0 / (OR_PI_D * h * h * h);
    or_fine fs;
    double acc = 0.0;
    /* The stencil is chosen from the FP32-rounded point; the FP32 and FP64
     * stencils can only differ for points within an ulp of a cell face, where
     * the difference is a cell at distance >= h and contributes nothing. */
    const float qf[3] = { (float)q[0], (float)q[1], (float)q[2] };
    if (or_fine_centre(g, qf, &fs)) {
        for (int dz = -2; dz <= 2; dz++)
            for (int dy = -2; dy <= 2; dy++) {
                uint32_t b, e;
                if (!or_fine_row(g, &fs, dy, dz, &b, &e)) continue;
                for (uint32_t s = b; s < e; s++) {
                    const float *p = g->fsorted + 4 * (uint64_t)s;
                    double rx = (double)p[0] - q[0], ry = (double)p[1] - q[1], rz = (double)p[2] - q[2];
                    double r2 = (rx * rx + ry * ry) + rz * rz;
                    if (r2 < h2) acc += W_d(r2, h2, hinv, sig);
                }
            }
    }
    return acc;
}

void or_density_batch_f(const or_grid *g, const float *q, uint64_t nq, int compensated,
                        float *rho, uint32_t *ncand)
{
    for (uint64_t i = 0; i < nq; i++)
        rho[i] = or_density_f(g, q + 3 * i, compensated, ncand ? ncand + i : NULL);
}

void or_density_batch_d(const or_grid *g, const double *q, uint64_t nq, double *rho)
{
    for (uint64_t i = 0; i < nq; i++) rho[i] = or_density_d(g, q + 3 * i);
}

/* ------------------------------------------------------------------------- */
/* Reference DensityGrid (Dataset.cpp:94-164, query :26-47).  Cell width h    */
/* over the padded AABB.  find_neighbors_box is a fork-only CompactNSearch    */
/* call whose source is absent; the box flag reads it as "some particle with  */
/* |x_j - centre|_inf < h", i.e. count * W(0) > 1.  The conservative flag     */
/* widens the box to 1.5h (+ margin) so that a cleared cell provably has      */
/* zero density everywhere inside it.                                         */
/* ------------------------------------------------------------------------- */

static uint32_t count_linf(const or_grid *g, const float c[3], float R)
{
    int32_t lo[3], hi[3];
    for (int a = 0; a < 3; a++) {
        lo[a] = or_key(c[a] - R, g->inv_h) - g->origin[a];
        hi[a] = or_key(c[a] + R, g->inv_h) - g->origin[a];
        if (lo[a] < 0) lo[a] = 0;
        if (hi[a] > g->dims[a] - 1) hi[a] = g->dims[a] - 1;
        if (lo[a] > hi[a]) return 0;
    }
    uint32_t cnt = 0;
    for (int32_t z = lo[2]; z <= hi[2]; z++)
        for (int32_t y = lo[1]; y <= hi[1]; y++) {
            uint32_t b = g->cell_start[cell_lin(g, lo[0], y, z)], e = g->cell_start[cell_lin(g, hi[0], y, z) + 1];
            for (uint32_t s = b; s < e; s++) {
                const float *p = g->sorted + 4 * (uint64_t)s;
                if (fabsf(p[0] - c[0]) < R && fabsf(p[1] - c[1]) < R && fabsf(p[2] - c[2]) < R) cnt++;
            }
        }
    return cnt;
}

static void densitygrid_build(or_grid *g)
{
    float cw = 1.0f * g->h;
    uint64_t tot = 1;
    for (int a = 0; a < 3; a++) {
        float ext = g->bmax[a] - g->bmin[a];
        g->dg_dims[a] = (int32_t)ceilf(ext / cw);
        tot *= (uint64_t)g->dg_dims[a];
    }
    g->dg_flag_box = (uint8_t *)calloc(tot, 1);
    g->dg_flag_cons = (uint8_t *)calloc(tot, 1);
    g->dg_count_box = (uint32_t *)calloc(tot, sizeof(uint32_t));
    float sig = or_sigma_f(g->h);
    for (int32_t z = 0; z < g->dg_dims[2]; z++)
        for (int32_t y = 0; y < g->dg_dims[1]; y++)
            for (int32_t x = 0; x < g->dg_dims[0]; x++) {
                float c[3] = { g->bmin[0] + ((float)x + 0.5f) * cw, g->bmin[1] + ((float)y + 0.5f) * cw,
                               g->bmin[2] + ((float)z + 0.5f) * cw };
                uint64_t id = (uint64_t)x + (uint64_t)g->dg_dims[0] * ((uint64_t)y + (uint64_t)g->dg_dims[1] * (uint64_t)z);
                uint32_t nb = count_linf(g, c, g->h);
                g->dg_count_box[id] = nb;
                g->dg_flag_box[id] = ((float)nb * sig > 1.0f);
                g->dg_flag_cons[id] = count_linf(g, c, 1.5f * g->h * 1.001f) > 0;
            }
}

/* empty27[c] = 1 when cells c-1..c+1 on every axis hold no particle; cells on
 * the grid border count their missing neighbours as empty. */
static void empty27_build(or_grid *g)
{
    g->empty27 = (uint8_t *)calloc(g->ncells, 1);
    for (int32_t z = 0; z < g->dims[2]; z++)
        for (int32_t y = 0; y < g->dims[1]; y++)
            for (int32_t x = 0; x < g->dims[0]; x++) {
                uint32_t cnt = 0;
                for (int dz = -1; dz <= 1; dz++)
                    for (int dy = -1; dy <= 1; dy++) {
                        int32_t yy = y + dy, zz = z + dz;
                        if (yy < 0 || yy >= g->dims[1] || zz < 0 || zz >= g->dims[2]) continue;
                        int32_t x0 = x > 0 ? x - 1 : 0, x1 = x < g->dims[0] - 1 ? x + 1 : g->dims[0] - 1;
                        cnt += g->cell_start[cell_lin(g, x1, yy, zz) + 1] - g->cell_start[cell_lin(g, x0, yy, zz)];
                    }
                g->empty27[cell_lin(g, x, y, z)] = cnt == 0;
            }
}

/* Empty-space skip of the configured march.  If the sample's cell has an empty
 * 27-neighbourhood, every sample the ray places inside that cell has zero
 * density and cannot be the crossing, so the march resumes at the last sample
 * before the ray leaves the cell box [key h, (key+1) h]; the first-crossing
 * sample, and so the whole result, is the same as without the skip.  Returns
 * the next sample index (k when there is nothing to skip). */
static int64_t skip_empty(const or_grid *g, const float o[3], const float d[3], const float x[3], float tn,
                          float step, int64_t k)
{
    int32_t c[3];
    if (!stencil_centre(g, x[0], x[1], x[2], c)) return k;
    if (!g->empty27[cell_lin(g, c[0], c[1], c[2])]) return k;
    float texit = INFINITY;
    for (int a = 0; a < 3; a++) {
        if (d[a] == 0.0f) continue;
        int32_t key = c[a] + g->origin[a];
        float lo = (float)key * g->h, hi = (float)(key + 1) * g->h;
        float ta = ((d[a] > 0.0f ? hi : lo) - o[a]) / d[a];
        if (ta < texit) texit = ta;
    }
    float kf = floorf((texit - tn) / step);
    int64_t kn = kf > (float)k ? (int64_t)kf : k + 1;
    return kn;
}

/* Sub-box density bound (configured-mode skip, no counterpart in the
 * reference).  Cell c is split into 8x8x8 sub-boxes; sub-box (fx,fy,fz) holds
 * the points whose u = inv_h * x lies in [key + f/8, key + (f+1)/8) on every
 * axis, and quarter box s = f/2 (4x4x4 per cell) the eight with the same s.
 * With B a box widened by delta = h (2^-10 + (max|key| + 1) 2^-18),
 * U = sum over the cell's 27-cell stencil of W(dist(x_j, B)) bounds the density
 * of every point inside (W falls with r; the stencil holds every particle
 * within h).  A sub-box's bit is set when U of its quarter box, or else its own
 * U, is below the limit iso (1 - 2^-9) -- except in cells whose centre row
 * (dz = dy = 0) alone takes every quarter box's U past that limit (inside the
 * fluid: no bit) and in quarter boxes whose U reaches 4 x the limit (no bit; a
 * clear bit only means the sample is evaluated; not for the zero-density bound,
 * iso = OR_ZERO_ISO, where a quarter box's U only says that some particle lies
 * within h of it).  A cell holds 8 words:
 * sub-box (fx,fy,fz) is bit
 * 8 l + (fx&1) + 2 (fy&1) + 4 (fz&1) (bit b of word b / 64) with
 * l = fx/2 + 4 (fy/2) + 16 (fz/2) its quarter box.  Border cells and cells with
 * an empty 27-neighbourhood get every bit.  Each U sums the centre row
 * (dz = dy = 0) first, then the other eight rows in (dz, dy) order, particles
 * in storage order, as the GPU build does (it stops a walk once every box it
 * sums is past the limit; adding terms >= 0 never lowers U again).  Cells are
 * split over threads by z-slab; each cell is computed by one thread, so the
 * bits do not depend on the thread count. */
typedef struct {
    or_grid *g;
    float lim;
    int quarter_only;
    int32_t next_z;
    pthread_mutex_t lock;
} bound_job;

/* U of the box lo..hi (already widened) over the first nrows rows of cell
 * (x,y,z)'s stencil, its walk stopped at the first row end with U >= lim. */
static float bound_sum(const or_grid *g, int32_t x, int32_t y, int32_t z, const float lo[3], const float hi[3],
                       float lim, int nrows)
{
    static const int8_t order[9][2] = { { 0, 0 }, { -1, -1 }, { -1, 0 }, { -1, 1 }, { 0, -1 },
                                        { 0, 1 }, { 1, -1 }, { 1, 0 }, { 1, 1 } };
    const float h = g->h, h2 = h * h, sig = or_sigma_f(h);
    float U = 0.0f;
    for (int r = 0; r < nrows && U < lim; r++) {
        const int dz = order[r][0], dy = order[r][1];
        uint64_t row = cell_lin(g, x - 1, y + dy, z + dz);
        for (uint32_t j = g->cell_start[row]; j < g->cell_start[row + 3]; j++) {
            const float *p = g->sorted + 4 * (uint64_t)j;
            float e[3];
            for (int a = 0; a < 3; a++) {
                float v = lo[a] - p[a], w = p[a] - hi[a];
                float m = v > w ? v : w;
                e[a] = m > 0.0f ? m : 0.0f;
            }
            float d2 = (e[0] * e[0] + e[1] * e[1]) + e[2] * e[2];
            U += W_f(d2, h2, g->inv_h, sig);
        }
    }
    return U;
}

static void bound_cell_bits(or_grid *g, float lim, int quarter_only, int32_t x, int32_t y, int32_t z)
{
    const uint64_t c = cell_lin(g, x, y, z);
    uint64_t *out = g->bound + 8 * c;
    if (x < 1 || y < 1 || z < 1 || x > g->dims[0] - 2 || y > g->dims[1] - 2 || z > g->dims[2] - 2 || g->empty27[c]) {
        for (int w = 0; w < 8; w++) out[w] = ~0ull;
        return;
    }
    const float h = g->h;
    int32_t k[3] = { x + g->origin[0], y + g->origin[1], z + g->origin[2] };
    int32_t km = 0;
    for (int a = 0; a < 3; a++) km = abs(k[a]) > km ? abs(k[a]) : km;
    const float delta = h * (0x1p-10f + (float)(km + 1) * 0x1p-18f);
    for (int w = 0; w < 8; w++) out[w] = 0;
    int inside = 1;
    for (int l = 0; l < 64 && inside; l++) {
        const int sq[3] = { l & 3, (l >> 2) & 3, l >> 4 };
        float lo[3], hi[3];
        for (int a = 0; a < 3; a++) {
            lo[a] = ((float)k[a] + 0.25f * (float)sq[a]) * h - delta;
            hi[a] = ((float)k[a] + 0.25f * (float)(sq[a] + 1)) * h + delta;
        }
        if (bound_sum(g, x, y, z, lo, hi, lim, 1) < lim) inside = 0;
    }
    if (inside) return;                          /* the centre row alone: every quarter box past the limit */
    for (int l = 0; l < 64; l++) {
        const int sq[3] = { l & 3, (l >> 2) & 3, l >> 4 };
        float lo[3], hi[3];
        for (int a = 0; a < 3; a++) {
            lo[a] = ((float)k[a] + 0.25f * (float)sq[a]) * h - delta;
            hi[a] = ((float)k[a] + 0.25f * (float)(sq[a] + 1)) * h + delta;
        }
        unsigned byte = 0;
        const float lim4 = g->bound_iso == OR_ZERO_ISO ? INFINITY : 4.0f * lim;
        const float Uq = bound_sum(g, x, y, z, lo, hi, quarter_only ? lim : lim4, 9);
        if (Uq < lim) {
            byte = 0xffu;                       /* the quarter box's bound covers its eight */
        } else if (!quarter_only && Uq < lim4) {
            for (int f = 0; f < 8; f++) {
                const int fi[3] = { 2 * sq[0] + (f & 1), 2 * sq[1] + ((f >> 1) & 1), 2 * sq[2] + (f >> 2) };
                for (int a = 0; a < 3; a++) {
                    lo[a] = ((float)k[a] + 0.125f * (float)fi[a]) * h - delta;
                    hi[a] = ((float)k[a] + 0.125f * (float)(fi[a] + 1)) * h + delta;
                }
                if (bound_sum(g, x, y, z, lo, hi, lim, 9) < lim) byte |= 1u << f;
            }
        }
        out[l >> 3] |= (uint64_t)byte << (8 * (l & 7));
    }
}

static void *bound_worker(void *arg)
{
    bound_job *j = (bound_job *)arg;
    for (;;) {
        pthread_mutex_lock(&j->lock);
        int32_t z = j->next_z++;
        pthread_mutex_unlock(&j->lock);
        if (z >= j->g->dims[2]) return NULL;
        for (int32_t y = 0; y < j->g->dims[1]; y++)
            for (int32_t x = 0; x < j->g->dims[0]; x++) bound_cell_bits(j->g, j->lim, j->quarter_only, x, y, z);
    }
}

int or_bound_build(or_grid *g, float iso, int32_t boxes)
{
    if (boxes != 0 && boxes != 1) return -1;
    free(g->bound);
    g->bound = (uint64_t *)calloc((g->ncells ? g->ncells : 1) * 8, sizeof(uint64_t));
    if (!g->bound) return -1;
    g->bound_iso = iso;
    g->bound_boxes = boxes;
    bound_job j;
    j.g = g;
    j.lim = iso * (1.0f - 0x1p-9f);
    j.quarter_only = boxes == 1;
    j.next_z = 0;
    pthread_mutex_init(&j.lock, NULL);
    long nc = sysconf(_SC_NPROCESSORS_ONLN);
    int nthreads = nc > 1 ? (int)(nc - 1 < 16 ? nc - 1 : 16) : 1;
    pthread_t th[16];
    for (int t = 0; t < nthreads; t++) pthread_create(&th[t], NULL, bound_worker, &j);
    for (int t = 0; t < nthreads; t++) pthread_join(th[t], NULL);
    pthread_mutex_destroy(&j.lock);
    return 0;
}

/* Bound skip of the configured march: the sample's sub-box (eighth box) has
 * its bit set, so the march resumes at the first sample past the box exit --
 * past the exit of its quarter box when all eight eighth boxes of that are set
 * (the exit index lowered by 2^-8 + |k| 2^-18 steps against rounding; slab
 * distances use the reciprocal direction and step, as the GPU computes them).
 * Returns the next sample index (k when there is nothing to skip). */
static int64_t skip_bound(const or_grid *g, const float o[3], const float d[3], const float x[3], float tn,
                          float step, int64_t k)
{
    int32_t key[3], c[3], fi[3];
    for (int a = 0; a < 3; a++) {
        float u = g->inv_h * x[a];
        key[a] = x[a] >= 0.0f ? (int32_t)u : (int32_t)u - 1;
        c[a] = key[a] - g->origin[a];
        if (c[a] < 0 || c[a] >= g->dims[a]) return k;
        int s8 = (int)((u - (float)key[a]) * 8.0f);
        fi[a] = s8 < 7 ? s8 : 7;
    }
    const int l = (fi[0] >> 1) + 4 * (fi[1] >> 1) + 16 * (fi[2] >> 1);
    const uint64_t *cb = g->bound + 8 * cell_lin(g, c[0], c[1], c[2]);
    const unsigned byte = (unsigned)(cb[l >> 3] >> (8 * (l & 7))) & 0xffu;
    float lo[3], span;
    if (byte == 0xffu) {
        for (int a = 0; a < 3; a++) lo[a] = (float)key[a] + 0.25f * (float)(fi[a] >> 1);
        span = 0.25f;
    } else {
        if (!((byte >> ((fi[0] & 1) + 2 * (fi[1] & 1) + 4 * (fi[2] & 1))) & 1u)) return k;
        for (int a = 0; a < 3; a++) lo[a] = (float)key[a] + 0.125f * (float)fi[a];
        span = 0.125f;
    }
    float texit = INFINITY;
    for (int a = 0; a < 3; a++) {
        if (d[a] == 0.0f) continue;
        float blo = lo[a] * g->h, bhi = (lo[a] + span) * g->h;
        float id = 1.0f / d[a];
        float ta = ((d[a] > 0.0f ? bhi : blo) - o[a]) * id;
        if (ta < texit) texit = ta;
    }
    float kq = (texit - tn) * (1.0f / step);
    kq -= 0x1p-8f + fabsf(kq) * 0x1p-18f;
    float kf = floorf(kq);
    return kf >= (float)k ? (int64_t)kf + 1 : k + 1;
}

/* QueryDensityGrid (Dataset.cpp:26-47): returns the node id or -1. */
static int64_t dg_query(const or_grid *g, const float p[3], float nmin[3], float nmax[3])
{
    float inv = 1.0f / g->h;
    int32_t k[3];
    for (int a = 0; a < 3; a++) {
        k[a] = (int32_t)floorf((p[a] - g->bmin[a]) * inv);
        if (!(k[a] >= 0 && k[a] < g->dg_dims[a])) return -1;
    }
    for (int a = 0; a < 3; a++) {
        nmin[a] = g->bmin[a] + (float)k[a] * g->h;      /* :132 */
        nmax[a] = nmin[a] + g->h;                       /* :133 */
    }
    return (int64_t)k[0] + (int64_t)g->dg_dims[0] * ((int64_t)k[1] + (int64_t)g->dg_dims[1] * (int64_t)k[2]);
}

static inline float gmin(float x, float y) { return (y < x) ? y : x; }
static inline float gmax(float x, float y) { return (x < y) ? y : x; }

/* intersectAABB (RayMarcher.cpp:51-62): t is in units of dir (= the step). */
static void intersect_aabb(const float o[3], const float d[3], const float bmin[3],
                           const float bmax[3], float out[3])
{
    float t1[3], t2[3];
    for (int a = 0; a < 3; a++) {
        float tmn = (bmin[a] - o[a]) / d[a], tmx = (bmax[a] - o[a]) / d[a];
        t1[a] = gmin(tmn, tmx);
        t2[a] = gmax(tmn, tmx);
    }
    float tnear = gmax(gmax(t1[0], t1[1]), t1[2]);
    float tfar = gmin(gmin(t2[0], t2[1]), t2[2]);
    float t = (tnear < tfar) ? tfar : tnear;
    if (t >= 0.0f) for (int a = 0; a < 3; a++) out[a] = o[a] + d[a] * t;
    else for (int a = 0; a < 3; a++) out[a] = o[a];
}

/* ------------------------------------------------------------------------- */
/* Ray generation (RayMarcher.cpp:268-274) and the AABB slab test.           */
/* mat4 * vec4 as glm evaluates it: (m0 x + m1 y) + (m2 z + m3 w).           */
/* ------------------------------------------------------------------------- */

static void unproject(const float M[16], float cx, float cy, float cz, float p[3])
{
    float w[4];
    for (int r = 0; r < 4; r++) {
        float a0 = M[0 * 4 + r] * cx, a1 = M[1 * 4 + r] * cy;
        float a2 = M[2 * 4 + r] * cz, a3 = M[3 * 4 + r] * 1.0f;
        w[r] = (a0 + a1) + (a2 + a3);
    }
    p[0] = w[0] / w[3]; p[1] = w[1] / w[3]; p[2] = w[2] / w[3];
}

static void pixel_clip(uint32_t idx, uint32_t W, uint32_t H, float *cx, float *cy)
{
    float tw = 2.0f / (float)W, th = 2.0f / (float)H;   /* RayMarcher.cpp:88-89 */
    *cx = (float)(idx % W) * tw - 1.0f;
    *cy = (float)(idx / W) * th - 1.0f;
}

static void normalize3(const float v[3], float out[3])
{
    float dd = (v[0] * v[0] + v[1] * v[1]) + v[2] * v[2];
    float inv = 1.0f / sqrtf(dd);
    out[0] = v[0] * inv; out[1] = v[1] * inv; out[2] = v[2] * inv;
}

/* Slab test against [bmin,bmax] for t >= 0.  Returns 0 on a miss. */
static int slab(const float o[3], const float d[3], const float bmin[3], const float bmax[3],
                float *tn, float *tf)
{
    float tnear = 0.0f, tfar = INFINITY;
    for (int a = 0; a < 3; a++) {
        if (d[a] != 0.0f) {
            float inv = 1.0f / d[a];
            float t0 = (bmin[a] - o[a]) * inv, t1 = (bmax[a] - o[a]) * inv;
            float lo = t0 < t1 ? t0 : t1, hi = t0 < t1 ? t1 : t0;
            if (lo > tnear) tnear = lo;
            if (hi < tfar) tfar = hi;
        } else if (o[a] < bmin[a] || o[a] > bmax[a]) {
            return 0;
        }
    }
    *tn = tnear; *tf = tfar;
    return tnear <= tfar;
}

void or_ray(const float invpv[16], const float cam[3], uint32_t W, uint32_t H, uint32_t idx,
            float o[3], float d[3])
{
    float cx, cy, wp[3], dun[3];
    pixel_clip(idx, W, H, &cx, &cy);
    unproject(invpv, cx, cy, 1.0f, wp);
    for (int a = 0; a < 3; a++) { o[a] = cam[a]; dun[a] = wp[a] - cam[a]; }
    normalize3(dun, d);
}

int or_slab(const float o[3], const float d[3], const float bmin[3], const float bmax[3], float tnf[2])
{
    return slab(o, d, bmin, bmax, &tnf[0], &tnf[1]);
}

/* ------------------------------------------------------------------------- */
/* Search-grid walk (cells of h) for the anisotropic density.                 */
/* ------------------------------------------------------------------------- */

#define OR_MAX_NEIGHBORS 8192   /* RayMarcher.cpp:14 */
#define OR_NEAR_LIM 1.0009765625f   /* 1 (1 + 2^-10): rows within h = 1 cell      */
#define OR_JACOBI_SWEEPS 5

/* The walk over the cells within R cells of q (R = 2 for the extended radius
 * 2h, R = 1 for h): q's cell (CompactNSearch key) and its offsets f inside the
 * cell, in cell units; rows (dy, dz), dz outer, each row the cells whose
 * distance to q is below R (1 + 2^-10) -- the same rule as the march grid's
 * walk, with cells of h -- clamped to the grid (cells outside hold no
 * particle).  0 when every cell of the walk is outside the grid. */
typedef struct { int32_t c[3]; float f[3]; } or_sw;

static int sw_centre(const or_grid *g, const float q[3], int R, or_sw *w)
{
    int ok = 1;
    for (int a = 0; a < 3; a++) {
        float u = g->inv_h * q[a];
        int32_t k = q[a] >= 0.0f ? (int32_t)u : (int32_t)u - 1;
        w->f[a] = u - (float)k;
        w->c[a] = k - g->origin[a];
        if (w->c[a] < -R || w->c[a] > g->dims[a] - 1 + R) ok = 0;
    }
    return ok;
}

static int sw_row(const or_grid *g, const or_sw *w, int R, float lim, int dy, int dz, uint32_t *b, uint32_t *e)
{
    float ty = fine_off(dy, w->f[1]), tz = fine_off(dz, w->f[2]);
    float r2 = ty * ty + tz * tz;
    if (!(r2 < lim)) return 0;
    int32_t y = w->c[1] + dy, z = w->c[2] + dz;
    if (y < 0 || y >= g->dims[1] || z < 0 || z >= g->dims[2]) return 0;
    int lo = 0, hi = 0;
    for (int d = -R; d < 0; d++) {
        float t = fine_off(d, w->f[0]);
        if (t * t + r2 < lim) { lo = d; break; }
    }
    for (int d = R; d > 0; d--) {
        float t = fine_off(d, w->f[0]);
        if (t * t + r2 < lim) { hi = d; break; }
    }
    int32_t xl = w->c[0] + lo, xh = w->c[0] + hi;
    if (xl < 0) xl = 0;
    if (xh > g->dims[0] - 1) xh = g->dims[0] - 1;
    if (xl > xh) return 0;
    *b = g->cell_start[cell_lin(g, xl, y, z)];
    *e = g->cell_start[cell_lin(g, xh, y, z) + 1];
    return 1;
}

typedef struct { float h, h2, hinv, hx, R2x, hxinv, sig, k_n, k_r, k_s; int32_t n_eps; } an_par_f;
typedef struct { double h, h2, hinv, hx, R2x, hxinv, sig, k_n, k_r, k_s; int32_t n_eps; } an_par_d;

/* ParticleRadius h, ParticleRadiusExt = multiplier * h (Dataset.cpp:169-177),
 * AnisotropicKernel's sig = 8 / pi (Kernel.cpp:57-61), the clamp knobs of
 * VisualizationSettings (RayMarcher.h:22-25). */
static void an_par_make_f(const or_grid *g, const or_settings *s, an_par_f *a)
{
    a->h = g->h; a->h2 = a->h * a->h; a->hinv = 1.0f / a->h;
    a->hx = s->ext_mult * a->h; a->R2x = a->hx * a->hx; a->hxinv = 1.0f / a->hx;
    a->sig = 8.0f / OR_PI_F;
    a->k_n = s->k_n; a->k_r = s->k_r; a->k_s = s->k_s; a->n_eps = s->n_eps;
}

static void an_par_make_d(const or_grid *g, const or_settings *s, an_par_d *a)
{
    a->h = (double)g->h; a->h2 = a->h * a->h; a->hinv = 1.0 / a->h;
    a->hx = (double)s->ext_mult * a->h; a->R2x = a->hx * a->hx; a->hxinv = 1.0 / a->hx;
    a->sig = 8.0 / OR_PI_D;
    a->k_n = s->k_n; a->k_r = s->k_r; a->k_s = s->k_s; a->n_eps = s->n_eps;
}

#define REAL float
#define SFX(x) x##_f
#define SQRT sqrtf
#define FABS fabsf
#include "pm_aniso.inc"
#undef REAL
#undef SFX
#undef SQRT
#undef FABS
#define REAL double
#define SFX(x) x##_d
#define SQRT sqrt
#define FABS fabs
#include "pm_aniso.inc"
#undef REAL
#undef SFX
#undef SQRT
#undef FABS

float or_density_aniso_f(const or_grid *g, const or_settings *s, const float q[3], int compensated, uint32_t *ncand)
{
    an_par_f a;
    an_par_make_f(g, s, &a);
    uint32_t cand = 0;
    float rho = an_sum_f(g, &a, q, q, compensated, &cand, NULL);
    if (ncand) *ncand = cand;
    return rho;
}

double or_density_aniso_d(const or_grid *g, const or_settings *s, const double q[3])
{
    an_par_d a;
    an_par_make_d(g, s, &a);
    uint32_t cand = 0;
    const float qf[3] = { (float)q[0], (float)q[1], (float)q[2] };
    return an_sum_d(g, &a, qf, q, 0, &cand, NULL);
}

void or_gradW_aniso_f(const or_grid *g, const or_settings *s, const float q[3], float out[3])
{
    an_par_f a;
    an_par_make_f(g, s, &a);
    uint32_t cand = 0;
    an_sum_f(g, &a, q, q, 0, &cand, out);
}

uint32_t or_aniso_G_f(const or_grid *g, const or_settings *s, const float q[3], float G[9], float *detG)
{
    an_par_f a;
    an_par_make_f(g, s, &a);
    uint32_t cand = 0;
    return an_wpca_f(g, &a, q, q, G, detG, &cand);
}

uint32_t or_aniso_G_d(const or_grid *g, const or_settings *s, const double q[3], double G[9], double *detG)
{
    an_par_d a;
    an_par_make_d(g, s, &a);
    uint32_t cand = 0;
    const float qf[3] = { (float)q[0], (float)q[1], (float)q[2] };
    return an_wpca_d(g, &a, qf, q, G, detG, &cand);
}

uint32_t or_aniso_C_f(const or_grid *g, const or_settings *s, const float q[3], float c[6])
{
    an_par_f a;
    an_par_make_f(g, s, &a);
    uint32_t cand = 0;
    return an_cov_f(g, &a, q, q, c, &cand);
}

uint32_t or_aniso_C_d(const or_grid *g, const or_settings *s, const double q[3], double c[6])
{
    an_par_d a;
    an_par_make_d(g, s, &a);
    uint32_t cand = 0;
    const float qf[3] = { (float)q[0], (float)q[1], (float)q[2] };
    return an_cov_d(g, &a, qf, q, c, &cand);
}

void or_aniso_G_from_C_f(const or_grid *g, const or_settings *s, const float c[6], uint32_t n, float G[9],
                         float *detG)
{
    an_par_f a;
    an_par_make_f(g, s, &a);
    an_gmat_f(c, n, &a, G, detG);
}

void or_eig3_f(const float c[6], float lam[3], float V[9]) { an_eig3_f(c, lam, (float(*)[3])V); }
void or_eig3_d(const double c[6], double lam[3], double V[9]) { an_eig3_d(c, lam, (double(*)[3])V); }

uint32_t or_neighbors_ext(const or_grid *g, const float q[3], float hx, float *out, uint32_t cap, uint32_t *ncand)
{
    or_sw w;
    const float R2x = hx * hx;
    uint32_t n = 0, cand = 0;
    if (sw_centre(g, q, 2, &w)) {
        for (int dz = -2; dz <= 2; dz++)
            for (int dy = -2; dy <= 2; dy++) {
                uint32_t b, e;
                if (!sw_row(g, &w, 2, OR_FINE_LIM, dy, dz, &b, &e)) continue;
                cand += e - b;
                for (uint32_t i = b; i < e; i++) {
                    const float *p = g->sorted + 4 * (uint64_t)i;
                    float rx = p[0] - q[0], ry = p[1] - q[1], rz = p[2] - q[2];
                    if (!((rx * rx + ry * ry) + rz * rz < R2x)) continue;
                    if (n < cap) { out[3 * n] = p[0]; out[3 * n + 1] = p[1]; out[3 * n + 2] = p[2]; }
                    n++;
                }
            }
    }
    if (ncand) *ncand = cand;
    return n;
}

/* ------------------------------------------------------------------------- */
/* Density providers: dense grid or hashed table, same canonical order.       */
/* ------------------------------------------------------------------------- */

typedef struct {
    const or_grid *g;
    const or_hash *hs;   /* non-NULL: reference-style hashed query */
    uint32_t *nbuf;      /* thread-local neighbour buffer (hashed)  */
    float h2, hinv, sig;
    int compensated;
    const or_settings *an;   /* non-NULL: anisotropic density (PerPixel_Anisotropic) */
} dens_ctx;

static float dens_f(dens_ctx *dc, const float q[3], or_counters *cnt)
{
    uint32_t cand = 0;
    float rho;
    if (dc->an) {
        /* RayMarcher.cpp:378-400: WPCA over the extended neighbours, then the
         * anisotropic kernel over those within h. */
        rho = or_density_aniso_f(dc->g, dc->an, q, dc->compensated, &cand);
    } else if (dc->hs) {
        /* RayMarcher.cpp:309-325: neighbour list, then sum of W over it. */
        uint32_t nn = hash_neighbors(dc->hs, q, dc->nbuf, OR_MAX_NEIGHBORS, &cand);
        if (nn > OR_MAX_NEIGHBORS) nn = OR_MAX_NEIGHBORS;
        neu_f acc = { 0.0f, 0.0f };
        float plain = 0.0f;
        for (uint32_t i = 0; i < nn; i++) {
            const float *p = dc->hs->xyz + 3 * (uint64_t)dc->nbuf[i];
            float rx = p[0] - q[0], ry = p[1] - q[1], rz = p[2] - q[2];
            float r2 = (rx * rx + ry * ry) + rz * rz;
            float w = W_f(r2, dc->h2, dc->hinv, dc->sig);
            if (dc->compensated) neu_add(&acc, w); else plain += w;
        }
        rho = dc->compensated ? acc.s + acc.c : plain;
    } else {
        rho = or_density_f(dc->g, q, dc->compensated, &cand);
    }
    cnt->samples++;
    cnt->cand += cand;
    return rho;
}

static void sum_gradW(dens_ctx *dc, const float q[3], float out[3], or_counters *cnt)
{
    /* RayMarcher.cpp:331-338: normal += gradW(x_j - x) over the neighbours. */
    const or_grid *g = dc->g;
    or_fine fs;
    out[0] = out[1] = out[2] = 0.0f;
    uint32_t cand = 0;
    if (dc->an) {
        /* RayMarcher.cpp:405-409: normal += gradW(G, det G, r) with the
         * sample's own G. */
        or_gradW_aniso_f(g, dc->an, q, out);
        return;
    }
    if (or_fine_centre(g, q, &fs)) {
        for (int dz = -2; dz <= 2; dz++)
            for (int dy = -2; dy <= 2; dy++) {
                uint32_t b, e;
                if (!or_fine_row(g, &fs, dy, dz, &b, &e)) continue;
                cand += e - b;
                for (uint32_t s = b; s < e; s++) {
                    const float *p = g->fsorted + 4 * (uint64_t)s;
                    float r[3] = { p[0] - q[0], p[1] - q[1], p[2] - q[2] };
                    float r2 = (r[0] * r[0] + r[1] * r[1]) + r[2] * r[2];
                    if (r2 < dc->h2) {
                        float gw[3];
                        gradW_f(r, g->h, dc->h2, dc->hinv, dc->sig, gw);
                        out[0] += gw[0]; out[1] += gw[1]; out[2] += gw[2];
                    }
                }
            }
    }
    (void)cnt; (void)cand;
}

uint32_t or_stencil_cand(const or_grid *g, const float q[3])
{
    or_fine fs;
    uint32_t cand = 0;
    if (!or_fine_centre(g, q, &fs)) return 0;
    for (int dz = -2; dz <= 2; dz++)
        for (int dy = -2; dy <= 2; dy++) {
            uint32_t b, e;
            if (or_fine_row(g, &fs, dy, dz, &b, &e)) cand += e - b;
        }
    return cand;
}

uint32_t or_cand27(const or_grid *g, const float q[3])
{
    int32_t c[3];
    uint32_t cand = 0;
    if (!stencil_centre(g, q[0], q[1], q[2], c)) return 0;
    for (int dz = -1; dz <= 1; dz++)
        for (int dy = -1; dy <= 1; dy++)
            cand += g->cell_start[cell_lin(g, c[0] + 1, c[1] + dy, c[2] + dz) + 1] -
                    g->cell_start[cell_lin(g, c[0] - 1, c[1] + dy, c[2] + dz)];
    return cand;
}

/* The per-sample counts of a configured isotropic march over the dense grid. */
static void count_sample(const dens_ctx *dc, or_counters *cnt, const float q[3], int march)
{
    if (dc->hs || dc->an) return;
    const uint32_t c = or_stencil_cand(dc->g, q);
    cnt->cand_stencil += c;
    cnt->cand27 += or_cand27(dc->g, q);
    if (march) cnt->march_cand += c;
}

/* ------------------------------------------------------------------------- */
/* One pixel of the configured march: AABB entry, fixed step, first sample    */
/* with rho >= iso, bisection (last iteration in FP64), normals.              */
/* ------------------------------------------------------------------------- */

static void pixel_config(dens_ctx *dc, const or_settings *s, const float M[16], const float cam[3],
                         uint32_t W, uint32_t H, uint32_t idx, float *pos, float *nrm, float *tout,
                         or_counters *cnt)
{
    const or_grid *g = dc->g;
    float *P = pos + 4 * (uint64_t)idx, *N = nrm + 4 * (uint64_t)idx;
    P[0] = P[1] = P[2] = P[3] = 0.0f;
    N[0] = N[1] = N[2] = N[3] = 0.0f;
    if (tout) tout[idx] = -1.0f;
    cnt->rays++;

    float cx, cy, wp[3], dun[3], d[3], o[3] = { cam[0], cam[1], cam[2] };
    pixel_clip(idx, W, H, &cx, &cy);
    unproject(M, cx, cy, 1.0f, wp);
    for (int a = 0; a < 3; a++) dun[a] = wp[a] - cam[a];
    normalize3(dun, d);

    float tn, tf;
    if (!slab(o, d, g->bmin, g->bmax, &tn, &tf)) return;
    cnt->rays_in_aabb++;

    float step = s->step;
    int64_t k = 0, hitk = -1;
    for (;;) {
        if (s->max_steps > 0 && k >= s->max_steps) break;
        float t = tn + (float)k * step;
        if (t > tf) break;
        float x[3] = { o[0] + t * d[0], o[1] + t * d[1], o[2] + t * d[2] };
        if (s->skip_mode == 1 || s->skip_mode == 4) {
            int64_t kn = s->skip_mode == 1 ? skip_empty(g, o, d, x, tn, step, k) : skip_bound(g, o, d, x, tn, step, k);
            if (kn != k) { k = kn; continue; }
        }
        float rho = dens_f(dc, x, cnt);
        cnt->march_samples++;
        count_sample(dc, cnt, x, 1);
        if (rho >= s->iso) { hitk = k; break; }
        k++;
    }
    if (hitk < 0) return;
    cnt->hits++;

    float xs[3], th;
    if (hitk == 0) {
        th = tn;
        for (int a = 0; a < 3; a++) xs[a] = o[a] + th * d[a];
    } else {
        float lo = tn + (float)(hitk - 1) * step, hi = tn + (float)hitk * step;
        int n32 = s->final_fp64 && s->bisect_iters > 0 ? s->bisect_iters - 1 : s->bisect_iters;
        for (int it = 0; it < n32; it++) {
            float tm = (lo + hi) * 0.5f;
            float x[3] = { o[0] + tm * d[0], o[1] + tm * d[1], o[2] + tm * d[2] };
            float rho = dens_f(dc, x, cnt);
            count_sample(dc, cnt, x, 0);
            if (rho >= s->iso) hi = tm; else lo = tm;
        }
        if (s->final_fp64 && s->bisect_iters > 0) {
            double dlo = lo, dhi = hi;
            double tm = (dlo + dhi) * 0.5;
            double x[3] = { (double)o[0] + tm * (double)d[0], (double)o[1] + tm * (double)d[1],
                            (double)o[2] + tm * (double)d[2] };
            double rho = dc->an ? or_density_aniso_d(g, dc->an, x) : or_density_d(g, x);
            cnt->samples++;
            const float xf[3] = { (float)x[0], (float)x[1], (float)x[2] };
            count_sample(dc, cnt, xf, 0);
            if (rho >= (double)s->iso) dhi = tm; else dlo = tm;
            double tt = (dlo + dhi) * 0.5;
            for (int a = 0; a < 3; a++) xs[a] = (float)((double)o[a] + tt * (double)d[a]);
            th = (float)tt;
        } else {
            th = (lo + hi) * 0.5f;
            for (int a = 0; a < 3; a++) xs[a] = o[a] + th * d[a];
        }
    }

    float n[3] = { 0.0f, 0.0f, 0.0f };
    if (s->normal_mode == 0) {
        /* The six samples' candidates are counted as the GPU's fused walk
         * reads them (or_cdiff_union); their densities are the single-sample
         * ones.  The hashed search keeps one count per sample. */
        uint32_t ucand;
        const int fused = !dc->hs && !dc->an && or_cdiff_union(dc->g, xs, s->normal_eps, &ucand, NULL, 0, NULL);
        or_counters one;
        memset(&one, 0, sizeof one);
        or_counters *c6 = fused ? &one : cnt;
        float gv[3];
        for (int a = 0; a < 3; a++) {
            float xp[3] = { xs[0], xs[1], xs[2] }, xm[3] = { xs[0], xs[1], xs[2] };
            xp[a] = xs[a] + s->normal_eps;
            xm[a] = xs[a] - s->normal_eps;
            float rp = dens_f(dc, xp, c6);
            float rm = dens_f(dc, xm, c6);
            count_sample(dc, cnt, xp, 0);
            count_sample(dc, cnt, xm, 0);
            gv[a] = rp - rm;
        }
        if (fused) {
            cnt->samples += 6;
            cnt->cand += ucand;
        }
        float dd = (gv[0] * gv[0] + gv[1] * gv[1]) + gv[2] * gv[2];
        if (dd > 0.0f) {
            float inv = 1.0f / sqrtf(dd);
            for (int a = 0; a < 3; a++) n[a] = -gv[a] * inv;
        }
    } else {
        float gs[3];
        sum_gradW(dc, xs, gs, cnt);
        normalize3(gs, n);
    }
    P[0] = xs[0]; P[1] = xs[1]; P[2] = xs[2]; P[3] = 1.0f;
    N[0] = n[0]; N[1] = n[1]; N[2] = n[2]; N[3] = 1.0f;
    if (tout) tout[idx] = th;
}

/* One pixel of the reference march, PerPixel_Isotropic (RayMarcher.cpp:256-344). */
static void pixel_reference(dens_ctx *dc, const or_settings *s, const float M[16], const float cam[3],
                            uint32_t W, uint32_t H, const float *depth, uint32_t idx, float *pos,
                            float *nrm, float *tout, or_counters *cnt)
{
    const or_grid *g = dc->g;
    float *P = pos + 4 * (uint64_t)idx, *N = nrm + 4 * (uint64_t)idx;
    float z = depth[idx];
    P[0] = P[1] = P[2] = P[3] = 0.0f;
    N[0] = N[1] = N[2] = N[3] = 0.0f;
    if (tout) tout[idx] = -1.0f;
    cnt->rays++;
    if (z == 1.0f) return;                                  /* :264 */
    cnt->rays_in_aabb++;

    float cx, cy, p[3], dun[3], dn[3], st[3];
    pixel_clip(idx, W, H, &cx, &cy);
    unproject(M, cx, cy, z, p);                             /* :268-273 */
    for (int a = 0; a < 3; a++) dun[a] = p[a] - cam[a];
    normalize3(dun, dn);
    for (int a = 0; a < 3; a++) st[a] = s->step * dn[a];    /* :274 */

    for (int i = 0; i < s->max_steps; i++) {
        for (int a = 0; a < 3; a++) p[a] += st[a];          /* :279 */
        if (s->skip_mode == 2 || s->skip_mode == 3) {       /* :284-304 */
            const uint8_t *flag = s->skip_mode == 2 ? g->dg_flag_box : g->dg_flag_cons;
            float nmin[3], nmax[3];
            int64_t id;
            while ((id = dg_query(g, p, nmin, nmax)) >= 0 && !flag[id]) {
                float q[3];
                intersect_aabb(p, st, nmin, nmax, q);
                for (int a = 0; a < 3; a++) p[a] = q[a] + st[a];
            }
        }
        float rho = dens_f(dc, p, cnt);
        cnt->march_samples++;
        if (rho >= s->iso) {                                /* :327 */
            float n[3] = { 0.0f, 0.0f, 0.0f };
            if (s->normal_mode == 1) {
                float gs[3];
                sum_gradW(dc, p, gs, cnt);
                normalize3(gs, n);                          /* :338 */
            } else {
                float gv[3];
                for (int a = 0; a < 3; a++) {
                    float xp[3] = { p[0], p[1], p[2] }, xm[3] = { p[0], p[1], p[2] };
                    xp[a] = p[a] + s->normal_eps;
                    xm[a] = p[a] - s->normal_eps;
                    gv[a] = dens_f(dc, xp, cnt) - dens_f(dc, xm, cnt);
                }
                float dd = (gv[0] * gv[0] + gv[1] * gv[1]) + gv[2] * gv[2];
                if (dd > 0.0f) {
                    float inv = 1.0f / sqrtf(dd);
                    for (int a = 0; a < 3; a++) n[a] = -gv[a] * inv;
                }
            }
            cnt->hits++;
            P[0] = p[0]; P[1] = p[1]; P[2] = p[2]; P[3] = 1.0f;
            N[0] = n[0]; N[1] = n[1]; N[2] = n[2]; N[3] = 1.0f;
            if (tout) {
                float dv[3] = { p[0] - cam[0], p[1] - cam[1], p[2] - cam[2] };
                tout[idx] = sqrtf((dv[0] * dv[0] + dv[1] * dv[1]) + dv[2] * dv[2]);
            }
            return;
        }
    }
}

/* ------------------------------------------------------------------------- */
/* Work scheduler: the reference's pool hands out one pixel per atomic        */
/* fetch_add (ThreadPool.cpp:38-55).  Same pattern, with every pixel covered  */
/* and completion defined by joining the workers.                             */
/* ------------------------------------------------------------------------- */

typedef struct {
    const or_grid *g;
    const or_hash *hs;
    const or_settings *s;
    const float *M, *cam, *depth;
    uint32_t W, H, row_step, nrows;
    float *pos, *nrm, *tout;
    _Atomic uint64_t next;
    uint64_t total;
    pthread_mutex_t lock;
    or_counters cnt;
} march_job;

static void *march_worker(void *arg)
{
    march_job *j = (march_job *)arg;
    dens_ctx dc;
    memset(&dc, 0, sizeof(dc));
    dc.g = j->g;
    dc.hs = j->hs;
    dc.h2 = j->g->h * j->g->h;
    dc.hinv = 1.0f / j->g->h;
    dc.sig = or_sigma_f(j->g->h);
    dc.compensated = j->s->compensated;
    dc.an = j->s->anisotropic ? j->s : NULL;
    dc.nbuf = j->hs ? (uint32_t *)malloc(OR_MAX_NEIGHBORS * sizeof(uint32_t)) : NULL;
    or_counters c;
    memset(&c, 0, sizeof(c));
    uint64_t i;
    while ((i = atomic_fetch_add(&j->next, 1)) < j->total) {
        uint32_t row = (uint32_t)(i / j->W) * j->row_step, col = (uint32_t)(i % j->W);
        uint32_t idx = row * j->W + col;
        if (j->s->mode == OR_MODE_REFERENCE)
            pixel_reference(&dc, j->s, j->M, j->cam, j->W, j->H, j->depth, idx, j->pos, j->nrm, j->tout, &c);
        else
            pixel_config(&dc, j->s, j->M, j->cam, j->W, j->H, idx, j->pos, j->nrm, j->tout, &c);
    }
    free(dc.nbuf);
    pthread_mutex_lock(&j->lock);
    j->cnt.rays += c.rays; j->cnt.rays_in_aabb += c.rays_in_aabb; j->cnt.hits += c.hits;
    j->cnt.samples += c.samples; j->cnt.cand += c.cand; j->cnt.march_samples += c.march_samples;
    j->cnt.march_cand += c.march_cand; j->cnt.cand_stencil += c.cand_stencil; j->cnt.cand27 += c.cand27;
    pthread_mutex_unlock(&j->lock);
    return NULL;
}

static int run_march(const or_grid *g, const or_hash *hs, const or_settings *s, const float M[16],
                     const float cam[3], uint32_t W, uint32_t H, const float *depth, uint32_t row_step,
                     int nthreads, float *pos, float *nrm, float *tout, or_counters *cnt)
{
    if (!g || !s || !M || !cam || !pos || !nrm || W == 0 || H == 0) return -1;
    if (s->mode == OR_MODE_REFERENCE && !depth) return -1;
    if (!(s->step > 0.0f)) return -1;
    /* or_bound_build first: for iso, or (anisotropic density) for OR_ZERO_ISO --
     * sub-boxes with no particle within h, where that density is exactly 0 */
    const float biso = s->anisotropic ? OR_ZERO_ISO : s->iso;
    if (s->skip_mode == 4 && (!g->bound || g->bound_iso != biso || g->bound_boxes != s->bound_boxes)) return -2;
    if (s->anisotropic && !(s->ext_mult >= 1.0f && s->ext_mult <= 2.0f)) return -1;
    if (row_step == 0) row_step = 1;
    march_job j;
    memset(&j, 0, sizeof(j));
    j.g = g; j.hs = hs; j.s = s; j.M = M; j.cam = cam; j.depth = depth;
    j.W = W; j.H = H; j.row_step = row_step;
    j.nrows = (H + row_step - 1) / row_step;
    j.pos = pos; j.nrm = nrm; j.tout = tout;
    j.total = (uint64_t)W * j.nrows;
    atomic_init(&j.next, 0);
    pthread_mutex_init(&j.lock, NULL);
    if (nthreads <= 1) {
        march_worker(&j);
    } else {
        pthread_t *th = (pthread_t *)malloc(sizeof(pthread_t) * (size_t)nthreads);
        for (int t = 0; t < nthreads; t++) pthread_create(&th[t], NULL, march_worker, &j);
        for (int t = 0; t < nthreads; t++) pthread_join(th[t], NULL);
        free(th);
    }
    pthread_mutex_destroy(&j.lock);
    if (cnt) *cnt = j.cnt;
    return 0;
}

int or_march(const or_grid *g, const or_settings *s, const float invpv[16], const float cam[3],
             uint32_t W, uint32_t H, const float *depth, uint32_t row_step, int nthreads,
             float *pos, float *nrm, float *tout, or_counters *cnt)
{
    return run_march(g, NULL, s, invpv, cam, W, H, depth, row_step, nthreads, pos, nrm, tout, cnt);
}

int or_march_hashed(const or_hash *hs, const or_grid *g, const or_settings *s,
                    const float invpv[16], const float cam[3], uint32_t W, uint32_t H,
                    uint32_t row_step, int nthreads, float *pos, float *nrm, float *tout,
                    or_counters *cnt)
{
    if (!hs || s->mode != OR_MODE_CONFIG || s->normal_mode != 0 || s->anisotropic) return -1;
    return run_march(g, hs, s, invpv, cam, W, H, NULL, row_step, nthreads, pos, nrm, tout, cnt);
}

/* ------------------------------------------------------------------------- */
/* FP64 configured march: the ground truth.  Same ray and sample definition   */
/* as the FP32 march, every operation in double, all bisection iterations in */
/* double, central-difference normals in double.                              */
/* ------------------------------------------------------------------------- */

typedef struct {
    const or_grid *g;
    const or_settings *s;
    const float *M, *cam;
    uint32_t W, H, row_step;
    double *pos, *nrm, *tout;
    _Atomic uint64_t next;
    uint64_t total;
} march64_job;

static double dens64(const or_grid *g, const or_settings *s, const double x[3])
{
    return s->anisotropic ? or_density_aniso_d(g, s, x) : or_density_d(g, x);
}

static void pixel_f64(const march64_job *j, uint32_t idx)
{
    const or_grid *g = j->g;
    const or_settings *s = j->s;
    double *P = j->pos + 4 * (uint64_t)idx, *N = j->nrm + 4 * (uint64_t)idx;
    for (int a = 0; a < 4; a++) P[a] = N[a] = 0.0;
    if (j->tout) j->tout[idx] = -1.0;
    double cx = (double)(idx % j->W) * (2.0 / (double)j->W) - 1.0;
    double cy = (double)(idx / j->W) * (2.0 / (double)j->H) - 1.0;
    double w[4];
    for (int r = 0; r < 4; r++)
        w[r] = ((double)j->M[r] * cx + (double)j->M[4 + r] * cy) +
               ((double)j->M[8 + r] * 1.0 + (double)j->M[12 + r]);
    double o[3] = { j->cam[0], j->cam[1], j->cam[2] }, d[3];
    double dd = 0.0;
    for (int a = 0; a < 3; a++) { d[a] = w[a] / w[3] - o[a]; dd += d[a] * d[a]; }
    double inv = 1.0 / sqrt(dd);
    for (int a = 0; a < 3; a++) d[a] *= inv;
    double tn = 0.0, tf = INFINITY;
    for (int a = 0; a < 3; a++) {
        if (d[a] != 0.0) {
            double t0 = ((double)g->bmin[a] - o[a]) / d[a], t1 = ((double)g->bmax[a] - o[a]) / d[a];
            double lo = t0 < t1 ? t0 : t1, hi = t0 < t1 ? t1 : t0;
            if (lo > tn) tn = lo;
            if (hi < tf) tf = hi;
        } else if (o[a] < g->bmin[a] || o[a] > g->bmax[a]) return;
    }
    if (tn > tf) return;
    double step = s->step, iso = s->iso;
    int64_t k, hitk = -1;
    for (k = 0;; k++) {
        if (s->max_steps > 0 && k >= s->max_steps) break;
        double t = tn + (double)k * step;
        if (t > tf) break;
        double x[3] = { o[0] + t * d[0], o[1] + t * d[1], o[2] + t * d[2] };
        if (dens64(g, s, x) >= iso) { hitk = k; break; }
    }
    if (hitk < 0) return;
    double th;
    if (hitk == 0) th = tn;
    else {
        double lo = tn + (double)(hitk - 1) * step, hi = tn + (double)hitk * step;
        for (int it = 0; it < s->bisect_iters; it++) {
            double tm = (lo + hi) * 0.5;
            double x[3] = { o[0] + tm * d[0], o[1] + tm * d[1], o[2] + tm * d[2] };
            if (dens64(g, s, x) >= iso) hi = tm; else lo = tm;
        }
        th = (lo + hi) * 0.5;
    }
    double xs[3] = { o[0] + th * d[0], o[1] + th * d[1], o[2] + th * d[2] };
    double gv[3];
    for (int a = 0; a < 3; a++) {
        double xp[3] = { xs[0], xs[1], xs[2] }, xm[3] = { xs[0], xs[1], xs[2] };
        xp[a] += s->normal_eps;
        xm[a] -= s->normal_eps;
        gv[a] = dens64(g, s, xp) - dens64(g, s, xm);
    }
    double gg = (gv[0] * gv[0] + gv[1] * gv[1]) + gv[2] * gv[2];
    double n[3] = { 0, 0, 0 };
    if (gg > 0) for (int a = 0; a < 3; a++) n[a] = -gv[a] / sqrt(gg);
    for (int a = 0; a < 3; a++) { P[a] = xs[a]; N[a] = n[a]; }
    P[3] = N[3] = 1.0;
    if (j->tout) j->tout[idx] = th;
}

static void *march64_worker(void *arg)
{
    march64_job *j = (march64_job *)arg;
    uint64_t i;
    while ((i = atomic_fetch_add(&j->next, 1)) < j->total) {
        uint32_t row = (uint32_t)(i / j->W) * j->row_step, col = (uint32_t)(i % j->W);
        pixel_f64(j, row * j->W + col);
    }
    return NULL;
}

int or_march_f64(const or_grid *g, const or_settings *s, const float invpv[16], const float cam[3],
                 uint32_t W, uint32_t H, uint32_t row_step, int nthreads,
                 double *pos, double *nrm, double *tout)
{
    if (!g || !s || s->mode != OR_MODE_CONFIG || !(s->step > 0.0f) || W == 0 || H == 0) return -1;
    if (s->anisotropic && !(s->ext_mult >= 1.0f && s->ext_mult <= 2.0f)) return -1;
    if (row_step == 0) row_step = 1;
    march64_job j;
    memset(&j, 0, sizeof(j));
    j.g = g; j.s = s; j.M = invpv; j.cam = cam; j.W = W; j.H = H; j.row_step = row_step;
    j.pos = pos; j.nrm = nrm; j.tout = tout;
    j.total = (uint64_t)W * ((H + row_step - 1) / row_step);
    atomic_init(&j.next, 0);
    if (nthreads <= 1) march64_worker(&j);
    else {
        pthread_t *th = (pthread_t *)malloc(sizeof(pthread_t) * (size_t)nthreads);
        for (int t = 0; t < nthreads; t++) pthread_create(&th[t], NULL, march64_worker, &j);
        for (int t = 0; t < nthreads; t++) pthread_join(th[t], NULL);
        free(th);
    }
    return 0;
}

/* ------------------------------------------------------------------------- */
/* Screen-space depth smoothing -- GaussRenderPass: ComputeGaussKernel        */
/* (GaussRenderPass.cpp:8-48) and gauss.frag's window sum, clamp-to-edge      */
/* texel-centre taps (BilateralBuffer.cpp:246-248, Spread = 1); each term is  */
/* accumulated with one fused multiply-add (`sum += k * t` in gauss.frag has  */
/* no `precise` qualifier, so the shader compiler may fuse it).               */
/* ------------------------------------------------------------------------- */

int or_gauss_kernel(int32_t N, float *K)
{
    if (!K || N < 0 || N > 31) return -1;
    const float E = 2.7182818284590452353602874713527f;     /* GaussRenderPass.cpp:10 */
    float sum = 0.0f;
    for (int i = 0; i <= N; i++)
        for (int j = 0; j <= N; j++) {
            float r = sqrtf((float)i * (float)i + (float)j * (float)j);   /* :17-20 */
            float y = powf(E, -0.5f * r * r);                             /* :12-15 */
            K[j + i * (N + 1)] = y;
            if (i == 0 && j == 0) sum += y;
            else if (i != 0 && j != 0) sum += 4.0f * y;
            else sum += 2.0f * y;
        }
    for (int k = 0; k < (N + 1) * (N + 1); k++) K[k] /= sum;
    return 0;
}

int or_depth_smooth(const float *depth, uint32_t W, uint32_t H, int32_t N, float *out)
{
    float K[32 * 32];
    if (!depth || !out || W == 0 || H == 0 || or_gauss_kernel(N, K)) return -1;
    for (uint32_t y = 0; y < H; y++)
        for (uint32_t x = 0; x < W; x++) {
            float s = 0.0f;
            for (int i = -N; i <= N; i++) {                 /* gauss.frag: i outer (u), j inner (v) */
                int gx = (int)x + i;
                gx = gx < 0 ? 0 : (gx > (int)W - 1 ? (int)W - 1 : gx);
                for (int j = -N; j <= N; j++) {
                    int gy = (int)y + j;
                    gy = gy < 0 ? 0 : (gy > (int)H - 1 ? (int)H - 1 : gy);
                    s = fmaf(K[(i < 0 ? -i : i) * (N + 1) + (j < 0 ? -j : j)], depth[(size_t)gy * W + (size_t)gx], s);
                }
            }
            out[(size_t)y * W + x] = s;
        }
    return 0;
}

/* ------------------------------------------------------------------------- */
/* Depth sprites -- DepthRenderPass (DepthRenderPass.cpp:243-250), the sprite */
/* quads of AdvancedRenderer::CollectRenderData (AdvancedRenderer.cpp:455-474)*/
/* and depth.vert / depth.frag, depth test LESS against a clear of 1.0: every */
/* particle a camera-facing quad p + R (+-x +-y), x and y the first two rows  */
/* of View (CameraController3D.cpp:80).  At the centre of pixel (px, py) the  */
/* pixel's ray meets the quad's plane at P; UV = ((P - p).x, -(P - p).y) / R  */
/* (the perspective-correct interpolation of the vertex UVs over the planar   */
/* quad); discard when |UV|^2 > 1; depth = projected z of                     */
/* View P - R (0, 0, cos(pi/2 |UV|)).  Plane depths outside [0, 1] are        */
/* clipped, written depths clamped to [0, 1]; a particle with a quad corner   */
/* behind the camera (w <= 0) is not drawn.  Each pixel keeps the least depth.*/
/* cos(pi/2 s) is its Taylor polynomial through s^12.                         */
/* ------------------------------------------------------------------------- */

static void sp_xform(const float *M, float x, float y, float z, float o[4])
{
    for (int r = 0; r < 4; r++) o[r] = (M[r] * x + M[4 + r] * y) + (M[8 + r] * z + M[12 + r] * 1.0f);
}

float or_cos_half_pi(float s)
{
    const float t = 1.57079632679f * s, z = t * t;
    float c = 1.0f / 479001600.0f;
    c = c * z + -1.0f / 3628800.0f;
    c = c * z + 1.0f / 40320.0f;
    c = c * z + -1.0f / 720.0f;
    c = c * z + 1.0f / 24.0f;
    c = c * z + -0.5f;
    return c * z + 1.0f;
}

int or_depth_sprites(const float *xyz, uint64_t n, const float view[16], const float proj[16], const float inv_pv[16],
                     const float cam[3], float R, uint32_t W, uint32_t H, float *depth)
{
    if (!view || !proj || !inv_pv || !cam || !depth || W == 0 || H == 0 || !(R > 0.0f)) return -1;
    uint32_t *bits = (uint32_t *)depth;
    for (uint64_t k = 0; k < (uint64_t)W * H; k++) bits[k] = 0x3f800000u;
    const float *V = view, *P = proj;
    const float ax = V[0], ay = V[4], az = V[8], bx = V[1], by = V[5], bz = V[9];
    const float nx = ay * bz - az * by, ny = az * bx - ax * bz, nz = ax * by - ay * bx;
    const float o0 = cam[0], o1 = cam[1], o2 = cam[2];
    const float invR = 1.0f / R;
    for (uint64_t i = 0; i < n; i++) {
        const float qx = xyz[3 * i], qy = xyz[3 * i + 1], qz = xyz[3 * i + 2];
        float lo_x = 3.0e38f, hi_x = -3.0e38f, lo_y = 3.0e38f, hi_y = -3.0e38f;
        int drawn = 1;
        for (int k = 0; k < 4; k++) {
            const float sx = (k & 1) ? 1.0f : -1.0f, sy = (k & 2) ? 1.0f : -1.0f;
            const float cx = qx + R * (sx * ax + sy * bx), cy = qy + R * (sx * ay + sy * by),
                        cz = qz + R * (sx * az + sy * bz);
            float v[4], c[4];
            sp_xform(V, cx, cy, cz, v);
            sp_xform(P, v[0], v[1], v[2], c);
            if (!(c[3] > 0.0f)) { drawn = 0; break; }
            const float fx = (c[0] / c[3] + 1.0f) * 0.5f * (float)W, fy = (c[1] / c[3] + 1.0f) * 0.5f * (float)H;
            lo_x = fminf(lo_x, fx); hi_x = fmaxf(hi_x, fx);
            lo_y = fminf(lo_y, fy); hi_y = fmaxf(hi_y, fy);
        }
        if (!drawn) continue;
        const float x0f = fmaxf(ceilf(lo_x - 0.5f), 0.0f), x1f = fminf(floorf(hi_x - 0.5f), (float)W - 1.0f);
        const float y0f = fmaxf(ceilf(lo_y - 0.5f), 0.0f), y1f = fminf(floorf(hi_y - 0.5f), (float)H - 1.0f);
        if (!(x0f <= x1f && y0f <= y1f)) continue;
        const float pn = ((qx - o0) * nx + (qy - o1) * ny) + (qz - o2) * nz;
        for (int py = (int)y0f; py <= (int)y1f; py++)
            for (int px = (int)x0f; px <= (int)x1f; px++) {
                const float ndx = ((float)px + 0.5f) * (2.0f / (float)W) - 1.0f;
                const float ndy = ((float)py + 0.5f) * (2.0f / (float)H) - 1.0f;
                float w[4];
                sp_xform(inv_pv, ndx, ndy, 1.0f, w);
                /* the ray direction scaled by w (no division): t d is unchanged; the
                 * far point of a perspective projection has w = 1 / far > 0 */
                if (!(w[3] > 0.0f)) continue;
                const float d0 = w[0] - o0 * w[3], d1 = w[1] - o1 * w[3], d2 = w[2] - o2 * w[3];
                const float den = (d0 * nx + d1 * ny) + d2 * nz;
                if (den == 0.0f) continue;
                const float t = pn / den;
                if (!(t > 0.0f)) continue;
                const float X = o0 + t * d0, Y = o1 + t * d1, Z = o2 + t * d2;
                const float rx = X - qx, ry = Y - qy, rz = Z - qz;
                const float u = ((rx * ax + ry * ay) + rz * az) * invR;
                const float v = -(((rx * bx + ry * by) + rz * bz) * invR);
                const float l2 = u * u + v * v;
                if (l2 > 1.0f) continue;
                float vp[4], cp[4], cs[4];
                sp_xform(V, X, Y, Z, vp);
                sp_xform(P, vp[0], vp[1], vp[2], cp);
                const float zp = cp[2] / cp[3];
                if (!(zp >= 0.0f && zp <= 1.0f)) continue;
                const float off = or_cos_half_pi(sqrtf(l2));
                sp_xform(P, vp[0], vp[1], vp[2] - R * off, cs);
                float z = cs[2] / cs[3];
                if (z != z) continue;
                z = z <= 0.0f ? 0.0f : (z > 1.0f ? 1.0f : z);
                uint32_t zb;
                memcpy(&zb, &z, 4);
                uint32_t *dst = &bits[(size_t)py * W + (size_t)px];
                if (zb < *dst) *dst = zb;
            }
    }
    return 0;
}
