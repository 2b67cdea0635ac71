// pm_march.hip -- the per-pixel iso-surface march on CDNA4.
//
// Replaces RayMarcher::PerPixel_Isotropic (RayMarcher.cpp:256-344) and the
// ThreadPool that drives it (ThreadPool.cpp:38-55).
//
// Work mapping: one wave64 per 8x8 pixel tile at a time, one ray per lane, four
// waves per 256-thread workgroup; persistent workgroups take the tiles from
// per-XCD bands (blockIdx % 8 names the XCD), so each XCD's L2 holds the
// particles of its band.
//
// Every lane runs a small state machine (march -> FP32 bisection; the learned
// form adds step proposals) and each loop iteration performs exactly one FP32
// density evaluation per busy lane, whatever its phase, so the candidate loop
// stays converged across the wave while rays hit at different steps.  A ray's
// march ends at its hit record; the FP64 step, the normal and the outputs
// follow per tile (finish_hit).  Schedule 1 (k_march_simple) marches one ray
// per lane to its end with a direct walk for every sample: the plain form the
// persistent one is checked against.
#include "pm_features.h"
#include "pm_internal.h"
#include "pm_ray.h"

namespace pm {

// The ray's hit is known up to what k_finish does: its record (at the
// pixel) holds t itself (fp64 = false), or the bracket [lo, hi] of the FP64
// bisection step.  Record: (1 | fp64 << 1, lo, hi, 0); a miss leaves zeros.
// lead: this lane writes (the others of a lane group share the ray).
__device__ __forceinline__ void to_emit(const MarchLaunch &L, RayState &r, bool fp64, float lo, float hi,
                                       bool lead = true)
{
    if (lead) L.hit_rec[r.idx] = make_uint4(fp64 ? 3u : 1u, __float_as_uint(lo), __float_as_uint(hi), 0u);
    r.phase = PH_DONE;
}

// The first sample k at or above iso is known: bracket [t_{k-1}, t_k] and
// start the bisection (or take t_0 when the ray starts inside the surface).
__device__ __forceinline__ void start_bracket(const MarchLaunch &L, RayState &r, int n32, bool lead = true)
{
    const pm_settings &s = L.s;
    if (r.k == 0) {
        to_emit(L, r, false, r.tn, r.tn, lead);
        return;
    }
    r.lo = r.tn + (float)(r.k - 1) * s.step;
    r.hi = r.tn + (float)r.k * s.step;
    r.it = 0;
    if (n32 > 0) r.phase = PH_BISECT;
    else if (s.final_fp64 && s.bisect_iters > 0) to_emit(L, r, true, r.lo, r.hi, lead);
    else { const float t = (r.lo + r.hi) * 0.5f; to_emit(L, r, false, t, t, lead); }
}

// First half of one step: the sample point of the lane's next FP32 density
// evaluation, after the skips of the march phase; false when the lane takes
// no sample this step (the ray ended, or it used up its HOPS skips).  Past an
// exactly skipped stretch the learned march may ask for proposals again.
template <int HOPS>
__device__ __forceinline__ bool ray_prepare(const MarchLaunch &L, RayState &r, float &qx, float &qy, float &qz)
{
    const pm_settings &s = L.s;
    const GridDev &g = L.g;
    const float o0 = L.cam[0], o1 = L.cam[1], o2 = L.cam[2];

    if (r.phase == PH_MARCH) {
        // at most HOPS skips per step: a lane crossing much skippable space
        // then idles for a step instead of holding the whole wave in this loop
        // (a skip costs a few dozen instructions, a step a density walk)
        for (int hop = 0;; hop++) {
            float t = r.tn + (float)r.k * s.step;
            if (t > r.tf || (s.max_steps > 0 && r.k >= s.max_steps)) {
                march_miss(L, r.idx);
                r.phase = PH_DONE;
                return false;
            }
            qx = o0 + t * r.d0; qy = o1 + t * r.d1; qz = o2 + t * r.d2;
            int kn;
            if (s.skip_mode == PM_SKIP_BOUND) kn = pm_skip_bound(g, o0, o1, o2, r, qx, qy, qz, L.inv_step);
            else if (s.skip_mode == PM_SKIP_EMPTY) kn = pm_skip_empty(g, o0, o1, o2, r, qx, qy, qz, s.step);
            else break;
            if (kn == r.k) break;
            r.k = kn;
            r.lrn = s.learned;
            if (hop == HOPS - 1) return false;
        }
    } else {  // PH_BISECT
        float tm = (r.lo + r.hi) * 0.5f;
        qx = o0 + tm * r.d0; qy = o1 + tm * r.d1; qz = o2 + tm * r.d2;
    }
    return true;
}

// The counts of one FP32 density evaluation at q (instrumented launches also
// count its own stencil and its 27 search cells).
template <bool INSTR>
__device__ __forceinline__ void count_sample(const GridDev &g, float qx, float qy, float qz, uint32_t cand, bool march,
                                             Counts &cn)
{
    cn.samples++;
    cn.cand += cand;
    if (march) {
        cn.march++;
        cn.mcand += cand;
    }
    if (INSTR) {
        cn.cst += cand;
        cn.c27 += pm_cand27(g, qx, qy, qz);
    }
}

// Second half: the sample's density rho (and its stencil candidate count)
// drives the lane's state machine.
template <bool INSTR>
__device__ __forceinline__ void ray_finish(const MarchLaunch &L, RayState &r, float rho, uint32_t cand, float qx,
                                           float qy, float qz, Counts &cn)
{
    const pm_settings &s = L.s;
    const bool fp64 = s.final_fp64 && s.bisect_iters > 0;
    const int n32 = fp64 ? s.bisect_iters - 1 : s.bisect_iters;
    count_sample<INSTR>(L.g, qx, qy, qz, cand, r.phase == PH_MARCH, cn);

    if (r.phase == PH_MARCH) {
        if (rho >= s.iso) {
            start_bracket(L, r, n32);
        } else if (r.lrn && rho < L.mlp_gate_rho) {
            r.rho = rho;
            r.phase = PH_PROPOSE;
        } else {
            r.k++;
        }
    } else {  // PH_BISECT
        float tm = (r.lo + r.hi) * 0.5f;
        if (rho >= s.iso) r.hi = tm; else r.lo = tm;
        r.it++;
        if (r.it >= n32) {
            if (fp64) to_emit(L, r, true, r.lo, r.hi);
            else { const float t = (r.lo + r.hi) * 0.5f; to_emit(L, r, false, t, t); }
        }
    }
}

// One density evaluation of this ray, whatever its phase, and the transition
// it implies.  Dens(q, cand) is the FP32 density.  HOPS: skips per step before
// a lane idles -- 16 in the single-view march, whose end is spent on few long
// rays with idle lanes beside them; 8 in the batch, whose lanes stay busy.
template <bool INSTR, int HOPS, class Dens>
__device__ __forceinline__ void ray_step(const MarchLaunch &L, RayState &r, Dens dens, Counts &cn)
{
    float qx, qy, qz;
    if (!ray_prepare<HOPS>(L, r, qx, qy, qz)) return;
    uint32_t cand = 0;
    float rho = dens(qx, qy, qz, cand);
    ray_finish<INSTR>(L, r, rho, cand, qx, qy, qz, cn);
}

// ---- lane groups: the end of a persistent launch ----
// Once no tile is left, a wave's remaining rays are spread over its idle lanes:
// G lanes (a power of two) per ray, every lane of a group holding the same
// RayState.  A march step then evaluates the G lattice samples k .. k+G-1 at
// once and takes the first at or above iso; a bisection step evaluates the
// 2^L - 1 midpoints of the next L levels (L = floor(log2(G + 1))) and descends
// through them.  Each sample is computed as the one-lane march computes it,
// and a sample the skip would pass has its density below iso, so the first
// crossing, the bracket and the image are those of the one-lane march.
// Counters count what the one-lane march evaluates: the samples the skip
// keeps, up to the crossing, and the bisection's path.

// The ray's state to LDS and back (region: 64 floats per field).
__device__ __forceinline__ void state_put(float *region, uint32_t slot, const RayState &r)
{
    float *p = region + slot;
    p[0 * 64] = r.d0; p[1 * 64] = r.d1; p[2 * 64] = r.d2; p[3 * 64] = r.tn; p[4 * 64] = r.tf;
    p[5 * 64] = r.lo; p[6 * 64] = r.hi; p[7 * 64] = __int_as_float(r.k); p[8 * 64] = __int_as_float(r.it);
    p[9 * 64] = __int_as_float(r.phase); p[10 * 64] = __uint_as_float(r.idx); p[11 * 64] = r.id0;
    p[12 * 64] = r.id1; p[13 * 64] = r.id2;
}

__device__ __forceinline__ void state_get(const float *region, uint32_t slot, RayState &r)
{
    const float *p = region + slot;
    r.d0 = p[0 * 64]; r.d1 = p[1 * 64]; r.d2 = p[2 * 64]; r.tn = p[3 * 64]; r.tf = p[4 * 64];
    r.lo = p[5 * 64]; r.hi = p[6 * 64]; r.k = __float_as_int(p[7 * 64]); r.it = __float_as_int(p[8 * 64]);
    r.phase = __float_as_int(p[9 * 64]); r.idx = __float_as_uint(p[10 * 64]); r.id0 = p[11 * 64];
    r.id1 = p[12 * 64]; r.id2 = p[13 * 64];
    r.lrn = 0;
}

// Widens the groups when the busy rays fit in fewer lanes: G becomes the
// largest power of two with (busy rays) x G <= 64.  Called by all 64 lanes.
__device__ __forceinline__ void regroup(RayState &r, int &G, float *region)
{
    const uint32_t lane = threadIdx.x & 63;
    const unsigned long long lead = __ballot(r.phase != PH_DONE && (lane & (uint32_t)(G - 1)) == 0);
    const int nb = __popcll(lead);
    int Gn = 1;
    while (Gn < 64 && nb * Gn * 2 <= 64) Gn *= 2;
    if (Gn <= G || nb == 0) return;
    if ((lead >> lane) & 1ull) state_put(region, (uint32_t)__popcll(lead & ((1ull << lane) - 1ull)), r);
    wave_lds_sync();
    if ((int)(lane / (uint32_t)Gn) < nb) state_get(region, lane / (uint32_t)Gn, r);
    else r.phase = PH_DONE;
    wave_lds_sync();
    G = Gn;
}

// Lane j of its group: its sample of this step; false when it takes none.
// nonskip: the sample is one the one-lane march would evaluate.
__device__ __forceinline__ bool group_prepare(const MarchLaunch &L, RayState &r, int G, int j, float &qx, float &qy,
                                              float &qz, bool &nonskip)
{
    const pm_settings &s = L.s;
    const GridDev &g = L.g;
    const float o0 = L.cam[0], o1 = L.cam[1], o2 = L.cam[2];
    nonskip = false;
    if (r.phase == PH_MARCH) {
        // the group's first sample, as ray_prepare finds it; the wave has few
        // rays left, so a group may take 8 G skips in one step
        const int hop_cap = 8 * G - 1;
        for (int hop = 0;; hop++) {
            float t = r.tn + (float)r.k * s.step;
            if (t > r.tf || (s.max_steps > 0 && r.k >= s.max_steps)) {
                if (j == 0) march_miss(L, r.idx);
                r.phase = PH_DONE;
                return false;
            }
            qx = o0 + t * r.d0; qy = o1 + t * r.d1; qz = o2 + t * r.d2;
            int kn = r.k;
            if (s.skip_mode == PM_SKIP_BOUND) kn = pm_skip_bound(g, o0, o1, o2, r, qx, qy, qz, L.inv_step);
            if (kn == r.k) break;
            r.k = kn;
            if (hop == hop_cap) return false;
        }
        if (j == 0) { nonskip = true; return true; }
        const int kj = r.k + j;
        const float t = r.tn + (float)kj * s.step;
        if (t > r.tf || (s.max_steps > 0 && kj >= s.max_steps)) return false;
        qx = o0 + t * r.d0; qy = o1 + t * r.d1; qz = o2 + t * r.d2;
        nonskip = s.skip_mode != PM_SKIP_BOUND || pm_skip_bound(g, o0, o1, o2, r, qx, qy, qz, L.inv_step) == r.k;
        return true;
    }
    if (r.phase == PH_BISECT) {
        const bool fp64 = s.final_fp64 && s.bisect_iters > 0;
        const int n32 = fp64 ? s.bisect_iters - 1 : s.bisect_iters;
        int lv = 31 - __clz(G + 1);                 // floor(log2(G + 1))
        if (lv > n32 - r.it) lv = n32 - r.it;
        const int node = j + 1;                     // heap order: 1 = the next midpoint
        if (node >= (1 << lv)) return false;
        const int depth = 31 - __clz(node);
        float lo = r.lo, hi = r.hi;
        for (int b = depth - 1; b >= 0; b--) {      // left child: the crossing lies below the midpoint
            const float tm = (lo + hi) * 0.5f;
            if ((node >> b) & 1) lo = tm; else hi = tm;
        }
        const float tm = (lo + hi) * 0.5f;
        qx = o0 + tm * r.d0; qy = o1 + tm * r.d1; qz = o2 + tm * r.d2;
        nonskip = true;
        return true;
    }
    return false;
}

// The group's transition from its lanes' densities.  Called by all 64 lanes.
template <bool INSTR>
__device__ __forceinline__ void group_finish(const MarchLaunch &L, RayState &r, int G, int j, bool want, bool nonskip,
                                             float rho, uint32_t cand, float qx, float qy, float qz, Counts &cn)
{
    const pm_settings &s = L.s;
    const bool fp64 = s.final_fp64 && s.bisect_iters > 0;
    const int n32 = fp64 ? s.bisect_iters - 1 : s.bisect_iters;
    const uint32_t lane = threadIdx.x & 63, gbase = lane & ~(uint32_t)(G - 1);
    const unsigned long long gmask = G == 64 ? ~0ull : ((1ull << G) - 1ull);
    const unsigned long long cross = (__ballot(want && rho >= s.iso) >> gbase) & gmask;
    const unsigned long long wants = (__ballot(want) >> gbase) & gmask;
    // bisection: descend through the evaluated midpoints (uniform trip count)
    const int lvmax = 31 - __clz(G + 1);
    const bool bis = r.phase == PH_BISECT && (wants & 1ull);
    int node = 1;
    for (int l = 0; l < lvmax; l++) {
        const float rn = __shfl(rho, (int)gbase + node - 1, 64);
        const uint32_t cn_n = __shfl(cand, (int)gbase + node - 1, 64);
        if (bis && r.it < n32 && ((wants >> (node - 1)) & 1ull)) {
            const float tm = (r.lo + r.hi) * 0.5f;
            if (j == 0)
                count_sample<INSTR>(L.g, L.cam[0] + tm * r.d0, L.cam[1] + tm * r.d1, L.cam[2] + tm * r.d2, cn_n, false,
                                    cn);
            if (rn >= s.iso) { r.hi = tm; node = 2 * node; } else { r.lo = tm; node = 2 * node + 1; }
            r.it++;
        }
    }
    if (bis && r.it >= n32) {
        if (fp64) to_emit(L, r, true, r.lo, r.hi, j == 0);
        else { const float t = (r.lo + r.hi) * 0.5f; to_emit(L, r, false, t, t, j == 0); }
        return;
    }
    if (r.phase != PH_MARCH || !(wants & 1ull)) return;
    const int jc = cross ? __ffsll((long long)cross) - 1 : G;
    if (want && nonskip && j <= jc) count_sample<INSTR>(L.g, qx, qy, qz, cand, true, cn);
    if (cross) {
        r.k += jc;
        start_bracket(L, r, n32, j == 0);
    } else {
        r.k += G;
    }
}

struct DensPlain {
    const GridDev *g;
    bool comp;
    __device__ float operator()(float x, float y, float z, uint32_t &cand) const
    {
        return pm_density32(*g, x, y, z, comp, cand);
    }
};

template <int CAP, int UNROLL>
struct DensLds {
    const GridDev *g;
    float *buf;
    __device__ float operator()(float x, float y, float z, uint32_t &cand) const
    {
        return pm_density32_lds<CAP, UNROLL>(*g, x, y, z, buf, cand);
    }
};

static constexpr int kLdsCap = 32;
static constexpr uint32_t kOwnTiles = 32;   // tiles a persistent wave finishes itself

// Band b of the T tiles is [T b / 8, T (b+1) / 8).
__device__ __forceinline__ uint32_t band_lo(uint32_t T, uint32_t b) { return (uint32_t)(((uint64_t)T * b) >> 3); }

// Tile pre-pass of the persistent schedules.  tile_buf = flag[T+1] | pos[T+1] |
// order[T+1] | scan scratch: flag = 1 for a tile one of whose rays enters the
// padded AABB (the ray set-up and slab test of ray_setup, bit for bit) -- in
// reference mode, one of whose pixels has a depth other than 1 -- pos
// its exclusive scan, order the active tiles in row order (order[T] = their
// count).  The persistent waves take only active tiles; k_miss_tiles writes
// the misses of the others.
__global__ __launch_bounds__(256) void k_tile_flags(MarchLaunch L, uint32_t T)
{
    // one wave per tile, one lane per ray
    const uint32_t t = blockIdx.x * 4 + (threadIdx.x >> 6), j = threadIdx.x & 63u;
    if (t > T) return;   // whole wave
    bool enters = false;
    if (t < T) {
        const uint32_t tiles_x = (L.W + 7) / 8;
        const uint32_t px = (t % tiles_x) * 8 + (j & 7u), py = (t / tiles_x) * 8 + (j >> 3);
        if (px < L.W && py < L.H && L.s.mode == PM_MODE_REFERENCE) {
            enters = L.depth[py * L.W + px] != 1.0f;   // the reference march starts where a sprite was drawn
        } else if (px < L.W && py < L.H) {
            const float cx = (float)px * (2.0f / (float)L.W) - 1.0f;
            const float cy = (float)py * (2.0f / (float)L.H) - 1.0f;
            float wx, wy, wz, d[3], tn, tf;
            pm_unproject(L.M, cx, cy, 1.0f, wx, wy, wz);
            pm_normalize(wx - L.cam[0], wy - L.cam[1], wz - L.cam[2], d[0], d[1], d[2]);
            const float o[3] = { L.cam[0], L.cam[1], L.cam[2] };
            enters = pm_slab(o, d, L.g.bmin, L.g.bmax, tn, tf);
        }
    }
    const unsigned long long m = __ballot(enters);
    if (j == 0) L.tile_buf[t] = m != 0ull ? 1u : 0u;   // flag[T] = 0 closes the scan
}

__global__ __launch_bounds__(256) void k_tile_compact(uint32_t *tile_buf, uint32_t T)
{
    const uint32_t t = blockIdx.x * 256 + threadIdx.x;
    if (t > T) return;
    const uint32_t *flag = tile_buf, *pos = tile_buf + (T + 1);
    uint32_t *order = tile_buf + 2 * (T + 1);
    if (t == T) order[T] = pos[T];
    else if (flag[t]) order[pos[t]] = t;
}

// Pixels of the inactive tiles are written as misses here.  (They are not
// counted on the device: one counter for every wave would serialise on a
// single address; the host knows the ray total.)
__global__ __launch_bounds__(256) void k_miss_tiles(MarchLaunch L, uint32_t T)
{
    const uint32_t t = blockIdx.x * 4 + (threadIdx.x >> 6), j = threadIdx.x & 63u;
    if (t >= T) return;   // whole wave
    if (L.tile_buf[t] == 0) {
        const uint32_t tiles_x = (L.W + 7) / 8;
        const uint32_t px = (t % tiles_x) * 8 + (j & 7u), py = (t / tiles_x) * 8 + (j >> 3);
        if (px < L.W && py < L.H) write_miss(L, py * L.W + px);
    }
}

// Schedule 1: one 8x8 tile per wave, one ray per lane to its end -- march,
// bisection, FP64 step, normal, outputs -- every sample with its own direct
// walk (no compaction, no lists, no hit records).  The plain form of the
// configured march the persistent schedule is checked against.
__global__ __launch_bounds__(256) void k_march_simple(MarchLaunch L)
{
    uint32_t px, py;
    Counts cn;
    const pm_settings &s = L.s;
    const GridDev &g = L.g;
    const bool comp = s.compensated != 0;
    if (pm_tile(L.W, L.H, px, py) && px < L.W && py < L.H) {
        RayState r{};
        cn.rays++;
        r.idx = py * L.W + px;
        float d[3];
        pm_ray_dir(L, px, py, d);
        r.d0 = d[0]; r.d1 = d[1]; r.d2 = d[2];
        r.id0 = 1.0f / d[0]; r.id1 = 1.0f / d[1]; r.id2 = 1.0f / d[2];
        const float o[3] = { L.cam[0], L.cam[1], L.cam[2] };
        bool hit = false;
        if (pm_slab(o, d, g.bmin, g.bmax, r.tn, r.tf)) {
            cn.in_aabb++;
            r.k = 0;
            for (;;) {
                const float t = r.tn + (float)r.k * s.step;
                if (t > r.tf || (s.max_steps > 0 && r.k >= s.max_steps)) break;
                const float qx = o[0] + t * d[0], qy = o[1] + t * d[1], qz = o[2] + t * d[2];
                int kn = r.k;
                if (s.skip_mode == PM_SKIP_BOUND) kn = pm_skip_bound(g, o[0], o[1], o[2], r, qx, qy, qz, L.inv_step);
                else if (s.skip_mode == PM_SKIP_EMPTY) kn = pm_skip_empty(g, o[0], o[1], o[2], r, qx, qy, qz, s.step);
                if (kn != r.k) { r.k = kn; continue; }
                uint32_t cand = 0;
                const float rho = pm_density32(g, qx, qy, qz, comp, cand);
                cn.samples++; cn.march++; cn.cand += cand; cn.mcand += cand;
                if (rho >= s.iso) { hit = true; break; }
                r.k++;
            }
        }
        if (!hit) {
            write_miss(L, r.idx);
        } else {
            const bool fp64 = s.final_fp64 && s.bisect_iters > 0;
            const int n32 = fp64 ? s.bisect_iters - 1 : s.bisect_iters;
            float lo = r.tn + (float)(r.k - 1) * s.step, hi = r.tn + (float)r.k * s.step;
            float xs[3], th;
            if (r.k == 0) {
                th = r.tn;
                xs[0] = o[0] + th * d[0]; xs[1] = o[1] + th * d[1]; xs[2] = o[2] + th * d[2];
            } else {
                for (int it = 0; it < n32; it++) {
                    const float tm = (lo + hi) * 0.5f;
                    uint32_t cand = 0;
                    const float rho = pm_density32(g, o[0] + tm * d[0], o[1] + tm * d[1], o[2] + tm * d[2], comp, cand);
                    cn.samples++; cn.cand += cand;
                    if (rho >= s.iso) hi = tm; else lo = tm;
                }
                if (fp64) {
                    const double tmd = ((double)lo + (double)hi) * 0.5;
                    const double rho = pm_density64(g, (double)o[0] + tmd * (double)d[0], (double)o[1] + tmd * (double)d[1],
                                                    (double)o[2] + tmd * (double)d[2]);
                    cn.samples++;
                    double dlo = lo, dhi = hi;
                    if (rho >= (double)s.iso) dhi = tmd; else dlo = tmd;
                    const double tt = (dlo + dhi) * 0.5;
#pragma unroll
                    for (int a = 0; a < 3; a++) xs[a] = (float)((double)o[a] + tt * (double)d[a]);
                    th = (float)tt;
                } else {
                    th = (lo + hi) * 0.5f;
                    xs[0] = o[0] + th * d[0]; xs[1] = o[1] + th * d[1]; xs[2] = o[2] + th * d[2];
                }
            }
            float nx, ny, nz;
            if (s.normal_mode == PM_NORMAL_GRADW) {
                float sx, sy, sz;
                uint32_t cand = 0;
                pm_sum_gradw(g, xs[0], xs[1], xs[2], sx, sy, sz, cand);
                cn.cand += cand;
                pm_normalize(sx, sy, sz, nx, ny, nz);
            } else {
                float gv[3];
                for (int j = 0; j < 6; j++) {
                    float q[3] = { xs[0], xs[1], xs[2] };
                    q[j >> 1] = (j & 1) ? xs[j >> 1] - s.normal_eps : xs[j >> 1] + s.normal_eps;
                    uint32_t cand = 0;
                    const float rho = pm_density32(g, q[0], q[1], q[2], comp, cand);
                    cn.samples++; cn.cand += cand;
                    if (j & 1) gv[j >> 1] -= rho; else gv[j >> 1] = rho;
                }
                const float dd = (gv[0] * gv[0] + gv[1] * gv[1]) + gv[2] * gv[2];
                nx = ny = nz = 0.0f;
                if (dd > 0.0f) {
                    const float inv = 1.0f / sqrtf(dd);
                    nx = -gv[0] * inv; ny = -gv[1] * inv; nz = -gv[2] * inv;
                }
            }
            L.pos[r.idx] = make_float4(xs[0], xs[1], xs[2], 1.0f);
            L.nrm[r.idx] = make_float4(nx, ny, nz, 1.0f);
            if (L.t) L.t[r.idx] = th;
            cn.hits++;
        }
    }
    flush_counts(L.counters, cn);
}

// ---- learned step proposal with certified jumps ----
// An upper bound U of the density over every point within `margin` of the
// segment [a, b]: the sum over the particles near the segment of W at their
// distance to the segment less the margin.  W falls with the distance, so
// rho(q) <= U for every such q.  The walk covers the march-grid cells near the
// segment's box (pm_box_begin); it stops once every lane's U reached lim.
// INFINITY when those cells leave the grid (nothing is certified).
__device__ __forceinline__ float pm_segment_bound(const GridDev &g, const float a[3], const float b[3], float margin,
                                                  float lim, uint32_t &cand)
{
    float lo[3], hi[3], u[3];
#pragma unroll
    for (int k = 0; k < 3; k++) {
        lo[k] = fminf(a[k], b[k]) - margin;
        hi[k] = fmaxf(a[k], b[k]) + margin;
        u[k] = b[k] - a[k];
    }
    BoxWalk w;
    if (!pm_box_begin(g, lo, hi, w)) return INFINITY;
    const float uu = (u[0] * u[0] + u[1] * u[1]) + u[2] * u[2];
    const float iuu = uu > 0.0f ? 1.0f / uu : 0.0f;
    float U = 0.0f;
    const int nr = w.ny * w.nz;
    int jy = 0, jz = 0;
    _Pragma("unroll 1") for (int r = 0; __ballot(r < nr && U < lim) != 0; r++) {
        uint32_t rb, re;
        pm_box_row(g, w, jy, jz, rb, re);
        if (++jy == w.ny) { jy = 0; jz++; }
        if (!(U < lim)) continue;
        cand += re - rb;
        for (uint32_t i = rb; i < re; i++) {
            const float4 p = g.fsorted[i];
            const float ax = p.x - a[0], ay = p.y - a[1], az = p.z - a[2];
            const float sp = fminf(fmaxf(((ax * u[0] + ay * u[1]) + az * u[2]) * iuu, 0.0f), 1.0f);
            const float dx = ax - sp * u[0], dy = ay - sp * u[1], dz = az - sp * u[2];
            const float dist = sqrtf((dx * dx + dy * dy) + dz * dz) - margin;
            if (dist < g.h) {
                const float dl = fmaxf(dist, 0.0f);
                U += pm_W_in(dl * dl, g.inv_h, g.sig);
            }
        }
    }
    return U;
}

// A step proposal for every lane in PH_PROPOSE (want), all at once: features at
// the lane's current sample k (its density rho is known; one fine-stencil pass
// for grad rho), one MLP evaluation for the whole wave (BF16 MFMA), a jump of
// n samples, and the certificate of the samples k+1 .. k+n-1 it passes over:
// the segment bound U < iso (1 - 2^-9), which covers the rounding of the bound
// and of the samples' own FP32 densities.  A certified lane next evaluates
// sample k+n exactly; the others step to k+1 and ask for no proposal until the
// exact skip next moves them.  Every sample the learned march does not
// evaluate lies below iso, so its first crossing -- and its image -- are the
// plain march's.  Called by all 64 lanes together, with the ray's state passed
// in and out by value.  (Kept inline: as a call, the kernel reserved 0.8 KB of
// stack per lane for the registers live across it and ran 30 % slower even
// without proposals.)
struct Proposal {
    int k, lrn;
    uint32_t mlp, cert, cskip, cand;
};

__device__ __forceinline__ Proposal wave_propose(const MarchLaunch &L, bool want, int k, float tn, float tf, float d0,
                                              float d1, float d2, float rho, uint16_t *X, uint16_t *H, float *Y)
{
    const int l = threadIdx.x & 63;
    const pm_settings &s = L.s;
    const GridDev &g = L.g;
    Proposal o{ k, 1, 0u, 0u, 0u, 0u };
    float f[PM_MLP_IN];
#pragma unroll
    for (int j = 0; j < PM_MLP_IN; j++) f[j] = 0.0f;
    const float t = tn + (float)k * s.step;
    if (want) {
        const float q0 = L.cam[0] + t * d0, q1 = L.cam[1] + t * d1, q2 = L.cam[2] + t * d2;
        float gx, gy, gz;
        uint32_t cand = 0;
        grad_feat(g, q0, q1, q2, gx, gy, gz, cand);
        o.cand += cand;
        const float d[3] = { d0, d1, d2 };
        mlp_features(g, L.cg, L.cam, d, t, tn, tf, rho, sqrtf((gx * gx + gy * gy) + gz * gz),
                     (float)pm_cand27(g, q0, q1, q2), s.iso, f);
    }
    wave_lds_sync();
#pragma unroll
    for (int j = 0; j < PM_MLP_K1; j++) X[l * PM_MLP_K1 + j] = j < PM_MLP_IN ? f2bf(f[j]) : (uint16_t)0;
    wave_lds_sync();
    const float y = mlp_wave_forward_halves(L.net, X, H, Y);
    wave_lds_sync();
    int n = 0;
    if (want) {
        o.mlp = 1;
        const float dist = g.h * (exp2f(fminf(y, 8.0f)) - 1.0f);
        const float sk = fminf(floorf(L.mlp_alpha * dist * L.inv_step), (float)L.mlp_max_skip);
        int kmax = (int)floorf((tf - tn) * L.inv_step);           // last sample inside the AABB
        while (kmax > 0 && tn + (float)kmax * s.step > tf) kmax--;
        if (sk >= (float)L.mlp_min_skip && k + (int)sk <= kmax) n = (int)sk;
        if (n < 2) {
            o.lrn = 0;
            o.k = k + 1;
        }
    }
    const bool cert = want && n >= 2;
    if (__ballot(cert) == 0) return o;
    const float lim = s.iso * (1.0f - 0x1p-9f);
    if (cert) {
        const float ta = tn + (float)(k + 1) * s.step, tb = tn + (float)(k + n - 1) * s.step;
        const float a[3] = { L.cam[0] + ta * d0, L.cam[1] + ta * d1, L.cam[2] + ta * d2 };
        const float b[3] = { L.cam[0] + tb * d0, L.cam[1] + tb * d1, L.cam[2] + tb * d2 };
        // the lattice samples between a and b lie within a few ulps of the
        // segment; the margin covers them and the rounding of the distances
        float m = g.h;
#pragma unroll
        for (int c = 0; c < 3; c++) m = fmaxf(m, fmaxf(fabsf(a[c]), fabsf(b[c])));
        uint32_t cand = 0;
        const float U = pm_segment_bound(g, a, b, m * 0x1p-18f, lim, cand);
        o.cand += cand;
        o.cert = 1;
        if (U < lim) {
            o.k = k + n;
            o.cskip = (uint32_t)(n - 1);
        } else {
            o.k = k + 1;
            o.lrn = 0;
        }
    }
    return o;
}

// The proposals of a wave and their transitions (all 64 lanes).
__device__ __forceinline__ void wave_propose_step(const MarchLaunch &L, RayState &r, Counts &cn, float *region)
{
    const bool want = r.phase == PH_PROPOSE;
    const Proposal o = wave_propose(L, want, r.k, r.tn, r.tf, r.d0, r.d1, r.d2, r.rho, (uint16_t *)region,
                                    (uint16_t *)region + 64 * PM_MLP_K1, region + 2048);
    if (want) {
        r.k = o.k;
        r.lrn = o.lrn;
        r.phase = PH_MARCH;
        cn.mlp += o.mlp;
        cn.cert += o.cert;
        cn.cskip += o.cskip;
        cn.cand += o.cand;
    }
}

// ---- finish ----
// Per hit record: the FP64 bisection step when the settings ask for it, the
// normal (the six central differences in one fused walk, or the sum of gradW)
// and the pixel's outputs.  Every lane of a wave does the same kind of work
// and none of the march's state is held, so this runs packed and short; the
// march kernels keep only the phases that decide where the hit is.
template <bool INSTR, bool LDS_NORMALS>
__device__ __forceinline__ void finish_hit(const MarchLaunch &L, uint32_t px, uint32_t py, uint4 rec, float *buf,
                                           Counts &cn)
{
    const pm_settings &s = L.s;
    const GridDev &g = L.g;
    const uint32_t idx = py * L.W + px;
    const float lo = __uint_as_float(rec.y), hi = __uint_as_float(rec.z);
    float d[3];
    pm_ray_dir(L, px, py, d);
    const float o0 = L.cam[0], o1 = L.cam[1], o2 = L.cam[2];
    float xs[3], th;
    if (rec.x & 2u) {
        // the last bisection step in double (RayMarcher.cpp:316-326 in FP64)
        const double tmd = ((double)lo + (double)hi) * 0.5;
        const double x = (double)o0 + tmd * (double)d[0];
        const double y = (double)o1 + tmd * (double)d[1];
        const double z = (double)o2 + tmd * (double)d[2];
        const double rho = pm_density64(g, x, y, z);
        cn.samples++;
        if (INSTR) {
            cn.cst += pm_stencil_cand(g, (float)x, (float)y, (float)z);
            cn.c27 += pm_cand27(g, (float)x, (float)y, (float)z);
        }
        double dlo = lo, dhi = hi;
        if (rho >= (double)s.iso) dhi = tmd; else dlo = tmd;
        const double tt = (dlo + dhi) * 0.5;
        xs[0] = (float)((double)o0 + tt * (double)d[0]);
        xs[1] = (float)((double)o1 + tt * (double)d[1]);
        xs[2] = (float)((double)o2 + tt * (double)d[2]);
        th = (float)tt;
    } else {
        th = lo;
        xs[0] = o0 + th * d[0]; xs[1] = o1 + th * d[1]; xs[2] = o2 + th * d[2];
    }
    float nx, ny, nz;
    if (s.normal_mode == PM_NORMAL_GRADW) {
        float sx, sy, sz;
        uint32_t cand = 0;
        pm_sum_gradw(g, xs[0], xs[1], xs[2], sx, sy, sz, cand);
        pm_normalize(sx, sy, sz, nx, ny, nz);
    } else {
        const bool comp = s.compensated != 0;
        const float e = s.normal_eps;
        float gv[3];
        const Cdiff6 o = LDS_NORMALS ? pm_cdiff6_lds<kLdsCap>(g, xs[0], xs[1], xs[2], e, buf, comp)
                                     : pm_cdiff6(g, xs[0], xs[1], xs[2], e, comp);
        if (o.cand != 0xffffffffu) {
            cn.samples += 6;
            cn.cand += o.cand;
            gv[0] = o.v[0] - o.v[1]; gv[1] = o.v[2] - o.v[3]; gv[2] = o.v[4] - o.v[5];
        } else {
            // the fused walk's cells leave the grid: +x, -x, +y, -y, +z, -z one at a time
#pragma unroll
            for (int k = 0; k < 6; k++) {
                float q[3] = { xs[0], xs[1], xs[2] };
                q[k >> 1] = (k & 1) ? xs[k >> 1] - e : xs[k >> 1] + e;
                uint32_t cand = 0;
                const float rho = comp ? pm_density32_lds<kLdsCap, 4>(g, q[0], q[1], q[2], buf, cand)
                                       : pm_density32(g, q[0], q[1], q[2], false, cand);
                cn.samples++;
                cn.cand += cand;
                if (k & 1) gv[k >> 1] -= rho; else gv[k >> 1] = rho;
            }
        }
        if (INSTR) {
#pragma unroll 1
            for (int k = 0; k < 6; k++) {
                float q[3] = { xs[0], xs[1], xs[2] };
                q[k >> 1] = (k & 1) ? xs[k >> 1] - e : xs[k >> 1] + e;
                cn.cst += pm_stencil_cand(g, q[0], q[1], q[2]);
                cn.c27 += pm_cand27(g, q[0], q[1], q[2]);
            }
        }
        // RayMarcher.cpp:331-338 with central differences
        const float dd = (gv[0] * gv[0] + gv[1] * gv[1]) + gv[2] * gv[2];
        nx = ny = nz = 0.0f;
        if (dd > 0.0f) {
            const float inv = 1.0f / sqrtf(dd);
            nx = -gv[0] * inv; ny = -gv[1] * inv; nz = -gv[2] * inv;
        }
    }
    L.pos[idx] = make_float4(xs[0], xs[1], xs[2], 1.0f);
    L.nrm[idx] = make_float4(nx, ny, nz, 1.0f);
    if (L.t) L.t[idx] = th;
    cn.hits++;
}

// One 8x8 tile per wave, one pixel per lane: the lanes whose pixel holds a
// hit record finish it.  LDS_NORMALS: the six-sample normal walk compacts its
// candidates through LDS first -- the finish kernel, where every lane of every
// wave finishes hits; the march kernel's own finishing adds them directly
// (measured: 1080p 1.11 -> 1.08 ms, 4K pool 10.4 -> 9.4 ms in the march
// kernel; 54 against 62 ms per batch replay in the finish kernel).
template <bool INSTR, bool LDS_NORMALS>
__device__ __forceinline__ void finish_tile(const MarchLaunch &L, uint32_t tile, float *buf, Counts &cn)
{
    const uint32_t lane = threadIdx.x & 63, tiles_x = (L.W + 7) / 8;
    const uint32_t px = (tile % tiles_x) * 8 + (lane & 7u), py = (tile / tiles_x) * 8 + (lane >> 3);
    if (px >= L.W || py >= L.H) return;
    const uint4 rec = L.hit_rec[py * L.W + px];
    if (rec.x & 1u) finish_hit<INSTR, LDS_NORMALS>(L, px, py, rec, buf, cn);
}

// Schedule 0: persistent waves.  The image is cut into 8x8 tiles and the tiles
// into 8 contiguous bands, one per XCD (workgroups are dispatched round-robin
// over the XCDs, so blockIdx % 8 names the XCD); a wave takes tiles from its
// band's counter and, once the band is drained, from the other bands.  A lane
// whose ray is finished takes the next ray of the wave's current tile before
// the next density evaluation, so all 64 lanes stay busy until the image runs
// out.  Rays that miss the AABB are retired during refill.  Every wave exits
// once no tile is left and its lanes are idle.
// LEARNED adds the step proposals; its LDS region per wave holds either the
// density compaction buffer or the network's activations (never both at once).
// INSTR adds the per-sample counts of instrumented launches.  The walk waits
// on the cache far more than it computes, so the kernel keeps 4 waves per SIMD
// (<= 128 VGPRs; all 1,024 workgroups resident at once).
template <bool LEARNED, bool INSTR>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4, 8))) void k_march_persist(
    MarchLaunch L, uint32_t *__restrict__ band_ctr)
{
    constexpr int CAP = kLdsCap;
    // floats per wave: the compaction buffer, or X[64][32] + H[32][64] bf16 + Y[64]
    constexpr int kRegion = LEARNED ? 2112 : CAP * 64;
    __shared__ float lds[4 * kRegion];
    __shared__ uint32_t own[4][kOwnTiles];       // tiles this wave marched (finished by it at the end)
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    float *region = lds + wave * kRegion;
    float *buf = region + lane;
    const uint32_t tiles_x = (L.W + 7) / 8, tiles_y = (L.H + 7) / 8, T = tiles_x * tiles_y;
    uint32_t nown = 0;
    const uint32_t xcd = blockIdx.x & 7u;
    const bool use_lds = L.s.compensated != 0;

    Counts cn;
    RayState r{};
    r.phase = PH_DONE;
    uint32_t tile = 0, steal = 0;
    int qpos = 64;
    bool exhausted = false;
    const unsigned long long lt_mask = (1ull << lane) - 1ull;
    // lane groups at the end of the launch (the one-lane march otherwise)
    int G = 1;
    const bool group_ok = L.s.skip_mode == PM_SKIP_BOUND || L.s.skip_mode == PM_SKIP_NONE;

    while (true) {
        // refill idle lanes from the wave's tile queue
        while (!exhausted) {
            unsigned long long needy = __ballot(r.phase == PH_DONE);
            if (needy == 0) break;
            if (qpos >= 64) {
                uint32_t t = 0xffffffffu;
                if (lane == 0) {
                    // the bands split the active tiles (the pre-pass lists them)
                    const uint32_t TA = L.tile_buf[2 * (T + 1) + T];
                    while (steal < 8) {
                        uint32_t b = (xcd + steal) & 7u;
                        uint32_t lo = band_lo(TA, b), hi = band_lo(TA, b + 1);
                        uint32_t v = atomicAdd(&band_ctr[b], 1u);
                        if (v < hi - lo) { t = L.tile_buf[2 * (T + 1) + lo + v]; break; }
                        steal++;
                    }
                }
                t = __shfl(t, 0, 64);
                steal = __shfl(steal, 0, 64);
                if (t == 0xffffffffu) { exhausted = true; break; }
                if (nown < kOwnTiles) {
                    if (lane == 0) own[wave][nown] = t;
                    nown++;
                }
                tile = t;
                qpos = 0;
            }
            int rank = __popcll(needy & lt_mask);
            int take = __popcll(needy);
            if (r.phase == PH_DONE && qpos + rank < 64) {
                uint32_t j = (uint32_t)(qpos + rank);
                uint32_t px = (tile % tiles_x) * 8 + (j & 7u), py = (tile / tiles_x) * 8 + (j >> 3);
                if (px < L.W && py < L.H) ray_setup(L, r, px, py, cn);
            }
            qpos = qpos + take < 64 ? qpos + take : 64;
        }
        if (__ballot(r.phase != PH_DONE) == 0) break;
        // The network is batched across the wave (the wave pays for it
        // whenever any lane needs it): it runs once half of the busy lanes wait
        // for a proposal, or at the end (before the lane groups, which take no
        // proposals).  Waiting lanes sit out.
        if (LEARNED) {
            const int nbusy = __popcll(__ballot(r.phase != PH_DONE));
            const int nwant = __popcll(__ballot(r.phase == PH_PROPOSE));
            if (nwant > 0 && (2 * nwant >= nbusy || exhausted))
                wave_propose_step(L, r, cn, region);
        }
        if (exhausted && group_ok) regroup(r, G, region);
        if (G == 1) {
            if (r.phase == PH_MARCH || r.phase == PH_BISECT) {
                if (use_lds) ray_step<INSTR, 16>(L, r, DensLds<CAP, 4>{ &L.g, buf }, cn);
                else ray_step<INSTR, 16>(L, r, DensPlain{ &L.g, false }, cn);
            }
        } else {
            const int j = (int)(lane & (uint32_t)(G - 1));
            float qx = 0.f, qy = 0.f, qz = 0.f, rho = 0.f;
            bool nonskip = false, want = false;
            uint32_t cand = 0;
            if (r.phase != PH_DONE) want = group_prepare(L, r, G, j, qx, qy, qz, nonskip);
            if (want) rho = use_lds ? pm_density32_lds<CAP, 4>(L.g, qx, qy, qz, buf, cand)
                                    : pm_density32(L.g, qx, qy, qz, false, cand);
            group_finish<INSTR>(L, r, G, j, want, nonskip, rho, cand, qx, qy, qz, cn);
        }
    }
    // The tiles this wave marched have all their rays done: their FP64 steps,
    // normals and outputs are computed here, while other waves still march
    // (k_finish then skips them).
    if (nown > 0) {
        wave_lds_sync();
        for (uint32_t i = 0; i < nown; i++) {
            const uint32_t t = own[wave][i];
            finish_tile<INSTR, false>(L, t, buf, cn);
            if (lane == 0) L.tile_buf[t] = 2u;
        }
    }
    flush_counts(L.counters, cn);
}

__global__ void k_zero_u32(uint32_t *p, int n)
{
    for (int i = (int)threadIdx.x; i < n; i += (int)blockDim.x) p[i] = 0u;
}

// PerPixel_Isotropic (RayMarcher.cpp:256-344): start on the depth-buffer
// surface, fixed step vector, MaxSteps samples, first sample with rho >= iso,
// normal = normalize(sum gradW) (or central differences on request).
struct RefCounts {
    uint32_t rays = 0, in = 0, hits = 0, samples = 0, ms = 0;
    unsigned long long cand = 0;
};

__device__ __forceinline__ void ref_pixel(const MarchLaunch &L, uint32_t px, uint32_t py, bool active, RefCounts &rc)
{
    const GridDev &g = L.g;
    const pm_settings &s = L.s;
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
                if (s.skip_mode == PM_SKIP_DG_BOX || s.skip_mode == PM_SKIP_DG_CONS) {   // RayMarcher.cpp:284-304
                    float nmin[3], nmax[3];
                    while (pm_dg_query(g, p0, p1, p2, nmin, nmax)) {
                        float pin[3] = { p0, p1, p2 }, stv[3] = { st0, st1, st2 }, q[3];
                        pm_intersect_aabb(pin, stv, nmin, nmax, q);
                        p0 = q[0] + st0; p1 = q[1] + st1; p2 = q[2] + st2;
                    }
                }
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
    rc.rays += c_rays; rc.in += c_in; rc.hits += c_hits; rc.samples += c_samples; rc.ms += c_ms;
    rc.cand += c_cand;
}

__global__ __launch_bounds__(256) void k_march_reference(MarchLaunch L)
{
    uint32_t px, py;
    if (!pm_tile(L.W, L.H, px, py)) return;
    RefCounts rc;
    ref_pixel(L, px, py, px < L.W && py < L.H, rc);
    flush_counters(L.counters, rc.rays, rc.in, rc.hits, rc.samples, rc.cand, rc.ms);
}

// The reference march over the active tiles of the tile pre-pass only (tiles
// with a pixel whose depth is not 1; the pre-pass wrote the others' misses),
// one tile per wave, the active list spread over the 8 XCDs as
// k_march_aniso_active spreads it.
__global__ __launch_bounds__(256) void k_march_reference_active(MarchLaunch L)
{
    const uint32_t T = ((L.W + 7) / 8) * ((L.H + 7) / 8);
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint32_t *order = L.tile_buf + 2 * (T + 1);
    const uint32_t count = order[T], chunk = (count + 7) / 8;
    const uint32_t local = (blockIdx.x >> 3) * 4 + wave, i = (blockIdx.x & 7u) * chunk + local;
    if (local >= chunk || i >= count) return;   // whole wave
    const uint32_t tile = order[i], tiles_x = (L.W + 7) / 8;
    const uint32_t px = (tile % tiles_x) * 8 + (lane & 7u), py = (tile / tiles_x) * 8 + (lane >> 3);
    RefCounts rc;
    ref_pixel(L, px, py, px < L.W && py < L.H, rc);
    flush_counters(L.counters, rc.rays, rc.in, rc.hits, rc.samples, rc.cand, rc.ms);
}

// Batched schedule: persistent waves take chunks of tiles (each chunk belongs to
// one (view, frame) job) from one counter; inside a chunk lanes are refilled
// ray by ray as in k_march_persist; a wave moves to its next chunk -- and
// loads that job's grid and camera -- once all its lanes are idle.
template <bool LEARNED>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4, 8))) void k_march_batch(BatchLaunch B)
{
    constexpr int kRegion = LEARNED ? 2112 : kLdsCap * 64;
    __shared__ float lds[4 * kRegion];
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    float *region = lds + wave * kRegion;
    float *buf = region + lane;
    const bool use_lds = B.base.s.compensated != 0;

    const MarchLaunch *Lp = B.jobs;
    Counts cn;
    RayState r{};
    r.phase = PH_DONE;
    uint32_t tile = 0, tile_i = 0, ntiles = 0, tile0 = 0;
    int qpos = 64;
    bool exhausted = false;
    const unsigned long long lt_mask = (1ull << lane) - 1ull;
    // lane groups while a chunk drains (its last rays, before the next job)
    int G = 1;
    const bool group_ok = B.base.s.skip_mode == PM_SKIP_BOUND || B.base.s.skip_mode == PM_SKIP_NONE;

    while (true) {
        while (!exhausted) {
            unsigned long long needy = __ballot(r.phase == PH_DONE);
            if (needy == 0) break;
            if (qpos >= 64) {
                if (tile_i < ntiles) {
                    tile = tile0 + tile_i++;
                    qpos = 0;
                } else {
                    if (__ballot(r.phase != PH_DONE) != 0) break;    // drain before changing job
                    uint32_t cidx = 0;
                    if (lane == 0) cidx = atomicAdd(B.chunk_ctr, 1u);
                    cidx = __builtin_amdgcn_readfirstlane(__shfl(cidx, 0, 64));
                    if (cidx >= B.nchunks) { exhausted = true; break; }
                    const BatchChunk ch = B.chunks[cidx];
                    Lp = B.jobs + __builtin_amdgcn_readfirstlane(ch.job);
                    tile0 = ch.tile0;
                    ntiles = ch.ntiles;
                    tile_i = 0;
                    G = 1;                                        // every lane its own ray again
                    continue;
                }
            }
            int rank = __popcll(needy & lt_mask);
            int take = __popcll(needy);
            if (r.phase == PH_DONE && qpos + rank < 64) {
                uint32_t j = (uint32_t)(qpos + rank);
                const uint32_t tiles_x = (Lp->W + 7) / 8;
                uint32_t px = (tile % tiles_x) * 8 + (j & 7u), py = (tile / tiles_x) * 8 + (j >> 3);
                if (px < Lp->W && py < Lp->H) ray_setup(*Lp, r, px, py, cn);
            }
            qpos = qpos + take < 64 ? qpos + take : 64;
        }
        if (__ballot(r.phase != PH_DONE) == 0) {
            if (exhausted) break;
            continue;
        }
        // batched proposals, as in k_march_persist (before the lane groups)
        const bool draining = exhausted || (qpos >= 64 && tile_i >= ntiles);
        if (LEARNED) {
            const int nbusy = __popcll(__ballot(r.phase != PH_DONE));
            const int nwant = __popcll(__ballot(r.phase == PH_PROPOSE));
            if (nwant > 0 && (2 * nwant >= nbusy || draining))
                wave_propose_step(*Lp, r, cn, region);
        }
        if (group_ok && draining) regroup(r, G, region);
        if (G == 1) {
            if (r.phase == PH_MARCH || r.phase == PH_BISECT) {
                if (use_lds) ray_step<false, 8>(*Lp, r, DensLds<kLdsCap, 4>{ &Lp->g, buf }, cn);
                else ray_step<false, 8>(*Lp, r, DensPlain{ &Lp->g, false }, cn);
            }
        } else {
            const int j = (int)(lane & (uint32_t)(G - 1));
            float qx = 0.f, qy = 0.f, qz = 0.f, rho = 0.f;
            bool nonskip = false, want = false;
            uint32_t cand = 0;
            if (r.phase != PH_DONE) want = group_prepare(*Lp, r, G, j, qx, qy, qz, nonskip);
            if (want) rho = use_lds ? pm_density32_lds<kLdsCap, 4>(Lp->g, qx, qy, qz, buf, cand)
                                    : pm_density32(Lp->g, qx, qy, qz, false, cand);
            group_finish<false>(*Lp, r, G, j, want, nonskip, rho, cand, qx, qy, qz, cn);
        }
    }
    flush_counts(B.base.counters, cn);
}

// Single launch: the active tiles no march wave finished; batch: the tiles of
// the chunks, one chunk per workgroup at a time.  Four waves per SIMD
// (registers capped at 128) hide more of the walks' latency than three would
// at 150 registers.
template <bool BATCH, bool INSTR>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4, 8))) void k_finish(MarchLaunch L0, const MarchLaunch *__restrict__ jobs,
                                                const BatchChunk *__restrict__ chunks, uint32_t nchunks)
{
    __shared__ float lds[4 * kLdsCap * 64];
    const uint32_t wave = threadIdx.x >> 6;
    float *buf = lds + wave * kLdsCap * 64 + (threadIdx.x & 63);
    Counts cn;
    if (BATCH) {
        for (uint32_t c = blockIdx.x; c < nchunks; c += gridDim.x) {
            const BatchChunk ch = chunks[c];
            const MarchLaunch *Lp = jobs + __builtin_amdgcn_readfirstlane(ch.job);
            for (uint32_t t = wave; t < ch.ntiles; t += 4) finish_tile<INSTR, true>(*Lp, ch.tile0 + t, buf, cn);
        }
    } else {
        const uint32_t T = ((L0.W + 7) / 8) * ((L0.H + 7) / 8);
        const uint32_t TA = L0.tile_buf[2 * (T + 1) + T];
        for (uint32_t i = blockIdx.x * 4 + wave; i < TA; i += gridDim.x * 4) {
            const uint32_t tile = L0.tile_buf[2 * (T + 1) + i];
            if (L0.tile_buf[tile] == 2u) continue;     // finished by the march
            finish_tile<INSTR, true>(L0, tile, buf, cn);
        }
    }
    flush_counts(L0.counters, cn);
}

static hipError_t launch_finish(const MarchLaunch &L0, hipStream_t st)
{
    const uint32_t T = ((L0.W + 7) / 8) * ((L0.H + 7) / 8);
    const uint32_t b = (T + 3) / 4;
    const dim3 grid(b < 2048 ? b : 2048), blk(256);
    if (L0.s.instrument) hipLaunchKernelGGL((k_finish<false, true>), grid, blk, 0, st, L0, nullptr, nullptr, 0u);
    else hipLaunchKernelGGL((k_finish<false, false>), grid, blk, 0, st, L0, nullptr, nullptr, 0u);
    return hipGetLastError();
}

hipError_t launch_march_batch(const BatchLaunch &B, uint32_t blocks, hipStream_t st)
{
    if (B.base.s.learned) hipLaunchKernelGGL((k_march_batch<true>), dim3(blocks), dim3(256), 0, st, B);
    else hipLaunchKernelGGL((k_march_batch<false>), dim3(blocks), dim3(256), 0, st, B);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    const uint32_t fb = B.nchunks < 8192 ? B.nchunks : 8192;
    hipLaunchKernelGGL((k_finish<true, false>), dim3(fb), dim3(256), 0, st, B.base, B.jobs, B.chunks, B.nchunks);
    return hipGetLastError();
}

__global__ void k_zero_u64(unsigned long long *p, uint32_t n)
{
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) p[i] = 0ull;
}

hipError_t launch_zero_u32(uint32_t *p, uint32_t n, hipStream_t st)
{
    hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, st, p, (int)n);
    return hipGetLastError();
}

hipError_t launch_zero_u64(unsigned long long *p, uint32_t n, hipStream_t st)
{
    hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, st, p, n);
    return hipGetLastError();
}

__global__ void k_init_counters(unsigned long long *p, uint32_t n, unsigned long long rays0)
{
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) p[i] = i == (uint32_t)C_RAYS ? rays0 : 0ull;   // slot 0 holds rays0
}

hipError_t launch_init_counters(unsigned long long *p, uint32_t n, unsigned long long rays0, hipStream_t st)
{
    hipLaunchKernelGGL(k_init_counters, dim3(1), dim3(64), 0, st, p, n, rays0);
    return hipGetLastError();
}

hipError_t launch_tile_flags(const MarchLaunch &L, uint32_t T, hipStream_t st)
{
    hipLaunchKernelGGL(k_tile_flags, dim3((T + 1 + 3) / 4), dim3(256), 0, st, L, T);
    return hipGetLastError();
}

hipError_t launch_miss_tiles(const MarchLaunch &L, uint32_t T, hipStream_t st)
{
    hipLaunchKernelGGL(k_miss_tiles, dim3((T + 3) / 4), dim3(256), 0, st, L, T);
    return hipGetLastError();
}

// The configured isotropic march of one view: schedule 1, or the tile
// pre-pass, the persistent march and the finish of the tiles its waves left
// (L.tile_buf and L.hit_rec set).
hipError_t launch_tile_prepass(const MarchLaunch &L, hipStream_t st)
{
    const uint32_t tiles = ((L.W + 7) / 8) * ((L.H + 7) / 8);
    uint32_t *flag = L.tile_buf, *pos = L.tile_buf + (tiles + 1);
    hipLaunchKernelGGL(k_tile_flags, dim3((tiles + 1 + 3) / 4), dim3(256), 0, st, L, tiles);
    hipError_t e = launch_scan_u32(flag, pos, tiles + 1, L.tile_buf + 3 * (tiles + 1), st);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_tile_compact, dim3((tiles + 1 + 255) / 256), dim3(256), 0, st, L.tile_buf, tiles);
    hipLaunchKernelGGL(k_miss_tiles, dim3((tiles + 3) / 4), dim3(256), 0, st, L, tiles);
    return hipGetLastError();
}

static hipError_t launch_march_config(const MarchLaunch &L, hipStream_t st)
{
    const uint32_t tiles = ((L.W + 7) / 8) * ((L.H + 7) / 8);
    if (L.s.schedule == 1 && !L.s.learned) {
        hipLaunchKernelGGL(k_march_simple, dim3((tiles + 3) / 4), dim3(256), 0, st, L);
        return hipGetLastError();
    }
    hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, st, L.band_ctr, 8);
    hipError_t e = launch_tile_prepass(L, st);
    if (e != hipSuccess) return e;
    // enough resident waves to fill the chip; extra waves find no tile and exit
    uint32_t blocks = L.persist_blocks ? L.persist_blocks : 1024;
    const uint32_t need = (tiles + 3) / 4;
    if (blocks > need) blocks = (need + 7) & ~7u;
    const dim3 grid(blocks), blk(256);
    const bool instr = L.s.instrument != 0;
    if (L.s.learned) {
        if (instr) hipLaunchKernelGGL((k_march_persist<true, true>), grid, blk, 0, st, L, L.band_ctr);
        else hipLaunchKernelGGL((k_march_persist<true, false>), grid, blk, 0, st, L, L.band_ctr);
    } else {
        if (instr) hipLaunchKernelGGL((k_march_persist<false, true>), grid, blk, 0, st, L, L.band_ctr);
        else hipLaunchKernelGGL((k_march_persist<false, false>), grid, blk, 0, st, L, L.band_ctr);
    }
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    return launch_finish(L, st);
}

hipError_t launch_march(const MarchLaunch &L, hipStream_t st)
{
    if (L.s.anisotropic) return launch_march_aniso(L, st);
    if (L.s.mode == PM_MODE_REFERENCE) {
        const uint32_t tiles = ((L.W + 7) / 8) * ((L.H + 7) / 8);
        if (L.tile_buf && L.s.schedule == 0) {
            hipError_t e = launch_tile_prepass(L, st);
            if (e != hipSuccess) return e;
            const uint32_t per_xcd = ((tiles + 7) / 8 + 3) / 4;
            hipLaunchKernelGGL(k_march_reference_active, dim3(8 * per_xcd), dim3(256), 0, st, L);
            return hipGetLastError();
        }
        hipLaunchKernelGGL(k_march_reference, dim3((tiles + 3) / 4), dim3(256), 0, st, L);
        return hipGetLastError();
    }
    return launch_march_config(L, st);
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
