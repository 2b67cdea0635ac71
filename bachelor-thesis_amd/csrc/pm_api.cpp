// pm_api.cpp -- the C ABI of libpmarch.so (include/pmarch.h).
//
// A context owns one HIP stream (default priority), the frames' device grids,
// the output images and a pinned staging area.  pm_march_async enqueues
// kernel + device-to-host copies and returns; pm_query / pm_wait complete the
// hand-off into the caller's buffers (the caller's buffers are never written
// by the device directly, so they may be ordinary pageable memory such as the
// mapped staging memory the reference hands to Prepare).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "pm_internal.h"
#include "pmarch.h"

using namespace pm;

namespace {

struct DevBuf {
    void *p = nullptr;
    size_t cap = 0;
    hipError_t ensure(size_t bytes)
    {
        if (bytes <= cap && p) return hipSuccess;
        if (p) { hipFree(p); p = nullptr; cap = 0; }
        if (bytes == 0) bytes = 256;
        hipError_t e = hipMalloc(&p, bytes);
        if (e == hipSuccess) cap = bytes;
        return e;
    }
    void release() { if (p) hipFree(p); p = nullptr; cap = 0; }
};

struct Frame {
    pm_frame_info info{};
    GridDev g{};
    DevBuf sorted, cell_start, cell_of, counts, cursor, scan_tmp, xyz, bounds_partial, bounds;
    void release()
    {
        sorted.release(); cell_start.release(); cell_of.release(); counts.release(); cursor.release();
        scan_tmp.release(); xyz.release(); bounds_partial.release(); bounds.release();
    }
};

}  // namespace

struct pm_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev_launch = nullptr, ev_k0 = nullptr, ev_k1 = nullptr, ev_done = nullptr;
    std::map<int32_t, Frame *> frames;
    std::string err;
    // outputs of the last march
    DevBuf d_pos, d_nrm, d_t, d_depth, d_counters, d_q, d_out;
    uint64_t npx = 0;
    bool have_t = false;
    // pinned staging for asynchronous hand-off
    void *h_stage = nullptr;
    size_t h_stage_cap = 0;
    unsigned long long *h_counters = nullptr;
    float *user_pos = nullptr, *user_nrm = nullptr, *user_t = nullptr;
    bool pending = false;
    bool finished = true;
    pm_stats stats{};
};

static int fail(pm_ctx *c, int code, const char *fmt, ...)
{
    if (c) {
        char buf[512];
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(buf, sizeof(buf), fmt, ap);
        va_end(ap);
        c->err = buf;
    }
    return code;
}

#define HIPCHK(c, expr)                                                                        \
    do {                                                                                       \
        hipError_t _e = (expr);                                                                \
        if (_e != hipSuccess) return fail((c), PM_ERR_HIP, "%s: %s", #expr, hipGetErrorString(_e)); \
    } while (0)

extern "C" {

int pm_create(int device, pm_ctx **out)
{
    if (!out) return PM_ERR_ARG;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return PM_ERR_NODEV;
    if (device < 0 || device >= n) return PM_ERR_NODEV;
    pm_ctx *c = new pm_ctx();
    c->device = device;
    if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&c->ev_launch) != hipSuccess || hipEventCreate(&c->ev_k0) != hipSuccess ||
        hipEventCreate(&c->ev_k1) != hipSuccess || hipEventCreate(&c->ev_done) != hipSuccess ||
        c->d_counters.ensure(sizeof(unsigned long long) * C_COUNT) != hipSuccess ||
        hipHostMalloc((void **)&c->h_counters, sizeof(unsigned long long) * C_COUNT, hipHostMallocDefault) != hipSuccess) {
        pm_destroy(c);
        return PM_ERR_HIP;
    }
    *out = c;
    return PM_OK;
}

void pm_destroy(pm_ctx *c)
{
    if (!c) return;
    hipSetDevice(c->device);
    if (c->stream) hipStreamSynchronize(c->stream);
    for (auto &kv : c->frames) { kv.second->release(); delete kv.second; }
    c->d_pos.release(); c->d_nrm.release(); c->d_t.release(); c->d_depth.release();
    c->d_counters.release(); c->d_q.release(); c->d_out.release();
    if (c->h_stage) hipHostFree(c->h_stage);
    if (c->h_counters) hipHostFree(c->h_counters);
    if (c->ev_launch) hipEventDestroy(c->ev_launch);
    if (c->ev_k0) hipEventDestroy(c->ev_k0);
    if (c->ev_k1) hipEventDestroy(c->ev_k1);
    if (c->ev_done) hipEventDestroy(c->ev_done);
    if (c->stream) hipStreamDestroy(c->stream);
    delete c;
}

const char *pm_last_error(const pm_ctx *c) { return c ? c->err.c_str() : "null context"; }

static int build_frame(pm_ctx *c, int32_t frame, const float *d_src, uint64_t n, float h)
{
    Frame *f;
    auto it = c->frames.find(frame);
    if (it == c->frames.end()) { f = new Frame(); c->frames[frame] = f; }
    else f = it->second;

    hipEvent_t e0, e1;
    HIPCHK(c, hipEventCreate(&e0));
    HIPCHK(c, hipEventCreate(&e1));
    HIPCHK(c, hipEventRecord(e0, c->stream));
    uint32_t np = bounds_partials(n);
    HIPCHK(c, f->bounds_partial.ensure(sizeof(float) * 6 * np));
    HIPCHK(c, f->bounds.ensure(sizeof(float) * 6));
    HIPCHK(c, launch_bounds_reduce(d_src, n, (float *)f->bounds_partial.p, np, (float *)f->bounds.p, c->stream));
    float b[6];
    HIPCHK(c, hipMemcpyAsync(b, f->bounds.p, sizeof(b), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    for (int a = 0; a < 6; a++)
        if (!std::isfinite(b[a])) return fail(c, PM_ERR_ARG, "pm_set_frame: non-finite particle coordinate");

    pm_frame_info &I = f->info;
    memset(&I, 0, sizeof(I));
    I.n = n;
    I.h = h;
    float inv_h = 1.0f / h;
    uint64_t nc = 1;
    for (int a = 0; a < 3; a++) {
        I.pmin[a] = b[a];
        I.pmax[a] = b[3 + a];
        float pad = 1.0f * h;                           // Frame::ComputeAABB, Dataset.cpp:89-91
        I.bmin[a] = I.pmin[a] - pad;
        I.bmax[a] = I.pmax[a] + pad;
        int32_t kmin = host_key(I.pmin[a], inv_h), kmax = host_key(I.pmax[a], inv_h);
        I.origin[a] = kmin - 2;
        I.dims[a] = kmax - kmin + 5;
        nc *= (uint64_t)I.dims[a];
    }
    if (nc >= (1ull << 31)) return fail(c, PM_ERR_SIZE, "pm_set_frame: grid of %llu cells exceeds 2^31", (unsigned long long)nc);
    I.ncells = nc;

    HIPCHK(c, f->sorted.ensure(sizeof(float4) * n));
    HIPCHK(c, f->cell_start.ensure(sizeof(uint32_t) * (nc + 1)));
    HIPCHK(c, f->cell_of.ensure(sizeof(uint32_t) * n));
    HIPCHK(c, f->counts.ensure(sizeof(uint32_t) * (nc + 1)));
    HIPCHK(c, f->cursor.ensure(sizeof(uint32_t) * nc));
    HIPCHK(c, f->scan_tmp.ensure(sizeof(uint32_t) * scan_tmp_size((uint32_t)(nc + 1))));

    GridBuildArgs a{};
    a.xyz = d_src; a.n = n; a.inv_h = inv_h;
    a.ox = I.origin[0]; a.oy = I.origin[1]; a.oz = I.origin[2];
    a.nx = I.dims[0]; a.ny = I.dims[1]; a.nz = I.dims[2];
    a.ncells = (uint32_t)nc;
    a.cell_of = (uint32_t *)f->cell_of.p;
    a.counts = (uint32_t *)f->counts.p;
    a.cursor = (uint32_t *)f->cursor.p;
    a.scan_tmp = (uint32_t *)f->scan_tmp.p;
    a.sorted = (float4 *)f->sorted.p;
    a.cell_start = (uint32_t *)f->cell_start.p;
    HIPCHK(c, launch_grid_build(a, c->stream));
    HIPCHK(c, hipEventRecord(e1, c->stream));
    HIPCHK(c, hipEventSynchronize(e1));
    float ms = 0.f;
    hipEventElapsedTime(&ms, e0, e1);
    I.build_ms = ms;
    hipEventDestroy(e0);
    hipEventDestroy(e1);

    GridDev &g = f->g;
    g.sorted = (const float4 *)f->sorted.p;
    g.cell_start = (const uint32_t *)f->cell_start.p;
    g.ox = I.origin[0]; g.oy = I.origin[1]; g.oz = I.origin[2];
    g.nx = I.dims[0]; g.ny = I.dims[1]; g.nz = I.dims[2];
    g.h = h;
    g.inv_h = inv_h;
    g.h2 = h * h;
    g.sig = 8.0f / (PM_PI_F * h * h * h);              // Kernel.cpp:13
    g.h_d = (double)h;
    g.h2_d = g.h_d * g.h_d;
    g.hinv_d = 1.0 / g.h_d;
    g.sig_d = 8.0 / (PM_PI_D * g.h_d * g.h_d * g.h_d);
    for (int a2 = 0; a2 < 3; a2++) { g.bmin[a2] = I.bmin[a2]; g.bmax[a2] = I.bmax[a2]; }
    return PM_OK;
}

int pm_set_frame(pm_ctx *c, int32_t frame, const float *xyz, uint64_t n, float h)
{
    if (!c) return PM_ERR_ARG;
    if (!xyz || n == 0 || n >= (1ull << 31) || !(h > 0.0f) || !std::isfinite(h))
        return fail(c, PM_ERR_ARG, "pm_set_frame: need xyz, 0 < n < 2^31 and finite h > 0");
    if (c->pending) return fail(c, PM_ERR_STATE, "pm_set_frame: a march is in flight; call pm_wait first");
    HIPCHK(c, hipSetDevice(c->device));
    auto it = c->frames.find(frame);
    Frame *f = it == c->frames.end() ? (c->frames[frame] = new Frame()) : it->second;
    HIPCHK(c, f->xyz.ensure(sizeof(float) * 3 * n));
    HIPCHK(c, hipMemcpyAsync(f->xyz.p, xyz, sizeof(float) * 3 * n, hipMemcpyHostToDevice, c->stream));
    return build_frame(c, frame, (const float *)f->xyz.p, n, h);
}

int pm_set_frame_device(pm_ctx *c, int32_t frame, const void *d_xyz, uint64_t n, float h)
{
    if (!c) return PM_ERR_ARG;
    if (!d_xyz || n == 0 || n >= (1ull << 31) || !(h > 0.0f) || !std::isfinite(h))
        return fail(c, PM_ERR_ARG, "pm_set_frame_device: need d_xyz, 0 < n < 2^31 and finite h > 0");
    if (c->pending) return fail(c, PM_ERR_STATE, "pm_set_frame_device: a march is in flight");
    HIPCHK(c, hipSetDevice(c->device));
    return build_frame(c, frame, (const float *)d_xyz, n, h);
}

static Frame *get_frame(pm_ctx *c, int32_t frame)
{
    auto it = c->frames.find(frame);
    return it == c->frames.end() ? nullptr : it->second;
}

int pm_frame_info_get(pm_ctx *c, int32_t frame, pm_frame_info *out)
{
    if (!c || !out) return PM_ERR_ARG;
    Frame *f = get_frame(c, frame);
    if (!f) return fail(c, PM_ERR_FRAME, "frame %d not set", frame);
    *out = f->info;
    return PM_OK;
}

int pm_get_grid(pm_ctx *c, int32_t frame, uint32_t *cell_start, float *sorted, uint32_t *index)
{
    if (!c) return PM_ERR_ARG;
    Frame *f = get_frame(c, frame);
    if (!f) return fail(c, PM_ERR_FRAME, "frame %d not set", frame);
    HIPCHK(c, hipSetDevice(c->device));
    uint64_t n = f->info.n;
    if (cell_start)
        HIPCHK(c, hipMemcpyAsync(cell_start, f->cell_start.p, sizeof(uint32_t) * (f->info.ncells + 1), hipMemcpyDeviceToHost, c->stream));
    std::vector<float> tmp;
    if (sorted || index) {
        tmp.resize(4 * n);
        HIPCHK(c, hipMemcpyAsync(tmp.data(), f->sorted.p, sizeof(float) * 4 * n, hipMemcpyDeviceToHost, c->stream));
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    for (uint64_t i = 0; i < (sorted || index ? n : 0); i++) {
        uint32_t bits;
        memcpy(&bits, &tmp[4 * i + 3], 4);
        if (sorted) { sorted[4 * i] = tmp[4 * i]; sorted[4 * i + 1] = tmp[4 * i + 1]; sorted[4 * i + 2] = tmp[4 * i + 2]; sorted[4 * i + 3] = 0.0f; }
        if (index) index[i] = bits;
    }
    return PM_OK;
}

static int validate_settings(pm_ctx *c, const pm_settings *s, const Frame *f)
{
    if (!s) return fail(c, PM_ERR_ARG, "settings is NULL");
    if (s->mode != PM_MODE_CONFIG && s->mode != PM_MODE_REFERENCE) return fail(c, PM_ERR_ARG, "unknown mode %d", s->mode);
    if (!(s->step > 0.0f) || !std::isfinite(s->step)) return fail(c, PM_ERR_ARG, "step must be finite and > 0");
    if (!std::isfinite(s->iso)) return fail(c, PM_ERR_ARG, "iso must be finite");
    if (s->bisect_iters < 0 || s->bisect_iters > 40) return fail(c, PM_ERR_ARG, "bisect_iters out of [0,40]");
    if (s->normal_mode != PM_NORMAL_CDIFF && s->normal_mode != PM_NORMAL_GRADW) return fail(c, PM_ERR_ARG, "unknown normal_mode");
    if (s->skip_mode != PM_SKIP_NONE) return fail(c, PM_ERR_ARG, "skip_mode %d is not available in this build", s->skip_mode);
    if (s->learned) return fail(c, PM_ERR_ARG, "learned step: no network loaded");
    if (s->mode == PM_MODE_REFERENCE && (s->max_steps < 0 || s->max_steps > (1 << 24)))
        return fail(c, PM_ERR_ARG, "max_steps out of range");
    if (s->mode == PM_MODE_CONFIG) {
        double diag = 0.0;
        for (int a = 0; a < 3; a++) {
            double e = (double)f->info.bmax[a] - (double)f->info.bmin[a];
            diag += e * e;
        }
        if (std::sqrt(diag) / (double)s->step > (double)(1 << 24))
            return fail(c, PM_ERR_ARG, "step too small for the frame extent (> 2^24 samples per ray)");
    }
    return PM_OK;
}

int pm_march_async(pm_ctx *c, int32_t frame, const pm_settings *s, const float inv_pv[16], const float cam[3],
                   uint32_t W, uint32_t H, const float *depth_in, float *positions_out, float *normals_out,
                   float *t_out)
{
    if (!c) return PM_ERR_ARG;
    if (c->pending) return fail(c, PM_ERR_STATE, "pm_march_async: previous march still in flight");
    if (!inv_pv || !cam) return fail(c, PM_ERR_ARG, "inv_pv and cam are required");
    if (W == 0 || H == 0 || (uint64_t)W * H > (1ull << 30)) return fail(c, PM_ERR_SIZE, "bad image size %ux%u", W, H);
    Frame *f = get_frame(c, frame);
    if (!f) return fail(c, PM_ERR_FRAME, "frame %d not set", frame);
    int rc = validate_settings(c, s, f);
    if (rc) return rc;
    if (s->mode == PM_MODE_REFERENCE && !depth_in) return fail(c, PM_ERR_ARG, "reference mode needs depth_in");
    HIPCHK(c, hipSetDevice(c->device));

    uint64_t npx = (uint64_t)W * H;
    HIPCHK(c, c->d_pos.ensure(sizeof(float4) * npx));
    HIPCHK(c, c->d_nrm.ensure(sizeof(float4) * npx));
    HIPCHK(c, c->d_t.ensure(sizeof(float) * npx));
    c->npx = npx;

    MarchLaunch L{};
    L.g = f->g;
    L.s = *s;
    memcpy(L.M, inv_pv, sizeof(L.M));
    memcpy(L.cam, cam, sizeof(L.cam));
    L.W = W; L.H = H;
    L.pos = (float4 *)c->d_pos.p;
    L.nrm = (float4 *)c->d_nrm.p;
    L.t = (float *)c->d_t.p;
    L.counters = (unsigned long long *)c->d_counters.p;

    HIPCHK(c, hipEventRecord(c->ev_launch, c->stream));
    if (s->mode == PM_MODE_REFERENCE) {
        HIPCHK(c, c->d_depth.ensure(sizeof(float) * npx));
        HIPCHK(c, hipMemcpyAsync(c->d_depth.p, depth_in, sizeof(float) * npx, hipMemcpyHostToDevice, c->stream));
        L.depth = (const float *)c->d_depth.p;
    }
    HIPCHK(c, hipMemsetAsync(c->d_counters.p, 0, sizeof(unsigned long long) * C_COUNT, c->stream));
    HIPCHK(c, hipEventRecord(c->ev_k0, c->stream));
    HIPCHK(c, launch_march(L, c->stream));
    HIPCHK(c, hipEventRecord(c->ev_k1, c->stream));
    HIPCHK(c, hipMemcpyAsync(c->h_counters, c->d_counters.p, sizeof(unsigned long long) * C_COUNT, hipMemcpyDeviceToHost, c->stream));

    c->user_pos = positions_out;
    c->user_nrm = normals_out;
    c->user_t = t_out;
    if (positions_out || normals_out || t_out) {
        size_t need = (positions_out ? 16 * npx : 0) + (normals_out ? 16 * npx : 0) + (t_out ? 4 * npx : 0);
        if (need > c->h_stage_cap) {
            if (c->h_stage) hipHostFree(c->h_stage);
            c->h_stage = nullptr;
            c->h_stage_cap = 0;
            HIPCHK(c, hipHostMalloc(&c->h_stage, need, hipHostMallocDefault));
            c->h_stage_cap = need;
        }
        char *p = (char *)c->h_stage;
        if (positions_out) { HIPCHK(c, hipMemcpyAsync(p, c->d_pos.p, 16 * npx, hipMemcpyDeviceToHost, c->stream)); p += 16 * npx; }
        if (normals_out) { HIPCHK(c, hipMemcpyAsync(p, c->d_nrm.p, 16 * npx, hipMemcpyDeviceToHost, c->stream)); p += 16 * npx; }
        if (t_out) { HIPCHK(c, hipMemcpyAsync(p, c->d_t.p, 4 * npx, hipMemcpyDeviceToHost, c->stream)); }
    }
    HIPCHK(c, hipEventRecord(c->ev_done, c->stream));
    c->pending = true;
    c->finished = false;
    return PM_OK;
}

static int complete(pm_ctx *c)
{
    uint64_t npx = c->npx;
    const char *p = (const char *)c->h_stage;
    if (c->user_pos) { memcpy(c->user_pos, p, 16 * npx); p += 16 * npx; }
    if (c->user_nrm) { memcpy(c->user_nrm, p, 16 * npx); p += 16 * npx; }
    if (c->user_t) { memcpy(c->user_t, p, 4 * npx); }
    c->user_pos = c->user_nrm = c->user_t = nullptr;
    pm_stats &S = c->stats;
    memset(&S, 0, sizeof(S));
    S.rays = c->h_counters[C_RAYS];
    S.rays_in_aabb = c->h_counters[C_IN_AABB];
    S.hits = c->h_counters[C_HITS];
    S.samples = c->h_counters[C_SAMPLES];
    S.cand = c->h_counters[C_CAND];
    S.march_samples = c->h_counters[C_MARCH_SAMPLES];
    S.mlp_evals = c->h_counters[C_MLP];
    hipEventElapsedTime(&S.kernel_ms, c->ev_k0, c->ev_k1);
    hipEventElapsedTime(&S.total_ms, c->ev_launch, c->ev_done);
    c->pending = false;
    c->finished = true;
    return PM_OK;
}

int pm_query(pm_ctx *c)
{
    if (!c) return PM_ERR_ARG;
    if (!c->pending) return c->finished ? PM_OK : fail(c, PM_ERR_STATE, "no march launched");
    hipError_t e = hipEventQuery(c->ev_done);
    if (e == hipErrorNotReady) return PM_PENDING;
    if (e != hipSuccess) { c->pending = false; return fail(c, PM_ERR_HIP, "march failed: %s", hipGetErrorString(e)); }
    return complete(c);
}

int pm_wait(pm_ctx *c)
{
    if (!c) return PM_ERR_ARG;
    if (!c->pending) return PM_OK;
    hipError_t e = hipEventSynchronize(c->ev_done);
    if (e != hipSuccess) { c->pending = false; return fail(c, PM_ERR_HIP, "march failed: %s", hipGetErrorString(e)); }
    return complete(c);
}

int pm_download(pm_ctx *c, float *positions, float *normals, float *t)
{
    if (!c) return PM_ERR_ARG;
    if (c->pending) return fail(c, PM_ERR_STATE, "pm_download: march in flight; call pm_wait first");
    if (c->npx == 0) return fail(c, PM_ERR_STATE, "pm_download: nothing marched yet");
    HIPCHK(c, hipSetDevice(c->device));
    if (positions) HIPCHK(c, hipMemcpy(positions, c->d_pos.p, 16 * c->npx, hipMemcpyDeviceToHost));
    if (normals) HIPCHK(c, hipMemcpy(normals, c->d_nrm.p, 16 * c->npx, hipMemcpyDeviceToHost));
    if (t) HIPCHK(c, hipMemcpy(t, c->d_t.p, 4 * c->npx, hipMemcpyDeviceToHost));
    return PM_OK;
}

int pm_get_stats(pm_ctx *c, pm_stats *out)
{
    if (!c || !out) return PM_ERR_ARG;
    *out = c->stats;
    return PM_OK;
}

static int run_query(pm_ctx *c, int32_t frame, uint64_t nq, Frame **fo)
{
    if (!c) return PM_ERR_ARG;
    if (nq == 0 || nq >= (1ull << 31)) return fail(c, PM_ERR_ARG, "query count out of range");
    if (c->pending) return fail(c, PM_ERR_STATE, "a march is in flight");
    Frame *f = get_frame(c, frame);
    if (!f) return fail(c, PM_ERR_FRAME, "frame %d not set", frame);
    HIPCHK(c, hipSetDevice(c->device));
    *fo = f;
    return PM_OK;
}

int pm_neighbor_counts(pm_ctx *c, int32_t frame, const float *q, uint64_t nq, uint32_t *counts, uint32_t *ncand)
{
    Frame *f;
    int rc = run_query(c, frame, nq, &f);
    if (rc) return rc;
    if (!q || !counts) return fail(c, PM_ERR_ARG, "q and counts are required");
    HIPCHK(c, c->d_q.ensure(12 * nq));
    HIPCHK(c, c->d_out.ensure(8 * nq));
    HIPCHK(c, hipMemcpyAsync(c->d_q.p, q, 12 * nq, hipMemcpyHostToDevice, c->stream));
    uint32_t *dc = (uint32_t *)c->d_out.p, *dn = dc + nq;
    HIPCHK(c, launch_neighbor_counts(f->g, (const float *)c->d_q.p, (uint32_t)nq, dc, dn, c->stream));
    HIPCHK(c, hipMemcpyAsync(counts, dc, 4 * nq, hipMemcpyDeviceToHost, c->stream));
    if (ncand) HIPCHK(c, hipMemcpyAsync(ncand, dn, 4 * nq, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return PM_OK;
}

int pm_density(pm_ctx *c, int32_t frame, const float *q, uint64_t nq, int compensated, float *rho)
{
    Frame *f;
    int rc = run_query(c, frame, nq, &f);
    if (rc) return rc;
    if (!q || !rho) return fail(c, PM_ERR_ARG, "q and rho are required");
    HIPCHK(c, c->d_q.ensure(12 * nq));
    HIPCHK(c, c->d_out.ensure(4 * nq));
    HIPCHK(c, hipMemcpyAsync(c->d_q.p, q, 12 * nq, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, launch_density(f->g, (const float *)c->d_q.p, (uint32_t)nq, compensated, (float *)c->d_out.p, c->stream));
    HIPCHK(c, hipMemcpyAsync(rho, c->d_out.p, 4 * nq, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return PM_OK;
}

int pm_density64(pm_ctx *c, int32_t frame, const double *q, uint64_t nq, double *rho)
{
    Frame *f;
    int rc = run_query(c, frame, nq, &f);
    if (rc) return rc;
    if (!q || !rho) return fail(c, PM_ERR_ARG, "q and rho are required");
    HIPCHK(c, c->d_q.ensure(24 * nq));
    HIPCHK(c, c->d_out.ensure(8 * nq));
    HIPCHK(c, hipMemcpyAsync(c->d_q.p, q, 24 * nq, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, launch_density64(f->g, (const double *)c->d_q.p, (uint32_t)nq, (double *)c->d_out.p, c->stream));
    HIPCHK(c, hipMemcpyAsync(rho, c->d_out.p, 8 * nq, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return PM_OK;
}

}  // extern "C"
