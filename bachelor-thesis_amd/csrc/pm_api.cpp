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
#include <algorithm>
#include <map>
#include <random>
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
    DevBuf sorted, cell_start, cell_of, counts, cursor, scan_tmp, xyz, bounds_partial, bounds, scattered;
    DevBuf empty27, dg_box, dg_cons, dg_count, coarse, bound;
    DevBuf fsorted, fcell_start;    // march grid (cells of h/2)
    bool bound_ok = false;      // bound holds the sub-box bits for bound_iso and bound_boxes
    hipEvent_t bound_ev[2] = { nullptr, nullptr };   // around the last bound build
    bool dg_ok = false;         // the DensityGrid flags are built
    float bound_iso = 0.0f;
    int32_t bound_boxes = 0;    // pm_settings.bound_boxes of those bits
    int32_t dg_dims[3] = { 0, 0, 0 };
    CoarseDev cg{};
    void release()
    {
        coarse.release();
        sorted.release(); cell_start.release(); cell_of.release(); counts.release(); cursor.release();
        scan_tmp.release(); xyz.release(); bounds_partial.release(); bounds.release(); scattered.release();
        empty27.release(); dg_box.release(); dg_cons.release(); dg_count.release(); bound.release();
        fsorted.release(); fcell_start.release();
        for (hipEvent_t &e : bound_ev)
            if (e) { hipEventDestroy(e); e = nullptr; }
    }
};

}  // namespace

struct pm_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipStream_t stream_build = nullptr;   // frame builds of pm_set_frame_stream
    int32_t pending_frame = 0;            // frame of the march in flight
    bool pending_batch = false;           // the march in flight is a batch (all its frames)
    bool outputs_on_device = false;       // the last march wrote the caller's device memory
    bool bound_used = false;              // a march has used the bound skip ...
    float bound_iso_last = 0.0f;          // ... with this iso (its bits' key)
    int32_t bound_boxes_last = 0;         // ... and these sub-boxes
    hipEvent_t ev_launch = nullptr, ev_k0 = nullptr, ev_k1 = nullptr, ev_done = nullptr;
    // a depth-sprite pass into the caller's device memory, queued on `stream`
    // and not yet waited for (pm_depth_sprites_device)
    hipEvent_t ev_sprites = nullptr;
    bool sprites_pending = false;
    int32_t sprites_frame = 0;
    // the static scratch array of pm_march_device_noalloc on this device
    void *scratch = nullptr;
    hipError_t scratch_rc = hipSuccess;
    std::map<int32_t, Frame *> frames;
    std::string err;
    // outputs of the last march
    DevBuf d_pos, d_nrm, d_t, d_depth, d_counters, d_q, d_out, d_band, d_tiles;
    DevBuf d_hits;                     // per-pixel hit records of one march
    DevBuf d_gauss;                    // the smoothing tables for N = 0..31, written once
    DevBuf d_sbig;                     // the sprite pass's large-footprint list (count + indices)
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
    // learned step proposal: FP32 master weights, Adam moments, BF16 copies
    DevBuf mlp_p, mlp_m, mlp_v, mlp_W1, mlp_W2, mlp_W2T, mlp_partial, mlp_out, mlp_x, mlp_idx;
    bool mlp_ready = false;
    int mlp_step = 0;
    // training set written by the trace logger
    DevBuf ds_feat, ds_label, ds_hit, ds_count, ds_meta;
    uint64_t ds_cap = 0;
    // batched views: job and tile-chunk tables, all views' images, the captured launch
    uint64_t frame_epoch = 0;           // bumped whenever a frame's device buffers may move
    uint64_t batch_epoch = ~0ull;
    DevBuf b_jobs, b_chunks, b_img, b_ctr, b_flags, b_hits;
    unsigned long long b_rays0 = 0;     // rays of inactive tiles (written once as misses)
    bool prepass = false;               // the last single-view march ran the tile pre-pass
    std::vector<uint64_t> b_off;        // per view: first pixel in b_img
    std::vector<uint32_t> b_w, b_h;
    uint64_t b_npx = 0;
    BatchLaunch b_launch{};
    uint32_t b_blocks = 0;
    hipGraph_t b_graph = nullptr;
    hipGraphExec_t b_exec = nullptr;
    bool b_ready = false;
};

static MlpDev mlp_dev(pm_ctx *c)
{
    MlpDev n{};
    n.W1 = (const uint16_t *)c->mlp_W1.p;
    n.W2 = (const uint16_t *)c->mlp_W2.p;
    n.W2T = (const uint16_t *)c->mlp_W2T.p;
    n.p = (const float *)c->mlp_p.p;
    return n;
}

// The learned step's network and knobs (defaults: 70 % of the predicted
// distance, jumps of 3 to 8 samples, proposals below a quarter of iso).
static void set_learned(pm_ctx *c, const pm_settings *s, MarchLaunch &L)
{
    L.net = mlp_dev(c);
    L.mlp_alpha = s->learned_alpha > 0.0f ? s->learned_alpha : 0.7f;
    L.mlp_min_skip = s->learned_min_skip > 0 ? s->learned_min_skip : 3;
    L.mlp_max_skip = s->learned_max_skip > 0 ? s->learned_max_skip : 8;
    L.mlp_gate_rho = (s->learned_gate > 0.0f ? s->learned_gate : 0.25f) * s->iso;
}

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

// Every transfer goes through the context stream and waits for it: the stream
// is non-blocking, so null-stream copies or fills would not be ordered with the
// kernels queued on it.
static hipError_t copy_sync(pm_ctx *c, void *dst, const void *src, size_t bytes, hipMemcpyKind kind)
{
    hipError_t e = hipMemcpyAsync(dst, src, bytes, kind, c->stream);
    if (e != hipSuccess) return e;
    return hipStreamSynchronize(c->stream);
}

static hipError_t zero_sync(pm_ctx *c, void *dst, size_t bytes)
{
    hipError_t e = hipMemsetAsync(dst, 0, bytes, c->stream);
    if (e != hipSuccess) return e;
    return hipStreamSynchronize(c->stream);
}

// Every point of a BGEO file, read as the data arrives (pm_bgeo.cpp).
int pm_bgeo_read_all(const char *path, std::vector<float> &out);

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
        hipStreamCreateWithFlags(&c->stream_build, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&c->ev_launch) != hipSuccess || hipEventCreate(&c->ev_k0) != hipSuccess ||
        hipEventCreate(&c->ev_k1) != hipSuccess || hipEventCreate(&c->ev_done) != hipSuccess ||
        hipEventCreate(&c->ev_sprites) != hipSuccess ||
        c->d_counters.ensure(sizeof(unsigned long long) * kCtrWords) != hipSuccess ||
        c->d_band.ensure(sizeof(uint32_t) * 64) != hipSuccess ||
        hipHostMalloc((void **)&c->h_counters, sizeof(unsigned long long) * kCtrWords, hipHostMallocDefault) != hipSuccess) {
        pm_destroy(c);
        return PM_ERR_HIP;
    }
    // set-up of pm_march_device_noalloc's scratch (once per device); a failure
    // is reported by that entry point
    c->scratch_rc = noalloc_scratch_init(device, &c->scratch);
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
    c->d_counters.release(); c->d_q.release(); c->d_out.release(); c->d_band.release(); c->d_tiles.release();
    c->d_sbig.release();
    c->d_gauss.release();
    c->d_hits.release();
    c->mlp_p.release(); c->mlp_m.release(); c->mlp_v.release(); c->mlp_W1.release(); c->mlp_W2.release();
    c->mlp_W2T.release(); c->mlp_partial.release(); c->mlp_out.release(); c->mlp_x.release(); c->mlp_idx.release();
    c->ds_feat.release(); c->ds_label.release(); c->ds_hit.release(); c->ds_count.release(); c->ds_meta.release();
    if (c->b_exec) hipGraphExecDestroy(c->b_exec);
    if (c->b_graph) hipGraphDestroy(c->b_graph);
    c->b_jobs.release(); c->b_chunks.release(); c->b_img.release(); c->b_ctr.release(); c->b_flags.release();
    c->b_hits.release();
    if (c->h_stage) hipHostFree(c->h_stage);
    if (c->h_counters) hipHostFree(c->h_counters);
    if (c->ev_launch) hipEventDestroy(c->ev_launch);
    if (c->ev_k0) hipEventDestroy(c->ev_k0);
    if (c->ev_k1) hipEventDestroy(c->ev_k1);
    if (c->ev_done) hipEventDestroy(c->ev_done);
    if (c->ev_sprites) hipEventDestroy(c->ev_sprites);
    if (c->stream) hipStreamDestroy(c->stream);
    if (c->stream_build) {
        hipStreamSynchronize(c->stream_build);
        hipStreamDestroy(c->stream_build);
    }
    delete c;
}

const char *pm_last_error(const pm_ctx *c) { return c ? c->err.c_str() : "null context"; }

static int build_frame(pm_ctx *c, int32_t frame, const float *d_src, uint64_t n, float h, hipStream_t st)
{
    Frame *f;
    auto it = c->frames.find(frame);
    if (it == c->frames.end()) { f = new Frame(); c->frames[frame] = f; }
    else f = it->second;
    c->frame_epoch++;   // a batch set before this holds the old grid's pointers
    f->bound_ok = false;

    hipEvent_t e0, e1;
    HIPCHK(c, hipEventCreate(&e0));
    HIPCHK(c, hipEventCreate(&e1));
    HIPCHK(c, hipEventRecord(e0, st));
    uint32_t np = bounds_partials(n);
    HIPCHK(c, f->bounds_partial.ensure(sizeof(float) * 6 * np));
    HIPCHK(c, f->bounds.ensure(sizeof(float) * 6));
    HIPCHK(c, launch_bounds_reduce(d_src, n, (float *)f->bounds_partial.p, np, (float *)f->bounds.p, st));
    float b[6];
    HIPCHK(c, hipMemcpyAsync(b, f->bounds.p, sizeof(b), hipMemcpyDeviceToHost, st));
    HIPCHK(c, hipStreamSynchronize(st));
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
    // march grid: cells of h/2 (inv 2/h, exact) and four empty cells of
    // margin, so the 5x5x5 fine stencil of any point within h of a particle
    // stays inside it
    const float inv_hf = 2.0f * inv_h;
    uint64_t fnc = 1;
    for (int ax = 0; ax < 3; ax++) {
        int32_t kmin = host_key(I.pmin[ax], inv_hf), kmax = host_key(I.pmax[ax], inv_hf);
        I.forigin[ax] = kmin - 4;
        I.fdims[ax] = kmax - kmin + 9;
        fnc *= (uint64_t)I.fdims[ax];
    }
    if (fnc >= (1ull << 31)) return fail(c, PM_ERR_SIZE, "pm_set_frame: march grid of %llu cells exceeds 2^31", (unsigned long long)fnc);
    I.fncells = fnc;

    // scratch of both builds sized once, before either is queued
    const uint64_t mc = fnc > nc ? fnc : nc;
    HIPCHK(c, f->sorted.ensure(sizeof(float4) * n));
    HIPCHK(c, f->cell_start.ensure(sizeof(uint32_t) * (nc + 1)));
    HIPCHK(c, f->fsorted.ensure(sizeof(float4) * n));
    HIPCHK(c, f->fcell_start.ensure(sizeof(uint32_t) * (fnc + 1)));
    HIPCHK(c, f->cell_of.ensure(sizeof(uint32_t) * n));
    HIPCHK(c, f->scattered.ensure(sizeof(float4) * n));
    HIPCHK(c, f->counts.ensure(sizeof(uint32_t) * (mc + 1)));
    HIPCHK(c, f->cursor.ensure(sizeof(uint32_t) * mc));
    HIPCHK(c, f->scan_tmp.ensure(sizeof(uint32_t) * scan_tmp_size((uint32_t)(mc + 1))));

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
    a.scattered = (float4 *)f->scattered.p;
    HIPCHK(c, launch_grid_build(a, st));

    GridBuildArgs fa = a;   // the march grid: same build, cells of h/2
    fa.inv_h = inv_hf;
    fa.ox = I.forigin[0]; fa.oy = I.forigin[1]; fa.oz = I.forigin[2];
    fa.nx = I.fdims[0]; fa.ny = I.fdims[1]; fa.nz = I.fdims[2];
    fa.ncells = (uint32_t)fnc;
    fa.counts = (uint32_t *)f->counts.p;
    fa.cursor = (uint32_t *)f->cursor.p;
    fa.scan_tmp = (uint32_t *)f->scan_tmp.p;
    fa.sorted = (float4 *)f->fsorted.p;
    fa.cell_start = (uint32_t *)f->fcell_start.p;
    HIPCHK(c, launch_grid_build(fa, st));
    HIPCHK(c, hipEventRecord(e1, st));
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
    g.fsorted = (const float4 *)f->fsorted.p;
    g.fcell_start = (const uint32_t *)f->fcell_start.p;
    g.fox = I.forigin[0]; g.foy = I.forigin[1]; g.foz = I.forigin[2];
    g.fnx = I.fdims[0]; g.fny = I.fdims[1]; g.fnz = I.fdims[2];
    g.inv_hf = inv_hf;

    // skip structures: empty 27-neighbourhoods and the reference DensityGrid
    HIPCHK(c, f->empty27.ensure(nc));
    HIPCHK(c, launch_empty27((const uint32_t *)f->cell_start.p, I.dims[0], I.dims[1], I.dims[2], (uint8_t *)f->empty27.p, st));
    g.empty27 = (const uint8_t *)f->empty27.p;
    uint64_t dgn = 1;
    for (int a2 = 0; a2 < 3; a2++) {
        float cw = 1.0f * h;                            // Dataset.cpp:96
        f->dg_dims[a2] = (int32_t)std::ceil((I.bmax[a2] - I.bmin[a2]) / cw);   // :102-104
        dgn *= (uint64_t)f->dg_dims[a2];
    }
    if (dgn >= (1ull << 31)) return fail(c, PM_ERR_SIZE, "pm_set_frame: DensityGrid too large");
    g.dg_nx = f->dg_dims[0]; g.dg_ny = f->dg_dims[1]; g.dg_nz = f->dg_dims[2];
    g.dg_flag = nullptr;
    f->dg_ok = false;   // built on first use (reference march, pm_get_skip_grids)
    // coarse occupancy (4x4x4 dense cells) for the step-proposal features
    f->cg.nx = (I.dims[0] + 3) / 4; f->cg.ny = (I.dims[1] + 3) / 4; f->cg.nz = (I.dims[2] + 3) / 4;
    HIPCHK(c, f->coarse.ensure(4ull * f->cg.nx * f->cg.ny * f->cg.nz));
    f->cg.count = (const uint32_t *)f->coarse.p;
    HIPCHK(c, launch_coarse_counts(g, f->cg, (uint32_t *)f->coarse.p, st));
    HIPCHK(c, hipStreamSynchronize(st));
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
    return build_frame(c, frame, (const float *)f->xyz.p, n, h, c->stream);
}

static int ensure_bound(pm_ctx *c, Frame *f, float iso, int32_t boxes, hipStream_t st);

int pm_set_frame_stream(pm_ctx *c, int32_t frame, const float *xyz, uint64_t n, float h)
{
    if (!c) return PM_ERR_ARG;
    if (!xyz || n == 0 || n >= (1ull << 31) || !(h > 0.0f) || !std::isfinite(h))
        return fail(c, PM_ERR_ARG, "pm_set_frame_stream: need xyz, 0 < n < 2^31 and finite h > 0");
    if (c->pending && (c->pending_batch || c->pending_frame == frame))
        return fail(c, PM_ERR_STATE, "pm_set_frame_stream: the march in flight reads frame %d", frame);
    HIPCHK(c, hipSetDevice(c->device));
    auto it = c->frames.find(frame);
    Frame *f = it == c->frames.end() ? (c->frames[frame] = new Frame()) : it->second;
    // a depth-sprite pass queued on the context stream may still read this
    // frame's grid: the rebuild on the build stream starts after it
    if (c->sprites_pending && c->sprites_frame == frame) HIPCHK(c, hipStreamWaitEvent(c->stream_build, c->ev_sprites, 0));
    HIPCHK(c, f->xyz.ensure(sizeof(float) * 3 * n));
    HIPCHK(c, hipMemcpyAsync(f->xyz.p, xyz, sizeof(float) * 3 * n, hipMemcpyHostToDevice, c->stream_build));
    int rc = build_frame(c, frame, (const float *)f->xyz.p, n, h, c->stream_build);
    if (rc || !c->bound_used) return rc;
    // Autoplay marches every frame with the same settings: the skip's bits for
    // the iso of the last bound-skip march are built here too, beside the march
    // in flight, instead of at the start of this frame's march.  A march with
    // another iso or sub-box size rebuilds them (the bits are keyed by both).
    rc = ensure_bound(c, f, c->bound_iso_last, c->bound_boxes_last, c->stream_build);
    if (rc) return rc;
    HIPCHK(c, hipStreamSynchronize(c->stream_build));
    return PM_OK;
}

int pm_load_bgeo(pm_ctx *c, int32_t frame, const char *path, float h)
{
    if (!c) return PM_ERR_ARG;
    std::vector<float> xyz;
    const int rc = pm_bgeo_read_all(path, xyz);
    if (rc) return fail(c, rc, "pm_load_bgeo: cannot read a BGEO V5 file at %s", path ? path : "(null)");
    const uint64_t n = xyz.size() / 3;
    if (n == 0) return fail(c, PM_ERR_ARG, "pm_load_bgeo: %s holds no points", path);
    return c->pending ? pm_set_frame_stream(c, frame, xyz.data(), n, h) : pm_set_frame(c, frame, xyz.data(), n, h);
}

int pm_set_frame_device(pm_ctx *c, int32_t frame, const void *d_xyz, uint64_t n, float h)
{
    if (!c) return PM_ERR_ARG;
    if (!d_xyz || n == 0 || n >= (1ull << 31) || !(h > 0.0f) || !std::isfinite(h))
        return fail(c, PM_ERR_ARG, "pm_set_frame_device: need d_xyz, 0 < n < 2^31 and finite h > 0");
    if (c->pending) return fail(c, PM_ERR_STATE, "pm_set_frame_device: a march is in flight");
    HIPCHK(c, hipSetDevice(c->device));
    return build_frame(c, frame, (const float *)d_xyz, n, h, c->stream);
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
    out->bound_ms = -1.0f;
    out->bound_iso = f->bound_ok ? f->bound_iso : 0.0f;
    if (f->bound_ok && f->bound_ev[1]) {
        HIPCHK(c, hipEventSynchronize(f->bound_ev[1]));
        HIPCHK(c, hipEventElapsedTime(&out->bound_ms, f->bound_ev[0], f->bound_ev[1]));
    }
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

// The reference's DensityGrid flags of a frame (Frame::BuildDensityGrid,
// Dataset.cpp:94-164), built on the context stream when first needed.
static int ensure_densitygrid(pm_ctx *c, Frame *f)
{
    if (f->dg_ok) return PM_OK;
    const uint64_t dgn = (uint64_t)f->dg_dims[0] * f->dg_dims[1] * f->dg_dims[2];
    HIPCHK(c, f->dg_box.ensure(dgn));
    HIPCHK(c, f->dg_cons.ensure(dgn));
    HIPCHK(c, f->dg_count.ensure(4 * dgn));
    HIPCHK(c, launch_densitygrid(f->g, f->dg_dims[0], f->dg_dims[1], f->dg_dims[2], (uint8_t *)f->dg_box.p,
                                 (uint8_t *)f->dg_cons.p, (uint32_t *)f->dg_count.p, c->stream));
    f->dg_ok = true;
    return PM_OK;
}

// Sub-box density bounds of a frame for one iso value, built on the context
// stream when first needed (a batch holding the old bits is invalidated).
static int ensure_bound(pm_ctx *c, Frame *f, float iso, int32_t boxes, hipStream_t st)
{
    if (f->bound_ok && f->bound_iso == iso && f->bound_boxes == boxes) return PM_OK;
    HIPCHK(c, f->bound.ensure(sizeof(uint64_t) * 8 * f->info.ncells));
    f->g.bound = (const uint64_t *)f->bound.p;
    for (hipEvent_t &e : f->bound_ev)
        if (!e) HIPCHK(c, hipEventCreate(&e));
    HIPCHK(c, hipEventRecord(f->bound_ev[0], st));
    HIPCHK(c, launch_bound_grid(f->g, iso, boxes != 0, (uint64_t *)f->bound.p, st));
    HIPCHK(c, hipEventRecord(f->bound_ev[1], st));
    f->bound_ok = true;
    f->bound_iso = iso;
    f->bound_boxes = boxes;
    c->frame_epoch++;
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
    if (s->skip_mode < PM_SKIP_NONE || s->skip_mode > PM_SKIP_BOUND) return fail(c, PM_ERR_ARG, "unknown skip_mode %d", s->skip_mode);
    if (s->bound_boxes != 0 && s->bound_boxes != 1) return fail(c, PM_ERR_ARG, "bound_boxes must be 0 or 1");
    if (s->mode == PM_MODE_CONFIG && (s->skip_mode == PM_SKIP_DG_BOX || s->skip_mode == PM_SKIP_DG_CONS))
        return fail(c, PM_ERR_ARG, "the DensityGrid skip belongs to the reference march");
    if (s->mode == PM_MODE_REFERENCE && (s->skip_mode == PM_SKIP_EMPTY || s->skip_mode == PM_SKIP_BOUND))
        return fail(c, PM_ERR_ARG, "the empty-stencil and bound skips belong to the configured march");
    if (s->schedule != 0 && s->schedule != 1) return fail(c, PM_ERR_ARG, "unknown kernel schedule %d", s->schedule);
    if (s->learned && !c->mlp_ready) return fail(c, PM_ERR_STATE, "learned step: no network loaded (pm_mlp_set)");
    if (s->learned && s->mode != PM_MODE_CONFIG) return fail(c, PM_ERR_ARG, "the learned step belongs to the configured march");
    if (s->learned && !(s->learned_alpha >= 0.0f && s->learned_alpha <= 4.0f)) return fail(c, PM_ERR_ARG, "learned_alpha out of [0,4]");
    if (s->learned && !(s->learned_gate >= 0.0f && s->learned_gate <= 1.0f)) return fail(c, PM_ERR_ARG, "learned_gate out of [0,1]");
    if (s->learned && (s->learned_min_skip < 0 || s->learned_min_skip == 1 || s->learned_max_skip < 0 ||
                       s->learned_max_skip > 64))
        return fail(c, PM_ERR_ARG, "learned: min_skip must be 0 or >= 2, max_skip in [0, 64]");
    if (s->instrument && (s->mode != PM_MODE_CONFIG || s->anisotropic || s->schedule != 0))
        return fail(c, PM_ERR_ARG, "instrument: configured isotropic march, schedule 0 only");
    if (s->mode == PM_MODE_REFERENCE && (s->max_steps < 0 || s->max_steps > (1 << 24)))
        return fail(c, PM_ERR_ARG, "max_steps out of range");
    if (s->anisotropic) {
        if (!(s->k_n > 0.0f && s->k_r > 0.0f && s->k_s > 0.0f) || !std::isfinite(s->k_n) || !std::isfinite(s->k_r) ||
            !std::isfinite(s->k_s))
            return fail(c, PM_ERR_ARG, "anisotropic: k_n, k_r and k_s must be finite and > 0");
        // the extended walk covers the cells within 2h of the sample
        if (s->ext_mult != 0.0f && !(s->ext_mult >= 1.0f && s->ext_mult <= 2.0f))
            return fail(c, PM_ERR_ARG, "anisotropic: ext_mult must be 0 (= 2) or in [1, 2]");
        if (s->learned) return fail(c, PM_ERR_ARG, "anisotropic: no learned step");
    }
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

// Anisotropic parameters of a frame (h = ParticleRadius) and settings, as
// Dataset.cpp:169-177 and Kernel.cpp:55-61 derive them, in FP32 and FP64.
static AnisoDev aniso_params(const Frame *f, const pm_settings *s)
{
    AnisoDev a{};
    const float mult = s->ext_mult != 0.0f ? s->ext_mult : 2.0f;
    AnParF &F = a.f;
    F.h = f->info.h; F.h2 = F.h * F.h; F.hinv = 1.0f / F.h;
    F.hx = mult * F.h; F.R2x = F.hx * F.hx; F.hxinv = 1.0f / F.hx;
    F.sig = 8.0f / PM_PI_F;
    F.k_n = s->k_n; F.k_r = s->k_r; F.k_s = s->k_s; F.n_eps = s->n_eps;
    AnParD &D = a.d;
    D.h = (double)f->info.h; D.h2 = D.h * D.h; D.hinv = 1.0 / D.h;
    D.hx = (double)mult * D.h; D.R2x = D.hx * D.hx; D.hxinv = 1.0 / D.hx;
    D.sig = 8.0 / PM_PI_D;
    D.k_n = s->k_n; D.k_r = s->k_r; D.k_s = s->k_s; D.n_eps = s->n_eps;
    return a;
}

// True when [p, p + bytes) is device (or managed) memory of `device`: a kernel
// handed a host address would fault.  The allocation's extent is checked where
// HIP reports it (memory imported from other APIs may not report one).
static bool device_range(int device, const void *p, size_t bytes)
{
    if (!p) return false;
    hipPointerAttribute_t a{};
    if (hipPointerGetAttributes(&a, p) != hipSuccess) { (void)hipGetLastError(); return false; }
    if ((a.type != hipMemoryTypeDevice && a.type != hipMemoryTypeManaged) || a.device != device) return false;
    hipDeviceptr_t base = nullptr;
    size_t size = 0;
    if (hipMemGetAddressRange(&base, &size, (hipDeviceptr_t)p) != hipSuccess) { (void)hipGetLastError(); return true; }
    return (const char *)p + bytes <= (const char *)base + size;
}

// The march of one view.  dev != nullptr: the outputs go straight into the
// caller's device memory dev[0..2] (t may be null) and depth_in is device
// memory; otherwise into the context's buffers, copied to the host pointers
// (any may be null) at completion.
// na != nullptr (pm_march_device_noalloc): every scratch buffer comes from the
// static array and nothing is allocated (the bound bits must exist, fills are
// kernels).
static int march_launch(pm_ctx *c, int32_t frame, const pm_settings *s, const float inv_pv[16], const float cam[3],
                        uint32_t W, uint32_t H, const float *depth_in, float *positions_out, float *normals_out,
                        float *t_out, void *const *dev, const NoAllocScratch *na = nullptr)
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

    if (s->skip_mode == PM_SKIP_BOUND) {
        // the anisotropic density is exactly 0 where no particle lies within h:
        // the bound bits for PM_ZERO_ISO mark those sub-boxes
        const float biso = s->anisotropic ? PM_ZERO_ISO : s->iso;
        if (na) {
            if (!(f->bound_ok && f->bound_iso == biso && f->bound_boxes == s->bound_boxes))
                return fail(c, PM_ERR_STATE, "pm_march_device_noalloc: frame %d has no bound bits for iso %g "
                            "and bound_boxes %d (build them first: pm_get_bound_grid or one pm_march_async_device)",
                            frame, (double)biso, s->bound_boxes);
        } else {
            rc = ensure_bound(c, f, biso, s->bound_boxes, c->stream);
            if (rc) return rc;
        }
        c->bound_iso_last = biso;
        c->bound_boxes_last = s->bound_boxes;
        c->bound_used = true;
    }
    if (s->skip_mode == PM_SKIP_DG_BOX || s->skip_mode == PM_SKIP_DG_CONS) {
        rc = ensure_densitygrid(c, f);
        if (rc) return rc;
    }
    uint64_t npx = (uint64_t)W * H;
    if (dev) {
        if (!device_range(c->device, dev[0], 16 * npx) || !device_range(c->device, dev[1], 16 * npx) ||
            (dev[2] && !device_range(c->device, dev[2], 4 * npx)) ||
            (depth_in && !device_range(c->device, depth_in, 4 * npx)))
            return fail(c, PM_ERR_ARG, "pm_march_async_device: images must be device memory of GPU %d, W*H long",
                        c->device);
    }
    if (!dev) {
        HIPCHK(c, c->d_pos.ensure(sizeof(float4) * npx));
        HIPCHK(c, c->d_nrm.ensure(sizeof(float4) * npx));
        HIPCHK(c, c->d_t.ensure(sizeof(float) * npx));
    }
    c->npx = npx;
    c->outputs_on_device = dev != nullptr;

    MarchLaunch L{};
    L.g = f->g;
    L.g.dg_flag = (const uint8_t *)(s->skip_mode == PM_SKIP_DG_BOX ? f->dg_box.p : f->dg_cons.p);
    L.s = *s;
    L.inv_step = 1.0f / s->step;
    memcpy(L.M, inv_pv, sizeof(L.M));
    memcpy(L.cam, cam, sizeof(L.cam));
    L.W = W; L.H = H;
    L.pos = dev ? (float4 *)dev[0] : (float4 *)c->d_pos.p;
    L.nrm = dev ? (float4 *)dev[1] : (float4 *)c->d_nrm.p;
    L.t = dev ? (float *)dev[2] : (float *)c->d_t.p;
    L.counters = (unsigned long long *)c->d_counters.p;
    L.band_ctr = (uint32_t *)c->d_band.p;
    if (na) {
        L.counters = na->counters;
        L.band_ctr = na->band;
        L.hit_rec = na->hit_rec;
    } else if (s->mode == PM_MODE_CONFIG) {   // hit records; the anisotropic march's hit list
        HIPCHK(c, c->d_hits.ensure(sizeof(uint4) * npx));
        L.hit_rec = (uint4 *)c->d_hits.p;
    }
    if (s->anisotropic) L.an = aniso_params(f, s);
    // tile pre-pass (active-tile list, bulk miss writes) for the persistent
    // schedule, the anisotropic march and the reference march
    const bool persistent = s->schedule == 0 || (!s->anisotropic && s->learned);
    if (persistent) {
        uint64_t tiles = (uint64_t)((W + 7) / 8) * ((H + 7) / 8);
        if (na) {
            L.tile_buf = na->tiles;
        } else {
            HIPCHK(c, c->d_tiles.ensure(sizeof(uint32_t) * (3 * (tiles + 1) + scan_tmp_size((uint32_t)(tiles + 1)))));
            L.tile_buf = (uint32_t *)c->d_tiles.p;
        }
    }
    c->prepass = L.tile_buf != nullptr;
    L.cg = f->cg;
    if (s->learned) set_learned(c, s, L);

    HIPCHK(c, hipEventRecord(c->ev_launch, c->stream));
    if (s->mode == PM_MODE_REFERENCE) {
        if (dev) {
            L.depth = depth_in;
        } else {
            HIPCHK(c, c->d_depth.ensure(sizeof(float) * npx));
            HIPCHK(c, hipMemcpyAsync(c->d_depth.p, depth_in, sizeof(float) * npx, hipMemcpyHostToDevice, c->stream));
            L.depth = (const float *)c->d_depth.p;
        }
    }
    if (na) HIPCHK(c, launch_zero_u64(L.counters, kCtrWords, c->stream));
    else HIPCHK(c, hipMemsetAsync(c->d_counters.p, 0, sizeof(unsigned long long) * kCtrWords, c->stream));
    HIPCHK(c, hipEventRecord(c->ev_k0, c->stream));
    HIPCHK(c, launch_march(L, c->stream));
    HIPCHK(c, hipEventRecord(c->ev_k1, c->stream));
    HIPCHK(c, hipMemcpyAsync(c->h_counters, L.counters, sizeof(unsigned long long) * kCtrWords, hipMemcpyDeviceToHost, c->stream));

    if (dev) positions_out = normals_out = t_out = nullptr;
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
    c->pending_frame = frame;
    c->pending_batch = false;
    c->finished = false;
    return PM_OK;
}

int pm_march_async(pm_ctx *c, int32_t frame, const pm_settings *s, const float inv_pv[16], const float cam[3],
                   uint32_t W, uint32_t H, const float *depth_in, float *positions_out, float *normals_out,
                   float *t_out)
{
    return march_launch(c, frame, s, inv_pv, cam, W, H, depth_in, positions_out, normals_out, t_out, nullptr);
}

int pm_march_async_device(pm_ctx *c, int32_t frame, const pm_settings *s, const float inv_pv[16],
                          const float cam[3], uint32_t W, uint32_t H, const float *d_depth, void *d_positions,
                          void *d_normals, void *d_t)
{
    if (!c) return PM_ERR_ARG;
    if (!d_positions || !d_normals) return fail(c, PM_ERR_ARG, "pm_march_async_device: need device positions and normals");
    void *dev[3] = { d_positions, d_normals, d_t };
    return march_launch(c, frame, s, inv_pv, cam, W, H, d_depth, nullptr, nullptr, nullptr, dev);
}

int pm_march_device_noalloc(pm_ctx *c, int32_t frame, const pm_settings *s, const float inv_pv[16],
                            const float cam[3], uint32_t W, uint32_t H, void *d_positions, void *d_normals, void *d_t)
{
    if (!c) return PM_ERR_ARG;
    if (!d_positions || !d_normals) return fail(c, PM_ERR_ARG, "pm_march_device_noalloc: need device positions and normals");
    if (!s) return fail(c, PM_ERR_ARG, "settings is NULL");
    if (s->mode != PM_MODE_CONFIG || s->anisotropic || s->schedule != 0)
        return fail(c, PM_ERR_ARG, "pm_march_device_noalloc: the configured isotropic march, schedule 0");
    if (W == 0 || H == 0 || W > 16384 || H > 16384 || (uint64_t)W * H > PM_NOALLOC_MAX_PIXELS)
        return fail(c, PM_ERR_SIZE, "pm_march_device_noalloc: %ux%u is beyond the scratch limits (W, H <= 16384, "
                    "W*H <= %u)", W, H, PM_NOALLOC_MAX_PIXELS);
    if (c->scratch_rc != hipSuccess || !c->scratch)
        return fail(c, PM_ERR_HIP, "pm_march_device_noalloc: no scratch on device %d: %s", c->device,
                    hipGetErrorString(c->scratch_rc));
    NoAllocScratch na{};
    if (!noalloc_layout(c->scratch, W, H, na))
        return fail(c, PM_ERR_SIZE, "pm_march_device_noalloc: %ux%u does not fit the scratch", W, H);
    void *dev[3] = { d_positions, d_normals, d_t };
    return march_launch(c, frame, s, inv_pv, cam, W, H, nullptr, nullptr, nullptr, nullptr, dev, &na);
}

uint64_t pm_noalloc_scratch_bytes(void) { return (uint64_t)noalloc_scratch_bytes(); }

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
    // every pixel is answered once; the rays of the tiles the pre-pass writes
    // as misses are not counted on the device, so a march with the pre-pass
    // reports all W*H rays
    unsigned long long sum[C_COUNT] = {};
    for (int k = 0; k < kCtrSlots; k++)
        for (int i = 0; i < C_COUNT; i++) sum[i] += c->h_counters[k * kCtrStride + i];
    S.rays = c->prepass ? c->npx : sum[C_RAYS];
    S.rays_in_aabb = sum[C_IN_AABB];
    S.hits = sum[C_HITS];
    S.samples = sum[C_SAMPLES];
    S.cand = sum[C_CAND];
    S.march_samples = sum[C_MARCH_SAMPLES];
    S.mlp_evals = sum[C_MLP];
    S.march_cand = sum[C_MARCH_CAND];
    S.cert_walks = sum[C_CERT];
    S.cert_skipped = sum[C_CERT_SKIP];
    S.cand_stencil = sum[C_CAND_STENCIL];
    S.cand27 = sum[C_CAND27];
    hipEventElapsedTime(&S.kernel_ms, c->ev_k0, c->ev_k1);
    hipEventElapsedTime(&S.total_ms, c->ev_launch, c->ev_done);
    c->pending = false;
    c->finished = true;
    c->sprites_pending = false;      // queued before the march on the same stream
    return PM_OK;
}

// A sprite pass into device memory completes with the context stream; pm_query
// and pm_wait report it (and a march queued after it completes after it).
static int sprites_query(pm_ctx *c, bool block)
{
    if (!c->sprites_pending) return PM_OK;
    hipError_t e = block ? hipEventSynchronize(c->ev_sprites) : hipEventQuery(c->ev_sprites);
    if (e == hipErrorNotReady) return PM_PENDING;
    c->sprites_pending = false;
    if (e != hipSuccess) return fail(c, PM_ERR_HIP, "depth-sprite pass failed: %s", hipGetErrorString(e));
    return PM_OK;
}

int pm_query(pm_ctx *c)
{
    if (!c) return PM_ERR_ARG;
    if (!c->pending) {
        const int rs = sprites_query(c, false);
        if (rs != PM_OK) return rs;
        return c->finished ? PM_OK : fail(c, PM_ERR_STATE, "no march launched");
    }
    hipError_t e = hipEventQuery(c->ev_done);
    if (e == hipErrorNotReady) return PM_PENDING;
    if (e != hipSuccess) { c->pending = false; return fail(c, PM_ERR_HIP, "march failed: %s", hipGetErrorString(e)); }
    return complete(c);
}

int pm_wait(pm_ctx *c)
{
    if (!c) return PM_ERR_ARG;
    if (!c->pending) return sprites_query(c, true);
    hipError_t e = hipEventSynchronize(c->ev_done);
    if (e != hipSuccess) { c->pending = false; return fail(c, PM_ERR_HIP, "march failed: %s", hipGetErrorString(e)); }
    return complete(c);
}

int pm_download(pm_ctx *c, float *positions, float *normals, float *t)
{
    if (!c) return PM_ERR_ARG;
    if (c->pending) return fail(c, PM_ERR_STATE, "pm_download: march in flight; call pm_wait first");
    if (c->npx == 0) return fail(c, PM_ERR_STATE, "pm_download: nothing marched yet");
    if (c->outputs_on_device) return fail(c, PM_ERR_STATE, "pm_download: the last march wrote the caller's device memory");
    HIPCHK(c, hipSetDevice(c->device));
    if (positions) HIPCHK(c, copy_sync(c, positions, c->d_pos.p, 16 * c->npx, hipMemcpyDeviceToHost));
    if (normals) HIPCHK(c, copy_sync(c, normals, c->d_nrm.p, 16 * c->npx, hipMemcpyDeviceToHost));
    if (t) HIPCHK(c, copy_sync(c, t, c->d_t.p, 4 * c->npx, hipMemcpyDeviceToHost));
    return PM_OK;
}

int pm_get_stats(pm_ctx *c, pm_stats *out)
{
    if (!c || !out) return PM_ERR_ARG;
    *out = c->stats;
    return PM_OK;
}

// ---- batched views x frames ----

static constexpr uint32_t kChunkTiles = 16;   // 8x8 tiles per queue entry (1024 rays)

static int batch_enqueue_launch(pm_ctx *c)
{
    HIPCHK(c, launch_init_counters(c->b_launch.base.counters, kCtrWords, c->b_rays0, c->stream));
    HIPCHK(c, launch_zero_u32(c->b_launch.chunk_ctr, 1, c->stream));
    HIPCHK(c, launch_march_batch(c->b_launch, c->b_blocks, c->stream));
    return PM_OK;
}

int pm_get_march_grid(pm_ctx *c, int32_t frame, uint32_t *cell_start, float *sorted)
{
    if (!c) return PM_ERR_ARG;
    if (c->pending) return fail(c, PM_ERR_STATE, "pm_get_march_grid: a march is in flight");
    Frame *f = get_frame(c, frame);
    if (!f) return fail(c, PM_ERR_FRAME, "frame %d not set", frame);
    HIPCHK(c, hipSetDevice(c->device));
    if (cell_start)
        HIPCHK(c, copy_sync(c, cell_start, f->fcell_start.p, sizeof(uint32_t) * (f->info.fncells + 1), hipMemcpyDeviceToHost));
    if (sorted) {
        uint64_t n = f->info.n;
        HIPCHK(c, copy_sync(c, sorted, f->fsorted.p, sizeof(float) * 4 * n, hipMemcpyDeviceToHost));
        for (uint64_t i = 0; i < n; i++) sorted[4 * i + 3] = 0.0f;
    }
    return PM_OK;
}

int pm_get_bound_grid(pm_ctx *c, int32_t frame, float iso, int32_t boxes, uint64_t *bits)
{
    if (!c || !bits) return PM_ERR_ARG;
    if (boxes != 0 && boxes != 1) return fail(c, PM_ERR_ARG, "pm_get_bound_grid: boxes must be 0 or 1");
    if (c->pending) return fail(c, PM_ERR_STATE, "pm_get_bound_grid: a march is in flight");
    if (!std::isfinite(iso)) return fail(c, PM_ERR_ARG, "iso must be finite");
    Frame *f = get_frame(c, frame);
    if (!f) return fail(c, PM_ERR_FRAME, "frame %d not set", frame);
    HIPCHK(c, hipSetDevice(c->device));
    int rc = ensure_bound(c, f, iso, boxes, c->stream);
    if (rc) return rc;
    HIPCHK(c, copy_sync(c, bits, f->bound.p, sizeof(uint64_t) * 8 * f->info.ncells, hipMemcpyDeviceToHost));
    return PM_OK;
}

int pm_batch_set(pm_ctx *c, const pm_settings *s, const pm_view *views, uint32_t n_views, int use_graph)
{
    if (!c) return PM_ERR_ARG;
    if (c->pending) return fail(c, PM_ERR_STATE, "pm_batch_set: a march is in flight; call pm_wait first");
    if (!views || n_views == 0 || n_views > (1u << 20)) return fail(c, PM_ERR_ARG, "pm_batch_set: need 1..2^20 views");
    if (!s) return fail(c, PM_ERR_ARG, "settings is NULL");
    if (s->mode != PM_MODE_CONFIG) return fail(c, PM_ERR_ARG, "pm_batch_set: the batch runs the configured march");
    if (s->anisotropic) return fail(c, PM_ERR_ARG, "pm_batch_set: the batch runs the isotropic march");
    if (s->instrument || s->schedule != 0)
        return fail(c, PM_ERR_ARG, "pm_batch_set: the batch runs schedule 0, not instrumented");
    HIPCHK(c, hipSetDevice(c->device));
    c->b_ready = false;
    if (c->b_exec) { hipGraphExecDestroy(c->b_exec); c->b_exec = nullptr; }
    if (c->b_graph) { hipGraphDestroy(c->b_graph); c->b_graph = nullptr; }

    std::vector<MarchLaunch> jobs(n_views);
    c->b_off.assign(n_views, 0);
    c->b_w.assign(n_views, 0);
    c->b_h.assign(n_views, 0);
    uint64_t npx = 0;
    for (uint32_t v = 0; v < n_views; v++) {
        const pm_view &vw = views[v];
        if (vw.width == 0 || vw.height == 0 || (uint64_t)vw.width * vw.height > (1ull << 30))
            return fail(c, PM_ERR_SIZE, "view %u: bad image size %ux%u", v, vw.width, vw.height);
        Frame *f = get_frame(c, vw.frame);
        if (!f) return fail(c, PM_ERR_FRAME, "view %u: frame %d not set", v, vw.frame);
        int rc = validate_settings(c, s, f);
        if (rc) return rc;
        if (s->skip_mode == PM_SKIP_BOUND) {
            rc = ensure_bound(c, f, s->iso, s->bound_boxes, c->stream);
            if (rc) return rc;
        }
        c->b_off[v] = npx;
        c->b_w[v] = vw.width;
        c->b_h[v] = vw.height;
        npx += (uint64_t)vw.width * vw.height;
        if (npx > (1ull << 34)) return fail(c, PM_ERR_SIZE, "pm_batch_set: more than 2^34 rays");
    }
    c->b_npx = npx;
    // images: pos (float4) | nrm (float4) | t (float), each over all views
    HIPCHK(c, c->b_img.ensure(36 * npx));
    float4 *pos = (float4 *)c->b_img.p;
    float4 *nrm = pos + npx;
    float *tt = (float *)(nrm + npx);
    for (uint32_t v = 0; v < n_views; v++) {
        const pm_view &vw = views[v];
        Frame *f = get_frame(c, vw.frame);
        MarchLaunch &J = jobs[v];
        J.g = f->g;
        J.cg = f->cg;
        memcpy(J.M, vw.inv_pv, sizeof(J.M));
        memcpy(J.cam, vw.cam, sizeof(J.cam));
        J.W = vw.width;
        J.H = vw.height;
        J.pos = pos + c->b_off[v];
        J.nrm = nrm + c->b_off[v];
        J.t = tt + c->b_off[v];
    }
    // queue: views by decreasing ray count (longest first), each as runs of
    // kChunkTiles consecutive tiles
    std::vector<uint32_t> order(n_views);
    for (uint32_t v = 0; v < n_views; v++) order[v] = v;
    std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) {
        return (uint64_t)c->b_w[a] * c->b_h[a] > (uint64_t)c->b_w[b] * c->b_h[b];
    });
    // Tiles none of whose rays enters the frame's AABB are answered once, here:
    // their misses are written now and the queue holds only the other tiles
    // (the images are identical; a run marches only tiles with work).
    std::vector<uint64_t> foff(n_views + 1, 0);
    for (uint32_t v = 0; v < n_views; v++)
        foff[v + 1] = foff[v] + (uint64_t)((c->b_w[v] + 7) / 8) * ((c->b_h[v] + 7) / 8) + 1;
    if (foff[n_views] >= (1ull << 32)) return fail(c, PM_ERR_SIZE, "pm_batch_set: too many tiles");
    HIPCHK(c, c->b_flags.ensure(sizeof(uint32_t) * foff[n_views]));
    for (uint32_t v = 0; v < n_views; v++) {
        MarchLaunch Lv{};
        Lv.g = jobs[v].g;
        memcpy(Lv.M, jobs[v].M, sizeof(Lv.M));
        memcpy(Lv.cam, jobs[v].cam, sizeof(Lv.cam));
        Lv.W = jobs[v].W; Lv.H = jobs[v].H;
        Lv.pos = jobs[v].pos; Lv.nrm = jobs[v].nrm; Lv.t = jobs[v].t;
        Lv.counters = (unsigned long long *)c->d_counters.p;
        Lv.tile_buf = (uint32_t *)c->b_flags.p + foff[v];
        uint32_t T = (uint32_t)(foff[v + 1] - foff[v] - 1);
        HIPCHK(c, launch_tile_flags(Lv, T, c->stream));
        HIPCHK(c, launch_miss_tiles(Lv, T, c->stream));
    }
    std::vector<uint32_t> flags(foff[n_views]);
    HIPCHK(c, copy_sync(c, flags.data(), c->b_flags.p, sizeof(uint32_t) * flags.size(), hipMemcpyDeviceToHost));
    std::vector<BatchChunk> chunks;
    c->b_rays0 = 0;
    for (uint32_t v : order) {
        const uint32_t W = c->b_w[v], H = c->b_h[v], tx = (W + 7) / 8;
        const uint32_t tiles = (uint32_t)(foff[v + 1] - foff[v] - 1);
        const uint32_t *fl = flags.data() + foff[v];
        for (uint32_t t = 0; t < tiles;) {
            if (!fl[t]) {
                const uint32_t x0 = (t % tx) * 8, y0 = (t / tx) * 8;
                c->b_rays0 += (uint64_t)std::min(8u, W - x0) * std::min(8u, H - y0);
                t++;
                continue;
            }
            uint32_t t0 = t, n = 0;
            while (t < tiles && fl[t] && n < kChunkTiles) { t++; n++; }
            chunks.push_back(BatchChunk{ v, t0, n, 0u });
        }
    }
    if (chunks.empty()) chunks.push_back(BatchChunk{ order[0], 0u, 0u, 0u });   // nothing to march
    if (chunks.size() >= (1ull << 32)) return fail(c, PM_ERR_SIZE, "pm_batch_set: too many tiles");

    BatchLaunch &B = c->b_launch;
    B = BatchLaunch{};
    B.base.s = *s;
    B.base.inv_step = 1.0f / s->step;
    B.base.counters = (unsigned long long *)c->d_counters.p;
    HIPCHK(c, c->b_hits.ensure(sizeof(uint4) * npx));
    if (s->learned) {
        set_learned(c, s, B.base);
    }
    for (uint32_t v = 0; v < n_views; v++) {   // the shared fields of every job
        MarchLaunch &J = jobs[v];
        J.hit_rec = (uint4 *)c->b_hits.p + c->b_off[v];
        J.s = B.base.s;
        J.inv_step = B.base.inv_step;
        J.counters = B.base.counters;
        J.net = B.base.net;
        J.mlp_alpha = B.base.mlp_alpha;
        J.mlp_min_skip = B.base.mlp_min_skip;
        J.mlp_max_skip = B.base.mlp_max_skip;
        J.mlp_gate_rho = B.base.mlp_gate_rho;
    }
    HIPCHK(c, c->b_jobs.ensure(sizeof(MarchLaunch) * jobs.size()));
    HIPCHK(c, c->b_chunks.ensure(sizeof(BatchChunk) * chunks.size()));
    HIPCHK(c, c->b_ctr.ensure(256));
    HIPCHK(c, copy_sync(c, c->b_jobs.p, jobs.data(), sizeof(MarchLaunch) * jobs.size(), hipMemcpyHostToDevice));
    HIPCHK(c, copy_sync(c, c->b_chunks.p, chunks.data(), sizeof(BatchChunk) * chunks.size(), hipMemcpyHostToDevice));
    B.jobs = (const MarchLaunch *)c->b_jobs.p;
    B.chunks = (const BatchChunk *)c->b_chunks.p;
    B.nchunks = (uint32_t)chunks.size();
    B.chunk_ctr = (uint32_t *)c->b_ctr.p;
    c->b_blocks = std::min<uint32_t>(1024u, (B.nchunks + 3) / 4 * 4);

    if (use_graph) {
        HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        int rc = batch_enqueue_launch(c);
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamEndCapture(c->stream, &g);
        if (rc) { if (g) hipGraphDestroy(g); return rc; }
        HIPCHK(c, e);
        c->b_graph = g;
        HIPCHK(c, hipGraphInstantiate(&c->b_exec, c->b_graph, nullptr, nullptr, 0));
    }
    c->batch_epoch = c->frame_epoch;
    c->b_ready = true;
    return PM_OK;
}

int pm_batch_run(pm_ctx *c)
{
    if (!c) return PM_ERR_ARG;
    if (c->pending) return fail(c, PM_ERR_STATE, "pm_batch_run: a march is in flight; call pm_wait first");
    if (!c->b_ready) return fail(c, PM_ERR_STATE, "pm_batch_run: no batch (pm_batch_set)");
    if (c->batch_epoch != c->frame_epoch)
        return fail(c, PM_ERR_STATE, "pm_batch_run: frames (or their bound bits) changed after pm_batch_set; set the batch again");
    if (c->b_launch.base.s.learned && !c->mlp_ready) return fail(c, PM_ERR_STATE, "learned step: no network loaded");
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, hipEventRecord(c->ev_launch, c->stream));
    HIPCHK(c, hipEventRecord(c->ev_k0, c->stream));
    if (c->b_exec) {
        HIPCHK(c, hipGraphLaunch(c->b_exec, c->stream));
    } else {
        int rc = batch_enqueue_launch(c);
        if (rc) return rc;
    }
    HIPCHK(c, hipEventRecord(c->ev_k1, c->stream));
    HIPCHK(c, hipMemcpyAsync(c->h_counters, c->d_counters.p, sizeof(unsigned long long) * kCtrWords, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipEventRecord(c->ev_done, c->stream));
    c->user_pos = c->user_nrm = c->user_t = nullptr;
    c->prepass = false;   // the batch counts its answered-once rays in the counters
    c->pending = true;
    c->pending_batch = true;
    c->finished = false;
    return PM_OK;
}

int pm_batch_download(pm_ctx *c, uint32_t view, float *positions, float *normals, float *t)
{
    if (!c) return PM_ERR_ARG;
    if (c->pending) return fail(c, PM_ERR_STATE, "pm_batch_download: march in flight; call pm_wait first");
    if (!c->b_ready || view >= c->b_off.size()) return fail(c, PM_ERR_ARG, "pm_batch_download: no view %u", view);
    HIPCHK(c, hipSetDevice(c->device));
    uint64_t n = (uint64_t)c->b_w[view] * c->b_h[view], o = c->b_off[view];
    const float4 *pos = (const float4 *)c->b_img.p;
    const float4 *nrm = pos + c->b_npx;
    const float *tt = (const float *)(nrm + c->b_npx);
    if (positions) HIPCHK(c, copy_sync(c, positions, pos + o, 16 * n, hipMemcpyDeviceToHost));
    if (normals) HIPCHK(c, copy_sync(c, normals, nrm + o, 16 * n, hipMemcpyDeviceToHost));
    if (t) HIPCHK(c, copy_sync(c, t, tt + o, 4 * n, hipMemcpyDeviceToHost));
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

static int aniso_query_settings(pm_ctx *c, const pm_settings *s)
{
    if (!s) return fail(c, PM_ERR_ARG, "settings is NULL");
    const pm_settings &t = *s;
    if (!(t.k_n > 0.0f && t.k_r > 0.0f && t.k_s > 0.0f) || !std::isfinite(t.k_n) || !std::isfinite(t.k_r) ||
        !std::isfinite(t.k_s))
        return fail(c, PM_ERR_ARG, "anisotropic: k_n, k_r and k_s must be finite and > 0");
    if (t.ext_mult != 0.0f && !(t.ext_mult >= 1.0f && t.ext_mult <= 2.0f))
        return fail(c, PM_ERR_ARG, "anisotropic: ext_mult must be 0 (= 2) or in [1, 2]");
    return PM_OK;
}

int pm_density_aniso(pm_ctx *c, int32_t frame, const pm_settings *s, const float *q, uint64_t nq, float *rho)
{
    Frame *f;
    int rc = run_query(c, frame, nq, &f);
    if (rc) return rc;
    if (!q || !rho) return fail(c, PM_ERR_ARG, "q and rho are required");
    if ((rc = aniso_query_settings(c, s))) return rc;
    const AnisoDev a = aniso_params(f, s);
    HIPCHK(c, c->d_q.ensure(12 * nq));
    HIPCHK(c, c->d_out.ensure(4 * nq));
    HIPCHK(c, hipMemcpyAsync(c->d_q.p, q, 12 * nq, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, launch_density_aniso(f->g, a.f, (const float *)c->d_q.p, (uint32_t)nq, s->compensated,
                                   (float *)c->d_out.p, c->stream));
    HIPCHK(c, hipMemcpyAsync(rho, c->d_out.p, 4 * nq, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return PM_OK;
}

int pm_density_aniso64(pm_ctx *c, int32_t frame, const pm_settings *s, const double *q, uint64_t nq, double *rho)
{
    Frame *f;
    int rc = run_query(c, frame, nq, &f);
    if (rc) return rc;
    if (!q || !rho) return fail(c, PM_ERR_ARG, "q and rho are required");
    if ((rc = aniso_query_settings(c, s))) return rc;
    const AnisoDev a = aniso_params(f, s);
    HIPCHK(c, c->d_q.ensure(24 * nq));
    HIPCHK(c, c->d_out.ensure(8 * nq));
    HIPCHK(c, hipMemcpyAsync(c->d_q.p, q, 24 * nq, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, launch_density_aniso64(f->g, a.d, (const double *)c->d_q.p, (uint32_t)nq, (double *)c->d_out.p,
                                     c->stream));
    HIPCHK(c, hipMemcpyAsync(rho, c->d_out.p, 8 * nq, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return PM_OK;
}

int pm_aniso_matrix(pm_ctx *c, int32_t frame, const pm_settings *s, const float *q, uint64_t nq, float *G, float *det,
                    uint32_t *n_ext)
{
    Frame *f;
    int rc = run_query(c, frame, nq, &f);
    if (rc) return rc;
    if (!q || !G || !det || !n_ext) return fail(c, PM_ERR_ARG, "q, G, det and n_ext are required");
    if ((rc = aniso_query_settings(c, s))) return rc;
    const AnisoDev a = aniso_params(f, s);
    HIPCHK(c, c->d_q.ensure(12 * nq));
    HIPCHK(c, c->d_out.ensure(48 * nq));
    HIPCHK(c, hipMemcpyAsync(c->d_q.p, q, 12 * nq, hipMemcpyHostToDevice, c->stream));
    float *dG = (float *)c->d_out.p, *dd = dG + 9 * nq;
    uint32_t *dn = (uint32_t *)(dd + nq);
    HIPCHK(c, launch_aniso_matrix(f->g, a.f, (const float *)c->d_q.p, (uint32_t)nq, dG, dd, dn, c->stream));
    HIPCHK(c, hipMemcpyAsync(G, dG, 36 * nq, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipMemcpyAsync(det, dd, 4 * nq, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipMemcpyAsync(n_ext, dn, 4 * nq, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return PM_OK;
}

// ------------------------------------------------------------ depth smoothing --

int pm_gauss_kernel(int32_t N, float *K)
{
    if (!K || N < 0 || N > 31) return PM_ERR_ARG;
    // ComputeGaussKernel (GaussRenderPass.cpp:20-48): gauss(r) = powf(E, -0.5 r r),
    // r = sqrtf(i^2 + j^2); the sum counts each off-axis entry 4 times and each
    // on-axis entry twice (the full window), then every entry is divided by it
    const float E = 2.7182818284590452353602874713527f;
    float sum = 0.0f;
    for (int i = 0; i <= N; i++)
        for (int j = 0; j <= N; j++) {
            const float r = sqrtf((float)i * (float)i + (float)j * (float)j);
            const float y = powf(E, -0.5f * r * r);
            K[j + i * (N + 1)] = y;
            if (i == 0 && j == 0) sum += y;
            else if (i != 0 && j != 0) sum += 4.0f * y;
            else sum += 2.0f * y;
        }
    for (int k = 0; k < (N + 1) * (N + 1); k++) K[k] /= sum;
    return PM_OK;
}

// Offset of N's (N + 1)^2 table among the tables for 0..31 stored back to back.
static inline size_t gauss_offset(int N) { return (size_t)N * (N + 1) * (2 * N + 1) / 6; }

// The smoothing tables on the device, uploaded on first use and never written
// again, so launches in flight on any stream may read them.
static int ensure_gauss(pm_ctx *c, const float **d_K)
{
    if (!c->d_gauss.p) {
        std::vector<float> all(gauss_offset(32));
        for (int n = 0; n <= 31; n++) pm_gauss_kernel(n, all.data() + gauss_offset(n));
        HIPCHK(c, c->d_gauss.ensure(sizeof(float) * all.size()));
        hipError_t e = copy_sync(c, c->d_gauss.p, all.data(), sizeof(float) * all.size(), hipMemcpyHostToDevice);
        if (e != hipSuccess) { c->d_gauss.release(); HIPCHK(c, e); }
    }
    *d_K = (const float *)c->d_gauss.p;
    return PM_OK;
}

int pm_depth_smooth(pm_ctx *c, const float *depth, uint32_t W, uint32_t H, int32_t N, float *out)
{
    if (!c) return PM_ERR_ARG;
    if (!depth || !out || N < 0 || N > 31) return fail(c, PM_ERR_ARG, "pm_depth_smooth: need images and 0 <= N <= 31");
    if (W == 0 || H == 0 || (uint64_t)W * H > (1ull << 30)) return fail(c, PM_ERR_SIZE, "bad image size %ux%u", W, H);
    if (c->pending) return fail(c, PM_ERR_STATE, "pm_depth_smooth: a march is in flight");
    HIPCHK(c, hipSetDevice(c->device));
    const uint64_t npx = (uint64_t)W * H;
    const float *d_K = nullptr;
    int rc = ensure_gauss(c, &d_K);
    if (rc) return rc;
    HIPCHK(c, c->d_q.ensure(4 * npx));
    HIPCHK(c, c->d_out.ensure(4 * npx));
    float *d_in = (float *)c->d_q.p;
    HIPCHK(c, hipMemcpyAsync(d_in, depth, 4 * npx, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, launch_depth_smooth(d_in, (float *)c->d_out.p, W, H, N, d_K + gauss_offset(N), c->stream));
    HIPCHK(c, hipMemcpyAsync(out, c->d_out.p, 4 * npx, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return PM_OK;
}

int pm_depth_smooth_device(pm_ctx *c, const float *d_depth, uint32_t W, uint32_t H, int32_t N, float *d_out,
                           void *stream)
{
    if (!c) return PM_ERR_ARG;
    if (N < 0 || N > 31) return fail(c, PM_ERR_ARG, "pm_depth_smooth_device: need 0 <= N <= 31");
    if (W == 0 || H == 0 || (uint64_t)W * H > (1ull << 30)) return fail(c, PM_ERR_SIZE, "bad image size %ux%u", W, H);
    const uint64_t nb = 4ull * W * H;
    HIPCHK(c, hipSetDevice(c->device));
    if (!device_range(c->device, d_depth, nb) || !device_range(c->device, d_out, nb))
        return fail(c, PM_ERR_ARG, "pm_depth_smooth_device: images must be device memory of GPU %d, W*H floats",
                    c->device);
    const char *a = (const char *)d_depth, *b = (const char *)d_out;
    if (a < b + nb && b < a + nb) return fail(c, PM_ERR_ARG, "pm_depth_smooth_device: input and output overlap");
    const float *d_K = nullptr;
    int rc = ensure_gauss(c, &d_K);
    if (rc) return rc;
    HIPCHK(c, launch_depth_smooth(d_depth, d_out, W, H, N, d_K + gauss_offset(N), (hipStream_t)stream));
    return PM_OK;
}

// d_depth != nullptr: into the caller's device memory, enqueued on the context
// stream without a wait; else into `depth` (host) and waited for.
static int sprites_launch(pm_ctx *c, int32_t frame, const float view[16], const float proj[16], const float inv_pv[16],
                          const float cam[3], float radius, uint32_t W, uint32_t H, float *depth, float *d_depth)
{
    if (!c) return PM_ERR_ARG;
    if (!view || !proj || !inv_pv || !cam || !(depth || d_depth)) return fail(c, PM_ERR_ARG, "pm_depth_sprites: null argument");
    if (!(radius > 0.0f) || !std::isfinite(radius)) return fail(c, PM_ERR_ARG, "pm_depth_sprites: radius must be finite and > 0");
    if (W == 0 || H == 0 || (uint64_t)W * H > (1ull << 30)) return fail(c, PM_ERR_SIZE, "bad image size %ux%u", W, H);
    if (c->pending) return fail(c, PM_ERR_STATE, "pm_depth_sprites: a march is in flight");
    Frame *f = get_frame(c, frame);
    if (!f) return fail(c, PM_ERR_FRAME, "frame %d not set", frame);
    SpriteLaunch S{};
    for (int k = 0; k < 16; k++) {
        S.view[k] = view[k]; S.proj[k] = proj[k]; S.inv_pv[k] = inv_pv[k];
        if (!std::isfinite(view[k]) || !std::isfinite(proj[k]) || !std::isfinite(inv_pv[k]))
            return fail(c, PM_ERR_ARG, "pm_depth_sprites: non-finite matrix entry");
    }
    for (int k = 0; k < 3; k++) S.cam[k] = cam[k];
    S.radius = radius;
    S.W = W;
    S.H = H;
    HIPCHK(c, hipSetDevice(c->device));
    const uint64_t npx = (uint64_t)W * H;
    // room to list every particle with a large footprint (a wave then shares its
    // pixels); a sprite pass still in flight may use the old list, so it ends
    // before the list grows
    const size_t big_bytes = sizeof(uint32_t) * (f->info.n + 1);
    if (c->d_sbig.cap < big_bytes && c->sprites_pending) HIPCHK(c, hipStreamSynchronize(c->stream));
    HIPCHK(c, c->d_sbig.ensure(big_bytes));
    if (d_depth) {
        if (!device_range(c->device, d_depth, 4 * npx))
            return fail(c, PM_ERR_ARG, "pm_depth_sprites_device: the image must be device memory of GPU %d, W*H floats",
                        c->device);
        HIPCHK(c, launch_depth_sprites(S, (const float4 *)f->sorted.p, (uint32_t)f->info.n, (uint32_t *)d_depth,
                                       (uint32_t *)c->d_sbig.p, (uint32_t)f->info.n, c->stream));
        HIPCHK(c, hipEventRecord(c->ev_sprites, c->stream));
        c->sprites_pending = true;
        c->sprites_frame = frame;
        return PM_OK;
    }
    HIPCHK(c, c->d_out.ensure(4 * npx));
    HIPCHK(c, launch_depth_sprites(S, (const float4 *)f->sorted.p, (uint32_t)f->info.n, (uint32_t *)c->d_out.p,
                                   (uint32_t *)c->d_sbig.p, (uint32_t)f->info.n, c->stream));
    HIPCHK(c, hipMemcpyAsync(depth, c->d_out.p, 4 * npx, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return PM_OK;
}

int pm_depth_sprites(pm_ctx *c, int32_t frame, const float view[16], const float proj[16], const float inv_pv[16],
                     const float cam[3], float radius, uint32_t W, uint32_t H, float *depth)
{
    if (!depth) return c ? fail(c, PM_ERR_ARG, "pm_depth_sprites: null argument") : PM_ERR_ARG;
    return sprites_launch(c, frame, view, proj, inv_pv, cam, radius, W, H, depth, nullptr);
}

int pm_depth_sprites_device(pm_ctx *c, int32_t frame, const float view[16], const float proj[16],
                            const float inv_pv[16], const float cam[3], float radius, uint32_t W, uint32_t H,
                            float *d_depth)
{
    if (!d_depth) return c ? fail(c, PM_ERR_ARG, "pm_depth_sprites_device: null image") : PM_ERR_ARG;
    return sprites_launch(c, frame, view, proj, inv_pv, cam, radius, W, H, nullptr, d_depth);
}

// ---------------------------------------------------------------- learned --

int pm_mlp_set(pm_ctx *c, const float *params)
{
    if (!c || !params) return PM_ERR_ARG;
    if (c->pending) return fail(c, PM_ERR_STATE, "a march is in flight");
    HIPCHK(c, hipSetDevice(c->device));
    const size_t nb = sizeof(float) * PM_MLP_NPARAM;
    HIPCHK(c, c->mlp_p.ensure(nb));
    HIPCHK(c, c->mlp_m.ensure(nb));
    HIPCHK(c, c->mlp_v.ensure(nb));
    HIPCHK(c, c->mlp_W1.ensure(2 * PM_MLP_H * PM_MLP_K1));
    HIPCHK(c, c->mlp_W2.ensure(2 * PM_MLP_H * PM_MLP_H));
    HIPCHK(c, c->mlp_W2T.ensure(2 * PM_MLP_H * PM_MLP_H));
    std::vector<float> p(params, params + PM_MLP_NPARAM);
    for (int i = 0; i < PM_MLP_H; i++)
        for (int j = PM_MLP_IN; j < PM_MLP_K1; j++) p[PM_MLP_OFF_W1 + i * PM_MLP_K1 + j] = 0.0f;
    for (float v : p)
        if (!std::isfinite(v)) return fail(c, PM_ERR_ARG, "pm_mlp_set: non-finite parameter");
    HIPCHK(c, copy_sync(c, c->mlp_p.p, p.data(), nb, hipMemcpyHostToDevice));
    HIPCHK(c, zero_sync(c, c->mlp_m.p, nb));
    HIPCHK(c, zero_sync(c, c->mlp_v.p, nb));
    HIPCHK(c, launch_mlp_sync((const float *)c->mlp_p.p, (uint16_t *)c->mlp_W1.p, (uint16_t *)c->mlp_W2.p,
                              (uint16_t *)c->mlp_W2T.p, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->mlp_ready = true;
    c->mlp_step = 0;
    return PM_OK;
}

int pm_mlp_get(pm_ctx *c, float *params)
{
    if (!c || !params) return PM_ERR_ARG;
    if (!c->mlp_ready) return fail(c, PM_ERR_STATE, "no network loaded");
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, copy_sync(c, params, c->mlp_p.p, sizeof(float) * PM_MLP_NPARAM, hipMemcpyDeviceToHost));
    return PM_OK;
}

int pm_mlp_forward(pm_ctx *c, const float *x, uint64_t n, float *y)
{
    if (!c || !x || !y || n == 0 || n >= (1ull << 31)) return PM_ERR_ARG;
    if (!c->mlp_ready) return fail(c, PM_ERR_STATE, "no network loaded");
    if (c->pending) return fail(c, PM_ERR_STATE, "a march is in flight");
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, c->mlp_x.ensure(sizeof(float) * PM_MLP_IN * n));
    HIPCHK(c, c->mlp_out.ensure(sizeof(float) * n));
    HIPCHK(c, copy_sync(c, c->mlp_x.p, x, sizeof(float) * PM_MLP_IN * n, hipMemcpyHostToDevice));
    HIPCHK(c, launch_mlp_forward(mlp_dev(c), (const float *)c->mlp_x.p, (uint32_t)n, (float *)c->mlp_out.p, c->stream));
    HIPCHK(c, hipMemcpyAsync(y, c->mlp_out.p, sizeof(float) * n, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return PM_OK;
}

// One forward/backward over device rows `feat`/`label` (ids from idx, or 0..n-1),
// partial gradients reduced in wave order; optionally one Adam step.
static int mlp_pass(pm_ctx *c, const float *feat, const float *label, const uint32_t *idx, uint32_t n, float w_over,
                    bool apply, float lr, float *loss_dev_out, float *grad_dev_out)
{
    uint32_t nw = n / 64;
    HIPCHK(c, c->mlp_partial.ensure(sizeof(float) * (size_t)nw * (PM_MLP_NPARAM + 1)));
    TrainArgs ta{};
    ta.net = mlp_dev(c);
    ta.feat = feat;
    ta.label = label;
    ta.idx = idx;
    ta.inv_batch = 1.0f / (float)n;
    ta.w_over = w_over;
    ta.partial = (float *)c->mlp_partial.p;
    HIPCHK(c, launch_mlp_train(ta, nw, c->stream));
    AdamArgs aa{};
    aa.partial = (const float *)c->mlp_partial.p;
    aa.nparts = nw;
    aa.p = (float *)c->mlp_p.p;
    aa.m = (float *)c->mlp_m.p;
    aa.v = (float *)c->mlp_v.p;
    aa.apply = apply ? 1 : 0;
    aa.grad_out = grad_dev_out;
    aa.loss_out = loss_dev_out;
    if (apply) {
        c->mlp_step++;
        aa.lr = lr;
        aa.b1 = 0.9f;
        aa.b2 = 0.999f;
        aa.eps = 1e-8f;
        aa.bc1 = 1.0f - std::pow(0.9f, (float)c->mlp_step);
        aa.bc2 = 1.0f - std::pow(0.999f, (float)c->mlp_step);
    }
    HIPCHK(c, launch_mlp_adam(aa, c->stream));
    if (apply)
        HIPCHK(c, launch_mlp_sync((const float *)c->mlp_p.p, (uint16_t *)c->mlp_W1.p, (uint16_t *)c->mlp_W2.p,
                                  (uint16_t *)c->mlp_W2T.p, c->stream));
    return PM_OK;
}

int pm_mlp_grad(pm_ctx *c, const float *x, const float *target, uint64_t n, float w_over, float *loss, float *grad)
{
    if (!c || !x || !target || !loss || !grad) return PM_ERR_ARG;
    if (n == 0 || n % 64 != 0 || n > (1u << 24)) return fail(c, PM_ERR_ARG, "pm_mlp_grad: n must be a positive multiple of 64");
    if (!c->mlp_ready) return fail(c, PM_ERR_STATE, "no network loaded");
    if (c->pending) return fail(c, PM_ERR_STATE, "a march is in flight");
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, c->mlp_x.ensure(sizeof(float) * (PM_MLP_IN + 1) * n));
    HIPCHK(c, c->mlp_out.ensure(sizeof(float) * (PM_MLP_NPARAM + 1)));
    float *dx = (float *)c->mlp_x.p, *dt = dx + PM_MLP_IN * n;
    HIPCHK(c, copy_sync(c, dx, x, sizeof(float) * PM_MLP_IN * n, hipMemcpyHostToDevice));
    HIPCHK(c, copy_sync(c, dt, target, sizeof(float) * n, hipMemcpyHostToDevice));
    float *dg = (float *)c->mlp_out.p;
    int rc = mlp_pass(c, dx, dt, nullptr, (uint32_t)n, w_over, false, 0.f, dg + PM_MLP_NPARAM, dg);
    if (rc) return rc;
    HIPCHK(c, hipMemcpyAsync(grad, dg, sizeof(float) * PM_MLP_NPARAM, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipMemcpyAsync(loss, dg + PM_MLP_NPARAM, sizeof(float), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return PM_OK;
}

int pm_dataset_clear(pm_ctx *c, uint64_t capacity)
{
    if (!c || capacity == 0 || capacity >= (1ull << 31)) return PM_ERR_ARG;
    if (c->pending) return fail(c, PM_ERR_STATE, "a march is in flight");
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, c->ds_feat.ensure(sizeof(float) * PM_MLP_IN * capacity));
    HIPCHK(c, c->ds_label.ensure(sizeof(float) * capacity));
    HIPCHK(c, c->ds_hit.ensure(capacity));
    HIPCHK(c, c->ds_meta.ensure(8 * capacity));
    HIPCHK(c, c->ds_count.ensure(sizeof(uint32_t)));
    HIPCHK(c, zero_sync(c, c->ds_count.p, sizeof(uint32_t)));
    c->ds_cap = capacity;
    return PM_OK;
}

static int ds_size(pm_ctx *c, uint64_t *n)
{
    uint32_t v = 0;
    HIPCHK(c, copy_sync(c, &v, c->ds_count.p, sizeof(v), hipMemcpyDeviceToHost));
    if (v > c->ds_cap) {                     // the logger overshoots its cap by the rows it dropped
        v = (uint32_t)c->ds_cap;
    }
    *n = v;
    return PM_OK;
}

int pm_trace_append(pm_ctx *c, int32_t frame, const pm_settings *s, const float inv_pv[16], const float cam[3], uint32_t W,
                    uint32_t H, uint32_t ray_stride, uint32_t k_stride, uint32_t seed, uint64_t max_new_rows,
                    uint64_t *rows_total)
{
    if (!c || !s || !inv_pv || !cam || ray_stride == 0 || k_stride == 0) return PM_ERR_ARG;
    if (c->ds_cap == 0) return fail(c, PM_ERR_STATE, "pm_trace_append: call pm_dataset_clear first");
    if (s->mode != PM_MODE_CONFIG || s->learned || s->anisotropic)
        return fail(c, PM_ERR_ARG, "pm_trace_append: traces come from the plain configured march");
    int rc = pm_march_async(c, frame, s, inv_pv, cam, W, H, nullptr, nullptr, nullptr, nullptr);
    if (rc) return rc;
    rc = pm_wait(c);
    if (rc) return rc;
    uint64_t before = 0;
    rc = ds_size(c, &before);
    if (rc) return rc;
    uint64_t cap = c->ds_cap;
    if (max_new_rows > 0 && before + max_new_rows < cap) cap = before + max_new_rows;
    Frame *f = get_frame(c, frame);
    TraceArgs a{};
    a.g = f->g;
    a.cg = f->cg;
    memcpy(a.M, inv_pv, sizeof(a.M));
    memcpy(a.cam, cam, sizeof(a.cam));
    a.W = W; a.H = H;
    a.step = s->step;
    a.iso = s->iso;
    a.t_hit = (const float *)c->d_t.p;
    a.ray_stride = ray_stride;
    a.k_stride = k_stride;
    a.seed = seed;
    a.count = (uint32_t *)c->ds_count.p;
    a.cap = (uint32_t)cap;
    a.feat = (float *)c->ds_feat.p;
    a.label = (float *)c->ds_label.p;
    a.is_hit = (uint8_t *)c->ds_hit.p;
    a.meta = (uint32_t *)c->ds_meta.p;
    HIPCHK(c, launch_trace(a, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    // rows past the cap were counted but not written: bring the counter back to the cap
    uint32_t v = 0;
    HIPCHK(c, copy_sync(c, &v, c->ds_count.p, sizeof(v), hipMemcpyDeviceToHost));
    if (v > cap) {
        v = (uint32_t)cap;
        HIPCHK(c, copy_sync(c, c->ds_count.p, &v, sizeof(v), hipMemcpyHostToDevice));
    }
    if (rows_total) *rows_total = v;
    return PM_OK;
}

int pm_dataset_get(pm_ctx *c, uint64_t offset, uint64_t n, float *feat, float *label, uint8_t *is_hit, uint32_t *meta)
{
    if (!c) return PM_ERR_ARG;
    uint64_t tot = 0;
    int rc = ds_size(c, &tot);
    if (rc) return rc;
    if (offset + n > tot) return fail(c, PM_ERR_ARG, "pm_dataset_get: range beyond the %llu rows", (unsigned long long)tot);
    if (feat) HIPCHK(c, copy_sync(c, feat, (float *)c->ds_feat.p + PM_MLP_IN * offset, sizeof(float) * PM_MLP_IN * n, hipMemcpyDeviceToHost));
    if (label) HIPCHK(c, copy_sync(c, label, (float *)c->ds_label.p + offset, sizeof(float) * n, hipMemcpyDeviceToHost));
    if (is_hit) HIPCHK(c, copy_sync(c, is_hit, (uint8_t *)c->ds_hit.p + offset, n, hipMemcpyDeviceToHost));
    if (meta) HIPCHK(c, copy_sync(c, meta, (uint32_t *)c->ds_meta.p + 2 * offset, 8 * n, hipMemcpyDeviceToHost));
    return PM_OK;
}

int pm_mlp_train(pm_ctx *c, int epochs, uint32_t batch, float lr, float w_over, uint64_t seed, float *loss_per_epoch)
{
    if (!c || epochs <= 0 || batch == 0 || batch % 64 != 0 || !(lr > 0.0f)) return PM_ERR_ARG;
    if (!c->mlp_ready) return fail(c, PM_ERR_STATE, "no network loaded (pm_mlp_set)");
    if (c->pending) return fail(c, PM_ERR_STATE, "a march is in flight");
    HIPCHK(c, hipSetDevice(c->device));
    uint64_t n = 0;
    int rc = ds_size(c, &n);
    if (rc) return rc;
    if (n < batch) return fail(c, PM_ERR_STATE, "pm_mlp_train: %llu rows, fewer than one batch", (unsigned long long)n);
    uint64_t nbatches = n / batch;
    HIPCHK(c, c->mlp_idx.ensure(sizeof(uint32_t) * nbatches * batch));
    HIPCHK(c, c->mlp_out.ensure(sizeof(float) * (nbatches + PM_MLP_NPARAM + 1)));
    std::vector<uint32_t> perm(n);
    std::vector<float> losses(nbatches);
    std::mt19937_64 rng(seed);
    for (int e = 0; e < epochs; e++) {
        for (uint64_t i = 0; i < n; i++) perm[i] = (uint32_t)i;
        std::shuffle(perm.begin(), perm.end(), rng);
        HIPCHK(c, copy_sync(c, c->mlp_idx.p, perm.data(), sizeof(uint32_t) * nbatches * batch, hipMemcpyHostToDevice));
        for (uint64_t b = 0; b < nbatches; b++) {
            rc = mlp_pass(c, (const float *)c->ds_feat.p, (const float *)c->ds_label.p,
                          (const uint32_t *)c->mlp_idx.p + b * batch, batch, w_over, true, lr, (float *)c->mlp_out.p + b,
                          nullptr);
            if (rc) return rc;
        }
        HIPCHK(c, copy_sync(c, losses.data(), c->mlp_out.p, sizeof(float) * nbatches, hipMemcpyDeviceToHost));
        double sum = 0.0;
        for (float v : losses) sum += v;
        if (loss_per_epoch) loss_per_epoch[e] = (float)(sum / (double)nbatches);
    }
    return PM_OK;
}

int pm_get_skip_grids(pm_ctx *c, int32_t frame, uint8_t *empty27, int32_t dg_dims[3], uint8_t *flag_box,
                      uint8_t *flag_cons, uint32_t *count_box)
{
    if (!c) return PM_ERR_ARG;
    Frame *f = get_frame(c, frame);
    if (!f) return fail(c, PM_ERR_FRAME, "frame %d not set", frame);
    HIPCHK(c, hipSetDevice(c->device));
    uint64_t dgn = (uint64_t)f->dg_dims[0] * f->dg_dims[1] * f->dg_dims[2];
    if (dg_dims) for (int a = 0; a < 3; a++) dg_dims[a] = f->dg_dims[a];
    if (flag_box || flag_cons || count_box) {
        int rc = ensure_densitygrid(c, f);
        if (rc) return rc;
    }
    if (empty27) HIPCHK(c, copy_sync(c, empty27, f->empty27.p, f->info.ncells, hipMemcpyDeviceToHost));
    if (flag_box) HIPCHK(c, copy_sync(c, flag_box, f->dg_box.p, dgn, hipMemcpyDeviceToHost));
    if (flag_cons) HIPCHK(c, copy_sync(c, flag_cons, f->dg_cons.p, dgn, hipMemcpyDeviceToHost));
    if (count_box) HIPCHK(c, copy_sync(c, count_box, f->dg_count.p, 4 * dgn, hipMemcpyDeviceToHost));
    return PM_OK;
}

int pm_stream_copy_gbps(pm_ctx *c, uint64_t bytes, int iters, float *gbps)
{
    if (!c || !gbps || bytes < 16 || iters <= 0) return PM_ERR_ARG;
    if (c->pending) return fail(c, PM_ERR_STATE, "a march is in flight");
    HIPCHK(c, hipSetDevice(c->device));
    bytes &= ~(uint64_t)15;
    void *a = nullptr, *b = nullptr;
    HIPCHK(c, hipMalloc(&a, bytes));
    if (hipMalloc(&b, bytes) != hipSuccess) { hipFree(a); return fail(c, PM_ERR_HIP, "hipMalloc failed"); }
    hipEvent_t e0, e1;
    hipEventCreate(&e0);
    hipEventCreate(&e1);
    hipMemsetAsync(a, 0, bytes, c->stream);
    launch_stream_copy(a, b, bytes, c->stream);   // warm-up
    hipEventRecord(e0, c->stream);
    for (int i = 0; i < iters; i++) launch_stream_copy(a, b, bytes, c->stream);
    hipEventRecord(e1, c->stream);
    hipError_t e = hipEventSynchronize(e1);
    float ms = 0.f;
    hipEventElapsedTime(&ms, e0, e1);
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    hipFree(a);
    hipFree(b);
    if (e != hipSuccess) return fail(c, PM_ERR_HIP, "copy failed: %s", hipGetErrorString(e));
    *gbps = (float)(2.0 * (double)bytes * iters / (ms * 1e-3) / 1e9);
    return PM_OK;
}

}  // extern "C"
