// pm_mlp.hip -- training and batch inference of the step-proposal MLP, the
// coarse occupancy grid its features read, and the trace logger that turns
// plain-march results into (features, distance-to-crossing) pairs.
//
// Training step for a batch of B samples (B a multiple of 64):
//   k_mlp_train : one wave per 64 samples -- forward, loss, backward with BF16
//                 MFMA GEMMs (FP32 accumulate) -> per-wave partial gradients
//   k_mlp_adam  : sums the partials in wave order (deterministic) and applies
//                 Adam to the FP32 master weights, then refreshes the BF16 copies
#include "pm_features.h"
#include "pm_internal.h"
#include "pm_mlp.h"

namespace pm {

// ------------------------------------------------------------ coarse grid --

__global__ __launch_bounds__(256) void k_coarse_counts(GridDev g, CoarseDev cg, uint32_t *__restrict__ out)
{
    uint32_t c = blockIdx.x * 256 + threadIdx.x;
    if (c >= (uint32_t)cg.nx * cg.ny * cg.nz) return;
    int32_t X = c % cg.nx, Y = (c / cg.nx) % cg.ny, Z = c / (cg.nx * cg.ny);
    uint32_t sum = 0;
    for (int z = 4 * Z; z < 4 * Z + 4 && z < g.nz; z++)
        for (int y = 4 * Y; y < 4 * Y + 4 && y < g.ny; y++) {
            int x0 = 4 * X, x1 = min(4 * X + 4, g.nx);
            uint32_t row = (uint32_t)g.nx * ((uint32_t)y + (uint32_t)g.ny * (uint32_t)z);
            sum += g.cell_start[row + x1] - g.cell_start[row + x0];
        }
    out[c] = sum;
}

hipError_t launch_coarse_counts(const GridDev &g, const CoarseDev &cg, uint32_t *out, hipStream_t st)
{
    uint32_t n = (uint32_t)cg.nx * cg.ny * cg.nz;
    hipLaunchKernelGGL(k_coarse_counts, dim3((n + 255) / 256), dim3(256), 0, st, g, cg, out);
    return hipGetLastError();
}

// ------------------------------------------------------------ trace logger --

__device__ __forceinline__ uint32_t hash32(uint32_t x)
{
    x ^= x >> 16; x *= 0x7feb352du; x ^= x >> 15; x *= 0x846ca68bu; x ^= x >> 16;
    return x;
}

// One thread per selected ray.  Rays are selected by hash(pixel ^ seed) % ray_stride == 0;
// samples k = k0, k0 + k_stride, ... (k0 = hash % k_stride) before the crossing.
__global__ __launch_bounds__(256) void k_trace(TraceArgs a)
{
    uint32_t idx = blockIdx.x * 256 + threadIdx.x;
    if (idx >= a.W * a.H) return;
    uint32_t hsh = hash32(idx ^ a.seed);
    if (hsh % a.ray_stride != 0) return;
    const GridDev &g = a.g;
    uint32_t px = idx % a.W, py = idx / a.W;
    float cx = (float)px * (2.0f / (float)a.W) - 1.0f, cy = (float)py * (2.0f / (float)a.H) - 1.0f;
    float w[4];
    for (int r = 0; r < 4; r++) w[r] = (a.M[r] * cx + a.M[4 + r] * cy) + (a.M[8 + r] * 1.0f + a.M[12 + r] * 1.0f);
    float dun[3] = { w[0] / w[3] - a.cam[0], w[1] / w[3] - a.cam[1], w[2] / w[3] - a.cam[2] };
    float inv = 1.0f / sqrtf((dun[0] * dun[0] + dun[1] * dun[1]) + dun[2] * dun[2]);
    float d[3] = { dun[0] * inv, dun[1] * inv, dun[2] * inv };
    float tn = 0.0f, tf = INFINITY;
    for (int ax = 0; ax < 3; ax++) {
        if (d[ax] != 0.0f) {
            float iv = 1.0f / d[ax];
            float t0 = (g.bmin[ax] - a.cam[ax]) * iv, t1 = (g.bmax[ax] - a.cam[ax]) * iv;
            tn = fmaxf(tn, fminf(t0, t1));
            tf = fminf(tf, fmaxf(t0, t1));
        } else if (a.cam[ax] < g.bmin[ax] || a.cam[ax] > g.bmax[ax]) {
            return;
        }
    }
    if (tn > tf) return;
    float thit = a.t_hit[idx];
    float tend = thit >= 0.0f ? thit : tf;
    uint32_t k = (hsh >> 8) % a.k_stride;
    for (;; k += a.k_stride) {
        float t = tn + (float)k * a.step;
        if (t >= tend) break;
        float q[3] = { a.cam[0] + t * d[0], a.cam[1] + t * d[1], a.cam[2] + t * d[2] };
        // the features the learned march computes at a sample below iso: its
        // (compensated) density, grad rho over the fine stencil, the 27-cell count
        uint32_t cand = 0;
        const float rho = pm_density32(g, q[0], q[1], q[2], true, cand);
        if (rho >= a.iso) break;                 // at or past the crossing
        uint32_t row = atomicAdd(a.count, 1u);
        if (row >= a.cap) break;
        float gx, gy, gz;
        grad_feat(g, q[0], q[1], q[2], gx, gy, gz, cand);
        float f[PM_MLP_IN];
        float gmag = sqrtf((gx * gx + gy * gy) + gz * gz);
        mlp_features(g, a.cg, a.cam, d, t, tn, tf, rho, gmag, (float)pm_cand27(g, q[0], q[1], q[2]), a.iso, f);
        for (int j = 0; j < PM_MLP_IN; j++) a.feat[(uint64_t)row * PM_MLP_IN + j] = f[j];
        float dist = (tend - t) * g.inv_h;
        a.label[row] = log2f(1.0f + fminf(dist, 255.0f));
        a.is_hit[row] = thit >= 0.0f ? 1 : 0;
        a.meta[2 * (uint64_t)row] = idx;
        a.meta[2 * (uint64_t)row + 1] = k;
    }
}

hipError_t launch_trace(const TraceArgs &a, hipStream_t st)
{
    uint32_t n = a.W * a.H;
    hipLaunchKernelGGL(k_trace, dim3((n + 255) / 256), dim3(256), 0, st, a);
    return hipGetLastError();
}

// --------------------------------------------------------------- training --

struct TrainLds {
    uint16_t X[64 * PM_MLP_K1];
    uint16_t XT[PM_MLP_K1 * 64];
    uint16_t H1[64 * 64], H1T[64 * 64];
    uint16_t G2[64 * 64], G2T[64 * 64];
    uint16_t G1T[64 * 64];
    float Y[64], DY[64];
};

__global__ __launch_bounds__(64) void k_mlp_train(TrainArgs a)
{
    __shared__ TrainLds S;
    const int l = threadIdx.x, col = l & 15, kb = l >> 4;
    const MlpDev &net = a.net;
    const uint32_t sample = a.idx ? a.idx[blockIdx.x * 64 + l] : blockIdx.x * 64 + l;
    float *part = a.partial + (uint64_t)blockIdx.x * (PM_MLP_NPARAM + 1);

    // inputs
    for (int j = 0; j < PM_MLP_K1; j++) {
        uint16_t v = j < PM_MLP_IN ? f2bf(a.feat[(uint64_t)sample * PM_MLP_IN + j]) : (uint16_t)0;
        S.X[l * PM_MLP_K1 + j] = v;
        S.XT[j * 64 + l] = v;
    }
    wave_lds_sync();

    // layer 1
    pm_f32x4 z1[4][4];
    wave_gemm_abt<4, 1>(S.X, PM_MLP_K1, net.W1, PM_MLP_K1, z1);
#pragma unroll
    for (int mi = 0; mi < 4; mi++)
#pragma unroll
        for (int ni = 0; ni < 4; ni++) {
            int u = 16 * ni + col;
            float b = net.p[PM_MLP_OFF_B1 + u];
#pragma unroll
            for (int r = 0; r < 4; r++) {
                int s = 16 * mi + 4 * kb + r;
                z1[mi][ni][r] += b;
                uint16_t hb = f2bf(mlp_act(z1[mi][ni][r]));
                S.H1[s * 64 + u] = hb;
                S.H1T[u * 64 + s] = hb;
            }
        }
    wave_lds_sync();

    // layer 2 and output
    pm_f32x4 z2[4][4];
    wave_gemm_abt<4, 2>(S.H1, 64, net.W2, 64, z2);
    float part_y[4][4];
#pragma unroll
    for (int mi = 0; mi < 4; mi++)
#pragma unroll
        for (int r = 0; r < 4; r++) part_y[mi][r] = 0.f;
#pragma unroll
    for (int ni = 0; ni < 4; ni++) {
        int u = 16 * ni + col;
        float b = net.p[PM_MLP_OFF_B2 + u], w3 = net.p[PM_MLP_OFF_W3 + u];
#pragma unroll
        for (int mi = 0; mi < 4; mi++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                z2[mi][ni][r] += b;
                part_y[mi][r] += w3 * mlp_act(z2[mi][ni][r]);
            }
    }
#pragma unroll
    for (int mi = 0; mi < 4; mi++)
#pragma unroll
        for (int r = 0; r < 4; r++) {
            float v = part_y[mi][r];
            v += __shfl_xor(v, 1, 64);
            v += __shfl_xor(v, 2, 64);
            v += __shfl_xor(v, 4, 64);
            v += __shfl_xor(v, 8, 64);
            part_y[mi][r] = v;
        }
    if (col == 0)
#pragma unroll
        for (int mi = 0; mi < 4; mi++)
#pragma unroll
            for (int r = 0; r < 4; r++) S.Y[16 * mi + 4 * kb + r] = part_y[mi][r] + net.p[PM_MLP_OFF_B3];
    wave_lds_sync();

    // loss: w * (y - t)^2, w = w_over when the prediction is too long
    float y = S.Y[l], t = a.label[sample];
    float diff = y - t;
    float wt = diff > 0.0f ? a.w_over : 1.0f;
    float loss = wt * diff * diff * a.inv_batch;
    float dy = 2.0f * wt * diff * a.inv_batch;
    S.DY[l] = dy;
    wave_lds_sync();

    // output layer and layer-2 pre-activation gradients
    float dw3[4] = { 0.f, 0.f, 0.f, 0.f }, db2[4] = { 0.f, 0.f, 0.f, 0.f };
#pragma unroll
    for (int ni = 0; ni < 4; ni++) {
        int u = 16 * ni + col;
        float w3 = net.p[PM_MLP_OFF_W3 + u];
#pragma unroll
        for (int mi = 0; mi < 4; mi++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                int s = 16 * mi + 4 * kb + r;
                float z = z2[mi][ni][r];
                float dys = S.DY[s];
                dw3[ni] += dys * mlp_act(z);
                float dz = dys * w3 * mlp_act_grad(z);
                db2[ni] += dz;
                uint16_t gb = f2bf(dz);
                S.G2[s * 64 + u] = gb;
                S.G2T[u * 64 + s] = gb;
            }
    }
#pragma unroll
    for (int ni = 0; ni < 4; ni++) {
        dw3[ni] += __shfl_xor(dw3[ni], 16, 64);
        dw3[ni] += __shfl_xor(dw3[ni], 32, 64);
        db2[ni] += __shfl_xor(db2[ni], 16, 64);
        db2[ni] += __shfl_xor(db2[ni], 32, 64);
    }
    if (kb == 0)
#pragma unroll
        for (int ni = 0; ni < 4; ni++) {
            part[PM_MLP_OFF_W3 + 16 * ni + col] = dw3[ni];
            part[PM_MLP_OFF_B2 + 16 * ni + col] = db2[ni];
        }
    wave_lds_sync();

    // dW2[o][i] = sum_s dZ2[s][o] A1[s][i]
    pm_f32x4 acc[4][4];
    wave_gemm_abt<4, 2>(S.G2T, 64, S.H1T, 64, acc);
#pragma unroll
    for (int mi = 0; mi < 4; mi++)
#pragma unroll
        for (int ni = 0; ni < 4; ni++)
#pragma unroll
            for (int r = 0; r < 4; r++) part[PM_MLP_OFF_W2 + (16 * mi + 4 * kb + r) * 64 + 16 * ni + col] = acc[mi][ni][r];

    // dA1[s][i] = sum_o dZ2[s][o] W2[o][i]; dZ1 = dA1 * act'(Z1)
    wave_gemm_abt<4, 2>(S.G2, 64, net.W2T, 64, acc);
    float db1[4] = { 0.f, 0.f, 0.f, 0.f };
#pragma unroll
    for (int mi = 0; mi < 4; mi++)
#pragma unroll
        for (int ni = 0; ni < 4; ni++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                int s = 16 * mi + 4 * kb + r, u = 16 * ni + col;
                float dz = acc[mi][ni][r] * mlp_act_grad(z1[mi][ni][r]);
                db1[ni] += dz;
                S.G1T[u * 64 + s] = f2bf(dz);
            }
#pragma unroll
    for (int ni = 0; ni < 4; ni++) {
        db1[ni] += __shfl_xor(db1[ni], 16, 64);
        db1[ni] += __shfl_xor(db1[ni], 32, 64);
    }
    if (kb == 0)
#pragma unroll
        for (int ni = 0; ni < 4; ni++) part[PM_MLP_OFF_B1 + 16 * ni + col] = db1[ni];
    wave_lds_sync();

    // dW1[i][f] = sum_s dZ1[s][i] X[s][f]
    pm_f32x4 acc1[4][2];
    wave_gemm_abt<2, 2>(S.G1T, 64, S.XT, 64, acc1);
#pragma unroll
    for (int mi = 0; mi < 4; mi++)
#pragma unroll
        for (int ni = 0; ni < 2; ni++)
#pragma unroll
            for (int r = 0; r < 4; r++) part[PM_MLP_OFF_W1 + (16 * mi + 4 * kb + r) * PM_MLP_K1 + 16 * ni + col] = acc1[mi][ni][r];

    // db3 and the loss: wave sums
    float v0 = dy, v1 = loss;
    for (int off = 32; off > 0; off >>= 1) {
        v0 += __shfl_xor(v0, off, 64);
        v1 += __shfl_xor(v1, off, 64);
    }
    if (l == 0) {
        part[PM_MLP_OFF_B3] = v0;
        part[PM_MLP_NPARAM] = v1;
    }
}

// Sum of the per-wave partials (wave order) for every parameter and the loss;
// with apply != 0, one Adam step on the FP32 master and the BF16 copies.
__global__ __launch_bounds__(256) void k_mlp_adam(AdamArgs a)
{
    uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i > PM_MLP_NPARAM) return;
    float gsum = 0.0f;
    for (uint32_t w = 0; w < a.nparts; w++) gsum += a.partial[(uint64_t)w * (PM_MLP_NPARAM + 1) + i];
    if (i == PM_MLP_NPARAM) {
        a.loss_out[0] = gsum;
        return;
    }
    if (a.grad_out) a.grad_out[i] = gsum;
    if (!a.apply) return;
    float m = a.b1 * a.m[i] + (1.0f - a.b1) * gsum;
    float v = a.b2 * a.v[i] + (1.0f - a.b2) * gsum * gsum;
    a.m[i] = m;
    a.v[i] = v;
    float mh = m / a.bc1, vh = v / a.bc2;
    float p = a.p[i] - a.lr * mh / (sqrtf(vh) + a.eps);
    if (i < PM_MLP_OFF_B1 && (i % PM_MLP_K1) >= PM_MLP_IN) p = 0.0f;   // padding inputs stay zero
    a.p[i] = p;
}

__global__ __launch_bounds__(256) void k_mlp_sync_bf16(const float *__restrict__ p, uint16_t *__restrict__ W1,
                                                       uint16_t *__restrict__ W2, uint16_t *__restrict__ W2T)
{
    uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i < PM_MLP_H * PM_MLP_K1) W1[i] = f2bf(p[PM_MLP_OFF_W1 + i]);
    if (i < PM_MLP_H * PM_MLP_H) {
        uint16_t b = f2bf(p[PM_MLP_OFF_W2 + i]);
        uint32_t o = i / PM_MLP_H, in = i % PM_MLP_H;
        W2[i] = b;
        W2T[in * PM_MLP_H + o] = b;
    }
}

hipError_t launch_mlp_train(const TrainArgs &a, uint32_t nwaves, hipStream_t st)
{
    hipLaunchKernelGGL(k_mlp_train, dim3(nwaves), dim3(64), 0, st, a);
    return hipGetLastError();
}

hipError_t launch_mlp_adam(const AdamArgs &a, hipStream_t st)
{
    hipLaunchKernelGGL(k_mlp_adam, dim3((PM_MLP_NPARAM + 1 + 255) / 256), dim3(256), 0, st, a);
    return hipGetLastError();
}

hipError_t launch_mlp_sync(const float *p, uint16_t *W1, uint16_t *W2, uint16_t *W2T, hipStream_t st)
{
    hipLaunchKernelGGL(k_mlp_sync_bf16, dim3((PM_MLP_H * PM_MLP_H + 255) / 256), dim3(256), 0, st, p, W1, W2, W2T);
    return hipGetLastError();
}

// ---------------------------------------------------------- batch forward --

__global__ __launch_bounds__(64) void k_mlp_forward(MlpDev net, const float *__restrict__ feat, uint32_t n,
                                                    float *__restrict__ y)
{
    __shared__ uint16_t X[64 * PM_MLP_K1];
    __shared__ uint16_t H[64 * 64];
    __shared__ float Y[64];
    const int l = threadIdx.x;
    uint32_t s = blockIdx.x * 64 + l;
    for (int j = 0; j < PM_MLP_K1; j++)
        X[l * PM_MLP_K1 + j] = (j < PM_MLP_IN && s < n) ? f2bf(feat[(uint64_t)s * PM_MLP_IN + j]) : (uint16_t)0;
    wave_lds_sync();
    float out = mlp_wave_forward(net, X, H, Y);
    if (s < n) y[s] = out;
}

hipError_t launch_mlp_forward(const MlpDev &net, const float *feat, uint32_t n, float *y, hipStream_t st)
{
    hipLaunchKernelGGL(k_mlp_forward, dim3((n + 63) / 64), dim3(64), 0, st, net, feat, n, y);
    return hipGetLastError();
}

}  // namespace pm
