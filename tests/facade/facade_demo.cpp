// facade_demo.cpp -- drives the RayMarcher facade the way AdvancedRenderer
// does (AdvancedRenderer.cpp:257-298): Prepare, Start, poll IsDone, read the
// images.  Input and output are little-endian binary files written/read by
// tests/test_gpu_facade.py.
//   in : "PMFD" u32 W H n, f32 h, f32 invPV[16], f32 cam[3], i32 max_steps,
//        f32 step, f32 iso, i32 anisotropy, f32 k_n k_r k_s, i32 N_eps,
//        f32 h_ext, f32 depth[W*H], f32 xyz[3n]
//   out: f32 positions[4*W*H], f32 normals[4*W*H], i32 polls
#include <cstdio>
#include <cstring>
#include <vector>

#include "RayMarcher.h"

PmVulkanStandIn Vulkan;

template <class T> static bool rd(FILE* f, T* p, size_t n) { return fread(p, sizeof(T), n, f) == n; }

int main(int argc, char** argv)
{
    if (argc != 3) { std::fprintf(stderr, "usage: facade_demo in out\n"); return 2; }
    FILE* f = std::fopen(argv[1], "rb");
    if (!f) return 2;
    char magic[4];
    uint32_t W, H, n;
    float h, inv[16], cam[3], step, iso, kn, kr, ks, hext;
    int32_t max_steps, aniso, neps;
    bool ok = rd(f, magic, 4) && !std::memcmp(magic, "PMFD", 4) && rd(f, &W, 1) && rd(f, &H, 1) && rd(f, &n, 1) &&
              rd(f, &h, 1) && rd(f, inv, 16) && rd(f, cam, 3) && rd(f, &max_steps, 1) && rd(f, &step, 1) && rd(f, &iso, 1) &&
              rd(f, &aniso, 1) && rd(f, &kn, 1) && rd(f, &kr, 1) && rd(f, &ks, 1) && rd(f, &neps, 1) && rd(f, &hext, 1);
    std::vector<float> depth((size_t)W * H), xyz(3 * (size_t)n);
    ok = ok && rd(f, depth.data(), depth.size()) && rd(f, xyz.data(), xyz.size());
    std::fclose(f);
    if (!ok) { std::fprintf(stderr, "bad input\n"); return 2; }

    Dataset ds;
    ds.ParticleRadius = h;
    ds.ParticleRadiusExt = hext;
    ds.Frames.resize(1);
    ds.Frames[0].m_Particles.resize(n);
    for (uint32_t i = 0; i < n; i++)
        for (int a = 0; a < 3; a++) ds.Frames[0].m_Particles[i][a] = xyz[3 * (size_t)i + a];
    Camera3D camera;
    for (int c = 0; c < 4; c++)
        for (int r = 0; r < 4; r++) camera.InvProjectionView[c][r] = inv[4 * c + r];
    CameraController3D ctl(camera);
    for (int a = 0; a < 3; a++) ctl.Position[a] = cam[a];
    Vulkan.SwapchainExtent.width = W;
    Vulkan.SwapchainExtent.height = H;

    VisualizationSettings vs{};
    vs.Frame = 0;
    vs.MaxSteps = max_steps;
    vs.StepSize = step;
    vs.IsoDensity = iso;
    vs.EnableAnisotropy = aniso != 0;
    vs.k_n = kn;
    vs.k_r = kr;
    vs.k_s = ks;
    vs.N_eps = neps;

    // the staging buffers hold stale data, as mapped Vulkan memory does: the
    // facade must overwrite every pixel (or answer all of them as misses)
    std::vector<glm::vec4> pos((size_t)W * H), nrm((size_t)W * H);
    for (size_t i = 0; i < pos.size(); i++)
        for (int a = 0; a < 4; a++) { pos[i][a] = -7.0f; nrm[i][a] = -7.0f; }
    RayMarcher rm;
    rm.Prepare(vs, ctl, &ds, pos.data(), nrm.data(), depth.data());
    rm.Start();
    int32_t polls = 0;
    while (!rm.IsDone()) polls++;
    if (rm.LastError()[0]) std::fprintf(stderr, "facade_demo: last error: %s\n", rm.LastError());
    rm.Exit();

    FILE* o = std::fopen(argv[2], "wb");
    if (!o) return 2;
    std::fwrite(pos.data(), sizeof(glm::vec4), pos.size(), o);
    std::fwrite(nrm.data(), sizeof(glm::vec4), nrm.size(), o);
    std::fwrite(&polls, sizeof(polls), 1, o);
    std::fclose(o);
    return 0;
}
