// RayMarcher.cpp -- the reference's marcher interface (RayMarcher.cpp:66-112)
// over libpmarch.so.  Prepare() records the view and uploads the frame's
// particles once (the dataset is immutable after load, Dataset.cpp:169-227);
// Start() launches the reference march the settings select (RayMarcher.cpp:
// 102-112): PerPixel_Isotropic (:256-344) or, with EnableAnisotropy,
// PerPixel_Anisotropic (:346-423) -- depth-buffer start, MaxSteps fixed steps,
// DensityGrid skip, first sample with rho >= IsoDensity, normal =
// normalize(sum gradW).
#include "RayMarcher.h"

#include <cstdio>
#include <cstring>

#include "pmarch.h"

#ifdef PM_FACADE_IN_APP
#include "app/Dataset.h"
#include <engine/renderer/Renderer.h>
#endif

RayMarcher::RayMarcher()
{
    if (pm_create(0, &m_Ctx) != PM_OK) {
        std::fprintf(stderr, "RayMarcher: no GPU available for the march\n");
        m_Ctx = nullptr;
    }
}

// Every pixel a miss: (0, 0, 0, 0) in both images, as the reference leaves a
// pixel whose ray finds no surface (RayMarcher.cpp:262-263).  Used whenever the
// GPU march cannot deliver the frame, so the caller never uploads what its
// staging buffers held before.
void RayMarcher::WriteMisses()
{
    const size_t n = (size_t)m_Width * m_Height;
    if (m_Positions) std::memset((void*)m_Positions, 0, n * sizeof(glm::vec4));
    if (m_Normals) std::memset((void*)m_Normals, 0, n * sizeof(glm::vec4));
}

void RayMarcher::Exit()
{
    if (m_Ctx) {
        pm_wait(m_Ctx);
        pm_destroy(m_Ctx);
        m_Ctx = nullptr;
    }
}

void RayMarcher::Prepare(const VisualizationSettings& settings,
                         const CameraController3D& camera,
                         Dataset* dataset,
                         glm::vec4* positions,
                         glm::vec4* normals,
                         float* depth)
{
    m_Settings = settings;
    m_Dataset = dataset;

    m_Width = Vulkan.SwapchainExtent.width;                  // RayMarcher.cpp:86-87
    m_Height = Vulkan.SwapchainExtent.height;

    m_Positions = positions;
    m_Normals = normals;
    m_Depth = depth;

    const glm::mat4& M = camera.Camera.GetInvProjectionView();   // RayMarcher.cpp:95-96
    for (int c = 0; c < 4; c++)
        for (int r = 0; r < 4; r++) m_InvProjectionView[4 * c + r] = M[c][r];
    for (int a = 0; a < 3; a++) m_CameraPosition[a] = camera.Position[a];

    m_FrameReady = false;
    if (!m_Ctx) return;
    const size_t f = (size_t)settings.Frame;
    if (m_FrameOnDevice.size() <= f) m_FrameOnDevice.resize(f + 1, false);
    if (!m_FrameOnDevice[f]) {
        const auto& particles = m_Dataset->Frames[f].m_Particles;
        if (pm_set_frame(m_Ctx, settings.Frame, (const float*)particles.data(), particles.size(),
                         m_Dataset->ParticleRadius) == PM_OK)
            m_FrameOnDevice[f] = true;
        else
            std::fprintf(stderr, "RayMarcher: %s\n", pm_last_error(m_Ctx));
    }
    m_FrameReady = m_FrameOnDevice[f];
}

void RayMarcher::Start()
{
    m_InFlight = false;
    if (!m_Ctx || !m_FrameReady) {
        // no context, or the frame could not be put on the GPU: the images are
        // answered as misses and LastError() says why
        WriteMisses();
        return;
    }
    pm_settings s{};
    s.mode = PM_MODE_REFERENCE;
    s.max_steps = m_Settings.MaxSteps;
    s.step = m_Settings.StepSize;
    s.iso = m_Settings.IsoDensity;
    s.normal_mode = PM_NORMAL_GRADW;
    s.skip_mode = PM_SKIP_DG_BOX;
    s.compensated = 0;
    if (m_Settings.EnableAnisotropy) {                          // RayMarcher.cpp:104-106
        s.anisotropic = 1;
        s.k_n = m_Settings.k_n;
        s.k_r = m_Settings.k_r;
        s.k_s = m_Settings.k_s;
        s.n_eps = m_Settings.N_eps;
        s.ext_mult = m_Dataset->ParticleRadiusExt / m_Dataset->ParticleRadius;   // Dataset.cpp:169-177
    }
    int rc = pm_march_async(m_Ctx, m_Settings.Frame, &s, m_InvProjectionView, m_CameraPosition, m_Width, m_Height,
                            m_Depth, (float*)m_Positions, (float*)m_Normals, nullptr);
    m_InFlight = rc == PM_OK;
    if (rc != PM_OK) {
        std::fprintf(stderr, "RayMarcher: %s\n", pm_last_error(m_Ctx));
        WriteMisses();
    }
}

bool RayMarcher::IsDone()
{
    if (!m_Ctx || !m_InFlight) return true;
    int rc = pm_query(m_Ctx);
    if (rc == PM_PENDING) return false;
    m_InFlight = false;
    if (rc != PM_OK) {
        std::fprintf(stderr, "RayMarcher: %s\n", pm_last_error(m_Ctx));
        WriteMisses();
    }
    return true;
}

const char* RayMarcher::LastError() const { return m_Ctx ? pm_last_error(m_Ctx) : "no GPU context"; }
