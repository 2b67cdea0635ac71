// RayMarcher.h -- drop-in for src/app/AdvancedRenderer/RayMarcher.h (RayMarcher.h:12-77).
//
// Same class name, settings struct and public methods as the reference, so
// AdvancedRenderer (AdvancedRenderer.cpp:257-298) compiles unchanged.  The
// march runs on the GPU through libpmarch.so (include/pmarch.h); Start()
// returns at once and IsDone() reports completion only when the positions and
// normals images are fully written (the reference's counter poll could report
// done while workers were still writing, ThreadPool.cpp:33-36).
//
// Inside the application define PM_FACADE_IN_APP so the application's own
// camera, dataset and glm types are used; the test harness supplies a
// stand-in with the same member names.
#pragma once

#include <cstdint>
#include <vector>

#ifdef PM_FACADE_IN_APP
#include "app/Kernel.h"
#include <engine/camera/Camera3D.h>
#include <engine/camera/CameraController3D.h>
#else
#include "pm_app_types.h"
#endif

struct pm_ctx;
class Dataset;

// RayMarcher.h:12-26
struct VisualizationSettings
{
    int Frame;

    int MaxSteps;
    float StepSize;
    float IsoDensity;

    bool EnableAnisotropy;

    float k_n;
    float k_r;
    float k_s;
    int N_eps;
};

class RayMarcher
{
public:
    RayMarcher();

    void Exit();

    void Prepare(const VisualizationSettings& settings,
                 const CameraController3D& camera,
                 Dataset* dataset,
                 glm::vec4* positions,
                 glm::vec4* normals,
                 float* depth);

    void Start();

    bool IsDone();

    // Not in the reference: the last error reported by the GPU library ("" if none).
    const char* LastError() const;

private:
    void WriteMisses();

    VisualizationSettings m_Settings{};
    Dataset* m_Dataset = nullptr;

    uint32_t m_Width = 0;
    uint32_t m_Height = 0;

    float m_InvProjectionView[16] = {};
    float m_CameraPosition[3] = {};

    glm::vec4* m_Positions = nullptr;
    glm::vec4* m_Normals = nullptr;
    float* m_Depth = nullptr;

    pm_ctx* m_Ctx = nullptr;
    std::vector<bool> m_FrameOnDevice;
    bool m_InFlight = false;
    bool m_FrameReady = false;   // the frame Prepare selected is on the GPU
};
