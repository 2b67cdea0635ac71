// pm_app_types.h -- stand-ins for the application types the RayMarcher facade
// reads (glm vectors/matrices, Camera3D, CameraController3D, Dataset/Frame and
// the Vulkan swapchain extent), with the member names the reference uses
// (RayMarcher.cpp:76-100, Dataset.h:37-105).  Test harness only; inside the
// application the real headers are used (PM_FACADE_IN_APP).
#pragma once

#include <cstdint>
#include <vector>

namespace glm {
struct vec3 {
    float x = 0, y = 0, z = 0;
    float& operator[](int i) { return (&x)[i]; }
    float operator[](int i) const { return (&x)[i]; }
};
struct vec4 {
    float x = 0, y = 0, z = 0, w = 0;
    float& operator[](int i) { return (&x)[i]; }
    float operator[](int i) const { return (&x)[i]; }
};
struct mat4 {                         // column-major, M[column][row]
    vec4 c[4];
    vec4& operator[](int i) { return c[i]; }
    const vec4& operator[](int i) const { return c[i]; }
};
}  // namespace glm

class Camera3D {
public:
    glm::mat4 InvProjectionView;
    const glm::mat4& GetInvProjectionView() const { return InvProjectionView; }
};

class CameraController3D {
public:
    explicit CameraController3D(Camera3D& camera) : Camera(camera) {}
    Camera3D& Camera;
    glm::vec3 Position;
};

class Frame {
public:
    std::vector<glm::vec3> m_Particles;
};

class Dataset {
public:
    float ParticleRadius = 0.1f;
    float ParticleRadiusExt = 0.2f;     // particleRadiusMultiplier * ParticleRadius
    std::vector<Frame> Frames;
};

struct PmExtent { uint32_t width = 0, height = 0; };
struct PmVulkanStandIn { PmExtent SwapchainExtent; };
extern PmVulkanStandIn Vulkan;
