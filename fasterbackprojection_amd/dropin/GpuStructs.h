// Host-side data model of the drop-in C++ surface: the reference's
// ReconstructionInfo / ReconstructionBuffers / CaptureSystem
// (FB/GpuStructs.cuh:3-45) over plain POD vectors with glm's memory layout.
#pragma once

#include <cstdint>

#include "fbp.h"

namespace fbp {
struct vec2 { float x = 0.f, y = 0.f; };
struct vec3 {
    float x = 0.f, y = 0.f, z = 0.f;
    vec3() = default;
    constexpr vec3(float a, float b, float c) : x(a), y(b), z(c) {}
    explicit constexpr vec3(float a) : x(a), y(a), z(a) {}
};
struct uvec3 {
    uint32_t x = 0, y = 0, z = 0;
    uvec3() = default;
    constexpr uvec3(uint32_t a, uint32_t b, uint32_t c) : x(a), y(b), z(c) {}
    explicit constexpr uvec3(uint32_t a) : x(a), y(a), z(a) {}
};
inline vec3 operator+(vec3 a, vec3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline vec3 operator-(vec3 a, vec3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline vec3 operator*(vec3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
inline vec3 operator/(vec3 a, vec3 b) { return {a.x / b.x, a.y / b.y, a.z / b.z}; }
static_assert(sizeof(vec3) == 12 && sizeof(uvec3) == 12 && sizeof(vec2) == 8, "glm layout");
}  // namespace fbp

enum CaptureSystem : uint32_t { Confocal = 0, Exhaustive = 1 };

// FB/GpuStructs.cuh:8-38, field for field; layout-identical to
// fbp_reconstruction_info (148 bytes).
struct ReconstructionInfo {
    fbp::vec3 _sensorPosition;
    CaptureSystem _captureSystem = Confocal;
    fbp::vec3 _laserPosition;
    float _timeOffset = 0.f;
    uint32_t _numLaserTargets = 0;
    uint32_t _numSensorTargets = 0;
    uint32_t _numTimeBins = 0;
    float _timeStep = 0.f;
    fbp::vec3 _hiddenVolumeMin;
    uint32_t _discardFirstLastBounces = 0;
    fbp::vec3 _hiddenVolumeMax;
    fbp::vec3 _hiddenVolumeSize;
    fbp::vec3 _hiddenVolumeVoxelSize;
    fbp::vec3 _relayWallMinPosition;
    fbp::vec3 _relayWallSize;
    fbp::vec3 _relayWallNormal;
    fbp::uvec3 _voxelResolution;

    // getFocusDepth (FB/GpuStructs.cuh:34-37)
    float getFocusDepth() const {
        const float m = (_hiddenVolumeMin.y + _hiddenVolumeMax.y) * 0.5f;
        return m < 0.f ? -m : m;
    }
    const fbp_reconstruction_info* abi() const {
        return reinterpret_cast<const fbp_reconstruction_info*>(this);
    }
};
static_assert(sizeof(ReconstructionInfo) == 148, "ReconstructionInfo must be 148 bytes");
static_assert(sizeof(ReconstructionInfo) == sizeof(fbp_reconstruction_info), "ABI mirror");

// FB/GpuStructs.cuh:40-45: device buffers filled by NLosData::toGpu; the
// caller frees them (FB/Laser.cu:64-66).
struct ReconstructionBuffers {
    fbp::vec3* _laserTargets = nullptr;
    fbp::vec3* _sensorTargets = nullptr;
    float* _intensity = nullptr;
};
