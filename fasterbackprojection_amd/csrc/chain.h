// Exact IEEE-754 restatement of the reference's per-update arithmetic, for the
// device (and the host, where the same source is compiled for window bounds).
//
// The reference evaluates (FB/transient_reconstruction.cuh:102-110, 128):
//   checkPos = (hmin + vec3(vx, z, vy) * vs) + 0.5f * vs           componentwise
//   r        = glm::distance(l, checkPos) = sqrt((dx*dx + dy*dy) + dz*dz),
//              d = checkPos - l
//   dist     = r * 2 - t0        [+ (|Lpos - l| + |Spos - l|) if !discard]
//   bin      = (int)(dist / Delta)
//   weight   = safeRCP(dist * dist)
// Every operation here is a separately rounded IEEE op (no contraction, no
// approximate sqrt/div): that sequence is what "bit-exact bins" is checked
// against.  The fast kernels compute an approximation and fall back to these
// functions whenever the approximation could round to a different bin.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace fbp {

#define FBP_HD __host__ __device__ __forceinline__

// Separately rounded IEEE operations.  The library is compiled with
// -ffp-contract=off, so no product/sum is fused.  On the device, HIP's f32
// division and sqrt are correctly rounded by default
// (-fhip-fp32-correctly-rounded-divide-sqrt); __fsqrt_rn is NOT used because
// without OCML_BASIC_ROUNDED_OPERATIONS it maps to the native (approximate)
// square root.  On the host, x86-64 SSE scalar arithmetic is IEEE.
FBP_HD float rn_add(float a, float b) { return a + b; }
FBP_HD float rn_sub(float a, float b) { return a - b; }
FBP_HD float rn_mul(float a, float b) { return a * b; }
FBP_HD float rn_div(float a, float b) { return a / b; }
FBP_HD float rn_sqrt(float a) { return __builtin_sqrtf(a); }

// Voxel centre coordinate along one axis: (hmin + float(i) * vs) + 0.5f * vs.
FBP_HD float voxel_center(float hmin, float vs, uint32_t i) {
    return rn_add(rn_add(hmin, rn_mul(static_cast<float>(i), vs)), rn_mul(0.5f, vs));
}

// fl(d*d), the per-axis term of glm::dot.
FBP_HD float sq(float d) { return rn_mul(d, d); }

// ((x2 + y2) + z2) in glm::dot's order.
FBP_HD float dot3(float x2, float y2, float z2) { return rn_add(rn_add(x2, y2), z2); }

// glm::distance(a, b) = length(b - a).
FBP_HD float distance3(float ax, float ay, float az, float bx, float by, float bz) {
    return rn_sqrt(dot3(sq(rn_sub(bx, ax)), sq(rn_sub(by, ay)), sq(rn_sub(bz, az))));
}

// Confocal dist from the exact squared distance s (FB/transient_reconstruction.cuh:108-110).
// off = (|Lpos - l| + |Spos - l|) already rounded; used only when !discard.
FBP_HD float confocal_dist(float s, float t0, uint32_t discard, float off) {
    float d = rn_sub(rn_mul(rn_sqrt(s), 2.0f), t0);
    if (discard == 0) d = rn_add(d, off);
    return d;
}

// Exhaustive dist from the two exact squared distances (:155-157).
FBP_HD float exhaustive_dist(float sl, float ss, float t0, uint32_t discard, float off) {
    float d = rn_sub(rn_add(rn_sqrt(sl), rn_sqrt(ss)), t0);
    if (discard == 0) d = rn_add(d, off);
    return d;
}

// (int)(dist / Delta) with C truncation toward zero.  Values outside int
// range cannot reach a staged window (callers bound dist first).
FBP_HD int bin_of(float dist, float dt) {
    float q = rn_div(dist, dt);
    return static_cast<int>(q);
}

// Half the voxel diagonal, glm::length(voxel size) / 2
// (FB/transient_reconstruction.cuh:113).
FBP_HD float voxel_extent(float vx, float vy, float vz) {
    return rn_div(rn_sqrt(dot3(sq(vx), sq(vy), sq(vz))), 2.0f);
}

// Bins visited by the voxel-scattering walk of one (voxel, scan point) pair
// (FB/transient_reconstruction.cuh:114-125, 235-250): d = max(0, dist - e),
// then d += (Delta + EPS) while d < dist + e, calling f(bin) with
// bin = (int)(d / Delta); the walk ends at the first bin outside [0, T)
// (d >= 0, so only bin >= T can occur).  At most T + 2 steps: a step below
// the spacing of floats at d would not advance the walk.
template <class F>
FBP_HD void scatter_walk(float dist, float ext, float dt, uint32_t T, F&& f) {
    const float lo = rn_sub(dist, ext);
    float d = 0.0f < lo ? lo : 0.0f;
    const float hi = rn_add(dist, ext);
    const float step = rn_add(dt, 0.00001f);
    for (uint32_t n = 0; n < T + 2u && d < hi; ++n) {
        const float q = rn_div(d, dt);
        if (!(q < static_cast<float>(T))) return;
        f(static_cast<int>(q));
        d = rn_add(d, step);
    }
}

// safeRCP (FB/math.cuh:23-29).
FBP_HD float safe_rcp(float x) {
    if (x > 0.00001f || x < -0.00001f) return rn_div(1.0f, x);
    return x >= 0.0f ? 3.402823466e+38f : -3.402823466e+38f;
}

}  // namespace fbp
