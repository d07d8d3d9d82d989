// Pieces shared by the laser-group kernels (exhaustive voxel backprojection
// and exhaustive depth planes): the exact sensor-distance square root, the
// two-term bin quotient and the host check that makes it exact, and the
// box bounds used for bin windows.
#pragma once

#include <cmath>
#include <cstdint>

#include "chain.h"
#include "fbp_internal.h"

namespace fbp {

// Lasers per group: the lasers whose distances a lane keeps in registers.
constexpr int kGroup = 8;

// Correctly rounded square root of a normal x (here x >= 1e-6: squared
// distances of pairs at least 1 mm apart): the hardware estimate s (within
// an ulp) and its neighbours, chosen by the signs of the residuals
// x - s' s computed exactly by FMA; the same selection as the compiler's
// IEEE square root, without its rescaling of tiny inputs and its
// special-value cases.
__device__ __forceinline__ float sqrt_rn_normal(float x) {
    const float s = __builtin_amdgcn_sqrtf(x);
    const float sd = __int_as_float(__float_as_int(s) - 1);
    const float su = __int_as_float(__float_as_int(s) + 1);
    float r = (fmaf(-sd, s, x) <= 0.f) ? sd : s;
    r = (fmaf(-su, s, x) > 0.f) ? su : r;
    return r;
}

// The reference's bin of dist from a two-term product, exact wherever the
// host's check passed: c_hi = fl(1 / dt), c_lo = fl(1 / dt - c_hi), so
// c_hi + c_lo is 1 / dt to about 2^-48 and q = fl(dist c_hi + fl(dist c_lo))
// is dist / dt rounded once from a value within about 2^-47 of it.  Host and
// device evaluate it with the same float operations.
__host__ __device__ inline float group_quotient(float dist, float c_hi, float c_lo) {
    return fmaf(dist, c_hi, dist * c_lo);
}

// c_lo of group_quotient for dt.
inline float group_inv_dt_lo(float dt) {
    return static_cast<float>(1.0 / double(dt) - double(1.0f / dt));
}

// Distance from p to the box [lo, hi] with the hardware square root (1 ulp):
// bounds only (window start, near test), all with margins.
__device__ __forceinline__ float box_dist_fast(const float p[3], const float lo[3],
                                               const float hi[3], bool far) {
    float s = 0.f;
    for (int ax = 0; ax < 3; ++ax) {
        const float d = far ? fmaxf(fabsf(lo[ax] - p[ax]), fabsf(hi[ax] - p[ax]))
                            : fmaxf(fmaxf(lo[ax] - p[ax], p[ax] - hi[ax]), 0.f);
        s += d * d;
    }
    return __builtin_amdgcn_sqrtf(s);
}

// Whether group_quotient truncates like the reference's fl(dist / dt) for
// every dist the bins 1 .. T + 1 can meet.  Both quotients are dist / dt
// rounded once (from values at most about 2^-47 apart), so they can only
// truncate differently when dist / dt is within an ulp of an integer k, i.e.
// dist within an ulp or so of k dt; every float within 16 ulps of each k dt
// is checked.
inline bool group_quotient_exact(float dt, uint32_t T) {
    const float c_hi = 1.0f / dt, c_lo = group_inv_dt_lo(dt);
    for (uint32_t k = 1; k <= T + 1; ++k) {
        float x = float(double(k) * double(dt));
        for (int i = 0; i < 16; ++i) x = std::nextafter(x, 0.0f);
        for (int i = 0; i <= 32; ++i, x = std::nextafter(x, INFINITY)) {
            if (static_cast<int>(group_quotient(x, c_hi, c_lo)) != static_cast<int>(x / dt))
                return false;
        }
    }
    return true;
}

// group_quotient_exact for (dt, T), kept on the context for the next call.
inline bool group_quotient_cached(fbp_context ctx, float dt, uint32_t T) {
    if (ctx->group_quot_dt != dt || ctx->group_quot_T != T) {
        ctx->group_quot_ok = group_quotient_exact(dt, T) ? 1 : 0;
        ctx->group_quot_dt = dt;
        ctx->group_quot_T = T;
    }
    return ctx->group_quot_ok != 0;
}

}  // namespace fbp
