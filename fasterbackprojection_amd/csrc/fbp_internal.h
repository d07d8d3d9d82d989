// Internal declarations shared by the HIP translation units of libfbp.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>

#include "fbp.h"

struct fbp_context_s {
    int device = 0;
    int num_cus = 0;
    int max_lds_per_block = 65536;
    // Confocal backprojection: 0 = the aligned-grid path may run when its
    // conditions hold (default), 1 = always the general kernel.
    int disable_aligned = 0;
    // 1: confocal voxel and MIS backprojection take the voxel-scattering
    // walk over every bin a voxel reaches (reference compile-time switch).
    int voxel_scattering = 0;
    // 1: the exhaustive voxel backprojection may take the laser-group kernel
    // (bp_exhaustive.hip) when the capture qualifies; 0: per-pair kernel only.
    int exhaustive_group_kernel = 1;
    // Kernel of the last exhaustive voxel call: 1 = laser-group kernel.
    int last_exhaustive_group = 0;
    // Whether the laser-group kernels' two-term quotient truncates like
    // fl(dist / dt) at every bin boundary, for the (dt, T) it was checked at.
    int group_quot_ok = 0;
    float group_quot_dt = 0.f;
    uint32_t group_quot_T = 0;
    // Device int: path flag of the last confocal voxel call (1 = aligned
    // path), valid when path_flag_valid is set.
    int* path_flag = nullptr;
    // Offset tables of the aligned confocal path, kept across calls.
    void* aligned_tables = nullptr;
    size_t aligned_tables_bytes = 0;
    bool path_flag_valid = false;
    std::string last_error;
    // Scratch arena: grown on demand, never shrunk.  Used for per-chunk
    // partial sums, window tables, FFT buffers.
    void* scratch = nullptr;
    size_t scratch_bytes = 0;
    // Second arena for FFT work areas (rocFFT) and LCT spectra.
    void* scratch2 = nullptr;
    size_t scratch2_bytes = 0;
    // Small host-pinned staging for reductions (min/max).
    float* pinned = nullptr;
    // Cached rocFFT plans and the LCT inverse-psf spectrum (owned by fourier.hip).
    void* fft_cache = nullptr;
    // Device address, on this context's device, of the module-scope partial
    // sum array of fbp_backproject_exhaustive_voxel_noalloc (resolved at
    // context creation; nullptr when that failed).
    void* noalloc_scratch = nullptr;
    // Threshold row of the pure confocal chain (t0 = 0, bounces discarded):
    // it depends on (T, dt) only, so it is kept for the next call with the
    // same pair (float[pure_thr_T + 1]).
    float* pure_thr = nullptr;
    uint32_t pure_thr_T = 0;
    float pure_thr_dt = 0.f;
    size_t pure_thr_cap = 0;
};


namespace fbp {

// Records a message on the context and returns the status.
fbp_status fail(fbp_context ctx, fbp_status st, const char* fmt, ...);

// hipError -> fbp_status with a message naming the call.
fbp_status check_hip(fbp_context ctx, hipError_t e, const char* what);

// Per-block (min, max) pairs of a volume: part[2 i], part[2 i + 1], i < count.
struct MinMaxPartials {
    float* part = nullptr;
    uint32_t count = 0;
};

// LoG (filters.hip); with mm, also the output's per-block min / max pairs
// (in ctx->scratch: valid until the next use of the scratch arena).
fbp_status log_impl(fbp_context ctx, const float* in, float* out, fbp_uvec3 size,
                    int32_t kernel_size, float sigma, hipStream_t s, MinMaxPartials* mm);

// Min and max over per-block pairs, reduced by the whole block (every
// thread gets the result).  Order-independent, so every block of a grid
// gets the same values.
__device__ __forceinline__ void block_minmax(const float* __restrict__ part, uint32_t count,
                                             float& mn, float& mx) {
    __shared__ float smn[256], smx[256];
    float a = 3.402823466e+38f, b = -3.402823466e+38f;
    for (uint32_t i = threadIdx.x; i < count; i += blockDim.x) {
        a = fminf(a, part[2 * i]);
        b = fmaxf(b, part[2 * i + 1]);
    }
    smn[threadIdx.x] = a;
    smx[threadIdx.x] = b;
    __syncthreads();
    for (int w = int(blockDim.x) / 2; w > 0; w >>= 1) {
        if (int(threadIdx.x) < w) {
            smn[threadIdx.x] = fminf(smn[threadIdx.x], smn[threadIdx.x + w]);
            smx[threadIdx.x] = fmaxf(smx[threadIdx.x], smx[threadIdx.x + w]);
        }
        __syncthreads();
    }
    mn = smn[0];
    mx = smx[0];
    __syncthreads();
}

// fbp_learned_calibrate on a stream (calibrate.hip).
fbp_status learned_calibrate_impl(fbp_context ctx, const fbp_reconstruction_info* info,
                                  const fbp_transient* tr, const fbp_vec3* laser_targets,
                                  float* volume, const float* reference, int32_t kernel_size,
                                  float sigma, uint32_t extremes, hipStream_t s);

// Ensures ctx->scratch holds at least `bytes` (device).  Synchronises the
// device only when it has to grow.
fbp_status ensure_scratch(fbp_context ctx, size_t bytes);
fbp_status ensure_scratch2(fbp_context ctx, size_t bytes);

// Device address of the no-alloc scratch array (bp_exhaustive.hip) on
// ctx->device, which must be current; looked up once per device.
fbp_status noalloc_scratch_address(fbp_context ctx, void** out);

// Set the context's device current for this host thread.
fbp_status bind(fbp_context ctx);

void fft_cache_destroy(fbp_context ctx);

inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

constexpr float kEps = 0.00001f;                // EPS, FB/math.cuh:11
constexpr float kFltMax = 3.402823466e+38f;     // FLT_MAX

}  // namespace fbp

#define FBP_CHECK_HIP(ctx, call)                                      \
    do {                                                              \
        hipError_t e__ = (call);                                      \
        if (e__ != hipSuccess) return fbp::check_hip(ctx, e__, #call); \
    } while (0)

#define FBP_CHECK(call)                      \
    do {                                     \
        fbp_status s__ = (call);             \
        if (s__ != FBP_OK) return s__;       \
    } while (0)
