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
};

namespace fbp {

// Records a message on the context and returns the status.
fbp_status fail(fbp_context ctx, fbp_status st, const char* fmt, ...);

// hipError -> fbp_status with a message naming the call.
fbp_status check_hip(fbp_context ctx, hipError_t e, const char* what);

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
