// Context, error reporting and scratch management of libfbp.
// Replaces CudaHelper / CudaContext (FB/CudaHelper.cpp, FB/CudaContext.cpp:22-58).
#include <cstdarg>
#include <cstdlib>

#include "fbp_internal.h"

namespace fbp {

fbp_status fail(fbp_context ctx, fbp_status st, const char* fmt, ...) {
    if (ctx) {
        char buf[512];
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(buf, sizeof(buf), fmt, ap);
        va_end(ap);
        ctx->last_error = buf;
    }
    return st;
}

fbp_status check_hip(fbp_context ctx, hipError_t e, const char* what) {
    if (e == hipSuccess) return FBP_OK;
    fbp_status st = (e == hipErrorOutOfMemory) ? FBP_ERR_OUT_OF_MEMORY : FBP_ERR_DEVICE;
    return fail(ctx, st, "%s: %s", what, hipGetErrorString(e));
}

fbp_status bind(fbp_context ctx) {
    if (!ctx) return FBP_ERR_INVALID_ARGUMENT;
    FBP_CHECK_HIP(ctx, hipSetDevice(ctx->device));
    return FBP_OK;
}

static fbp_status grow(fbp_context ctx, void** ptr, size_t* have, size_t bytes) {
    if (*have >= bytes) return FBP_OK;
    // Round up to 64 MiB steps to avoid repeated reallocation.
    size_t want = (bytes + (size_t(64) << 20) - 1) & ~((size_t(64) << 20) - 1);
    if (*ptr) {
        FBP_CHECK_HIP(ctx, hipDeviceSynchronize());
        FBP_CHECK_HIP(ctx, hipFree(*ptr));
        *ptr = nullptr;
        *have = 0;
    }
    FBP_CHECK_HIP(ctx, hipMalloc(ptr, want));
    *have = want;
    return FBP_OK;
}

fbp_status ensure_scratch(fbp_context ctx, size_t bytes) {
    return grow(ctx, &ctx->scratch, &ctx->scratch_bytes, bytes);
}

fbp_status ensure_scratch2(fbp_context ctx, size_t bytes) {
    return grow(ctx, &ctx->scratch2, &ctx->scratch2_bytes, bytes);
}

}  // namespace fbp

extern "C" {

const char* fbp_status_string(int status) {
    switch (status) {
        case FBP_OK: return "ok";
        case FBP_ERR_INVALID_ARGUMENT: return "invalid argument";
        case FBP_ERR_DEVICE: return "device (HIP runtime) error";
        case FBP_ERR_OUT_OF_MEMORY: return "out of device memory";
        case FBP_ERR_FFT: return "FFT error";
        case FBP_ERR_UNSUPPORTED: return "unsupported shape or mode";
        case FBP_ERR_PROBABILITIES: return "invalid probabilities";
        case FBP_ERR_IO: return "I/O error";
        default: return "unknown status";
    }
}

fbp_status fbp_context_create(int device, fbp_context* out) {
    if (!out) return FBP_ERR_INVALID_ARGUMENT;
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0) return FBP_ERR_DEVICE;
    if (device >= count) return FBP_ERR_INVALID_ARGUMENT;
    if (device < 0) {
        // Largest clock rate x multiprocessor count, as CudaContext::setDevice.
        long long best = -1;
        for (int d = 0; d < count; ++d) {
            hipDeviceProp_t p;
            if (hipGetDeviceProperties(&p, d) != hipSuccess) continue;
            long long score = (long long)p.clockRate * p.multiProcessorCount;
            if (score > best) {
                best = score;
                device = d;
            }
        }
        if (device < 0) return FBP_ERR_DEVICE;
    }
    fbp_context ctx = new fbp_context_s();
    ctx->device = device;
    if (hipSetDevice(device) != hipSuccess) {
        delete ctx;
        return FBP_ERR_DEVICE;
    }
    hipDeviceProp_t p;
    if (hipGetDeviceProperties(&p, device) != hipSuccess) {
        delete ctx;
        return FBP_ERR_DEVICE;
    }
    ctx->num_cus = p.multiProcessorCount;
    ctx->max_lds_per_block = int(p.sharedMemPerBlock);
    if (hipHostMalloc((void**)&ctx->pinned, 64 * sizeof(float), hipHostMallocDefault) != hipSuccess) {
        delete ctx;
        return FBP_ERR_OUT_OF_MEMORY;
    }
    // Library set-up for the no-alloc entry point, with the device current:
    // the per-device address of its scratch array.  A failure here only
    // disables that entry point (it reports the recorded message).
    if (fbp::noalloc_scratch_address(ctx, &ctx->noalloc_scratch) != FBP_OK)
        ctx->noalloc_scratch = nullptr;
    *out = ctx;
    return FBP_OK;
}

fbp_status fbp_context_set_option(fbp_context ctx, int32_t option, int32_t value) {
    if (!ctx) return FBP_ERR_INVALID_ARGUMENT;
    switch (option) {
        case FBP_OPTION_CONFOCAL_ALIGNED:
            if (value != 0 && value != 1)
                return fbp::fail(ctx, FBP_ERR_INVALID_ARGUMENT, "option value must be 0 or 1");
            ctx->disable_aligned = value ? 0 : 1;
            return FBP_OK;
        case FBP_OPTION_VOXEL_SCATTERING:
            if (value != 0 && value != 1)
                return fbp::fail(ctx, FBP_ERR_INVALID_ARGUMENT, "option value must be 0 or 1");
            ctx->voxel_scattering = value;
            return FBP_OK;
        case FBP_OPTION_EXHAUSTIVE_LASER_GROUPS:
            if (value != 0 && value != 1)
                return fbp::fail(ctx, FBP_ERR_INVALID_ARGUMENT, "option value must be 0 or 1");
            ctx->exhaustive_group_kernel = value;
            return FBP_OK;
        default:
            return fbp::fail(ctx, FBP_ERR_INVALID_ARGUMENT, "unknown option %d", int(option));
    }
}

fbp_status fbp_context_get_option(fbp_context ctx, int32_t option, int32_t* value) {
    if (!ctx || !value) return FBP_ERR_INVALID_ARGUMENT;
    switch (option) {
        case FBP_OPTION_CONFOCAL_ALIGNED:
            *value = ctx->disable_aligned ? 0 : 1;
            return FBP_OK;
        case FBP_OPTION_VOXEL_SCATTERING:
            *value = ctx->voxel_scattering;
            return FBP_OK;
        case FBP_OPTION_EXHAUSTIVE_LASER_GROUPS:
            *value = ctx->exhaustive_group_kernel;
            return FBP_OK;
        default:
            return fbp::fail(ctx, FBP_ERR_INVALID_ARGUMENT, "unknown option %d", int(option));
    }
}

fbp_status fbp_context_last_exhaustive_kernel(fbp_context ctx, int32_t* laser_groups) {
    if (!ctx || !laser_groups) return FBP_ERR_INVALID_ARGUMENT;
    *laser_groups = ctx->last_exhaustive_group;
    return FBP_OK;
}

fbp_status fbp_context_last_confocal_path(fbp_context ctx, int32_t* aligned) {
    if (!ctx || !aligned) return FBP_ERR_INVALID_ARGUMENT;
    FBP_CHECK(fbp::bind(ctx));
    *aligned = 0;
    if (!ctx->path_flag_valid || !ctx->path_flag) return FBP_OK;
    int v = 0;
    FBP_CHECK_HIP(ctx, hipDeviceSynchronize());
    FBP_CHECK_HIP(ctx, hipMemcpy(&v, ctx->path_flag, sizeof(int), hipMemcpyDeviceToHost));
    *aligned = v;
    return FBP_OK;
}

fbp_status fbp_context_destroy(fbp_context ctx) {
    if (!ctx) return FBP_ERR_INVALID_ARGUMENT;
    hipSetDevice(ctx->device);
    hipDeviceSynchronize();
    fbp::fft_cache_destroy(ctx);
    if (ctx->scratch) hipFree(ctx->scratch);
    if (ctx->scratch2) hipFree(ctx->scratch2);
    if (ctx->path_flag) hipFree(ctx->path_flag);
    if (ctx->aligned_tables) hipFree(ctx->aligned_tables);
    if (ctx->pinned) hipHostFree(ctx->pinned);
    if (ctx->pure_thr) hipFree(ctx->pure_thr);
    delete ctx;
    return FBP_OK;
}

const char* fbp_context_last_error(fbp_context ctx) {
    return ctx ? ctx->last_error.c_str() : "null context";
}

int fbp_context_device(fbp_context ctx) { return ctx ? ctx->device : -1; }

size_t fbp_context_scratch_bytes(fbp_context ctx) {
    return ctx ? ctx->scratch_bytes + ctx->scratch2_bytes : 0;
}

}  // extern "C"
