// Exact backprojection at listed voxels, and the learned form's calibration
// of its output.
//
// fbp_backproject_confocal_points: the confocal voxel backprojection
// (backprojectConfocalVoxel, FB/transient_reconstruction.cuh:93-136) evaluated
// only at the listed voxels, with the reference's IEEE chain per pair: one
// wave per voxel, lanes over scan points, FP64 wave sum.
//
// fbp_learned_calibrate: the refiner predicts the min-max normalised LoG of
// the backprojection up to a global affine map (the normalisation of its
// sampled input is set by noisy extremes, which its convolutions cannot see).
// Two facts fix the map.  (1) The LoG in physical units: regressing an
// independent unbiased estimate of it (the caller's `reference`, e.g. the
// LoG of a second MIS draw) on the refined volume over all voxels gives
// L ~ a v + b; the reference's noise is independent of v, so it averages
// out of the fit.  (2) The true range: the exact LoG at the k lowest and k
// highest voxels of the refined volume (exact backprojections of their 5^3
// clamp-to-edge neighbourhoods).  Output: (a v + b - min) / (max - min).
// No host round trip.
#include <algorithm>
#include <vector>

#include "bp_common.h"

namespace fbp {

namespace {

constexpr int kTaps = 125;        // 5^3 LoG neighbourhood
constexpr int kMaxExtremes = 32;  // k lowest and k highest candidates

struct Weights {
    float w[kTaps];
};

struct PointArgs {
    const float* I;      // [L][T] time-fastest
    const fbp_vec3* laser;
    uint32_t L, T;
    uint32_t X, Y, Z;
    float hmin[3], vs[3];
    float t0, dt;
    uint32_t discard;
    fbp_vec3 lpos, spos;
};

// Exact confocal backprojection sample of one (voxel centre, scan point) pair.
__device__ __forceinline__ double point_sample(const PointArgs& a, uint32_t l, float px, float py,
                                               float pz) {
    const fbp_vec3 lp = a.laser[l];
    float dist = rn_sub(rn_mul(distance3(lp.x, lp.y, lp.z, px, py, pz), 2.0f), a.t0);
    if (a.discard == 0)
        dist = rn_add(dist, rn_add(distance3(lp.x, lp.y, lp.z, a.lpos.x, a.lpos.y, a.lpos.z),
                                   distance3(lp.x, lp.y, lp.z, a.spos.x, a.spos.y, a.spos.z)));
    const float q = rn_div(dist, a.dt);
    if (!(q > -1.0f) || !(q < float(a.T))) return 0.0;
    const float light = a.I[size_t(l) * a.T + uint32_t(int(q))];
    return light > kEps ? double(rn_mul(light, safe_rcp(rn_mul(dist, dist)))) : 0.0;
}

// One wave per voxel; 4 voxels per 256-thread block.
__global__ __launch_bounds__(256) void bp_points_kernel(PointArgs a, const uint32_t* __restrict__ vox,
                                                        uint32_t count, double* __restrict__ out) {
    const uint32_t p = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (p >= count) return;  // whole wave exits together
    const uint32_t v = vox[p];
    const uint32_t vx = v % a.X, vy = (v / a.X) % a.Y, z = v / (a.X * a.Y);
    const float px = voxel_center(a.hmin[0], a.vs[0], vx);
    const float py = voxel_center(a.hmin[1], a.vs[1], z);
    const float pz = voxel_center(a.hmin[2], a.vs[2], vy);
    double acc = 0.0;
    for (uint32_t l = uint32_t(lane); l < a.L; l += 64) acc += point_sample(a, l, px, py, pz);
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off, 64);
    if (lane == 0) out[p] = acc;
}

// The k smallest (blockIdx.y = 0) or largest (1) (value, index) pairs of
// each wave's segment of 64 * PER_LANE entries, by k rounds of a wave-wide
// arg-min (ties by index).  idx == nullptr: the indices are the positions
// in the input.  Entries beyond n never win while n >= k.
template <int PER_LANE>
__global__ __launch_bounds__(256) void wave_extremes_kernel(const float* __restrict__ in_lo,
                                                            const uint32_t* __restrict__ idx_lo,
                                                            const float* __restrict__ in_hi,
                                                            const uint32_t* __restrict__ idx_hi,
                                                            uint32_t n, uint32_t k,
                                                            float* __restrict__ lo_v,
                                                            uint32_t* __restrict__ lo_i,
                                                            float* __restrict__ hi_v,
                                                            uint32_t* __restrict__ hi_i) {
    const bool high = blockIdx.y == 1;
    const float* in = high ? in_hi : in_lo;
    const uint32_t* in_idx = high ? idx_hi : idx_lo;
    float* out_v = high ? hi_v : lo_v;
    uint32_t* out_i = high ? hi_i : lo_i;
    const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t seg = wave * (64u * PER_LANE);
    if (seg >= n) return;  // whole wave
    // The largest are the smallest of -value: one arg-min loop serves both.
    const float sgn = high ? -1.f : 1.f;
    float v[PER_LANE];
    uint32_t id[PER_LANE];
#pragma unroll
    for (int j = 0; j < PER_LANE; ++j) {
        const uint32_t g = seg + uint32_t(j) * 64 + lane;
        const bool ok = g < n;
        v[j] = ok ? sgn * in[g] : __builtin_inff();
        id[j] = ok ? (in_idx ? in_idx[g] : g) : 0xffffffffu;
    }
    auto better = [](float a, uint32_t ia, float b, uint32_t ib) {
        return a < b || (a == b && ia < ib);
    };
    for (uint32_t r = 0; r < k; ++r) {
        float bv = v[0];
        uint32_t bi = id[0];
#pragma unroll
        for (int j = 1; j < PER_LANE; ++j)
            if (better(v[j], id[j], bv, bi)) {
                bv = v[j];
                bi = id[j];
            }
        for (int off = 32; off > 0; off >>= 1) {
            const float ov = __shfl_xor(bv, off, 64);
            const uint32_t oi = __shfl_xor(bi, off, 64);
            if (better(ov, oi, bv, bi)) {
                bv = ov;
                bi = oi;
            }
        }
#pragma unroll
        for (int j = 0; j < PER_LANE; ++j)
            if (id[j] == bi && bi != 0xffffffffu) {
                v[j] = __builtin_inff();
                id[j] = 0xffffffffu;
            }
        if (lane == 0) {
            out_v[size_t(wave) * k + r] = sgn * bv;
            out_i[size_t(wave) * k + r] = bi;
        }
    }
}

// Exact LoG contribution of one chunk of scan points at extreme candidate ci
// (ci < k: lows, else highs): block (ci, chunk); thread t < 125 backprojects
// neighbour t = (dz, dy, dx) of the candidate (clamped to the volume) over
// the chunk.  All threads read the same scan point's row at nearby bins
// (coalesced).  partial[ci * chunks + chunk] = sum_t w[t] bp_t (FP64, fixed
// order).
__global__ __launch_bounds__(128) void calib_neigh_kernel(PointArgs a, Weights w,
                                                          const uint32_t* __restrict__ lo,
                                                          const uint32_t* __restrict__ hi, uint32_t k,
                                                          uint32_t chunk_len,
                                                          double* __restrict__ partial) {
    __shared__ double red[128];
    const uint32_t ci = blockIdx.x, chunk = blockIdx.y, t = threadIdx.x;
    double acc = 0.0;
    if (t < uint32_t(kTaps)) {
        const uint32_t c = ci < k ? lo[ci] : hi[ci - k];
        const int x = int(c % a.X), y = int((c / a.X) % a.Y), z = int(c / (a.X * a.Y));
        const int nx = min(max(x + int(t % 5) - 2, 0), int(a.X) - 1);
        const int ny = min(max(y + int((t / 5) % 5) - 2, 0), int(a.Y) - 1);
        const int nz = min(max(z + int(t / 25) - 2, 0), int(a.Z) - 1);
        const float px = voxel_center(a.hmin[0], a.vs[0], uint32_t(nx));
        const float py = voxel_center(a.hmin[1], a.vs[1], uint32_t(nz));
        const float pz = voxel_center(a.hmin[2], a.vs[2], uint32_t(ny));
        const uint32_t l0 = chunk * chunk_len, l1 = min(a.L, l0 + chunk_len);
        for (uint32_t l = l0; l < l1; ++l) acc += point_sample(a, l, px, py, pz);
        acc *= double(w.w[t]);
    }
    red[t] = acc;
    __syncthreads();
    for (int s = 64; s > 0; s >>= 1) {
        if (int(t) < s) red[t] += red[t + s];
        __syncthreads();
    }
    if (t == 0) partial[size_t(ci) * gridDim.y + chunk] = red[0];
}

// Least-squares sums of reference (r) on volume (v): per block, in FP64:
// sums[block][0..4] = (n, sum v, sum v^2, sum r, sum v r).
constexpr int kFitBlocks = 256;
__global__ __launch_bounds__(256) void calib_sums_kernel(const float* __restrict__ v,
                                                         const float* __restrict__ r, size_t n,
                                                         double* __restrict__ sums) {
    __shared__ double red[256][5];
    double a[5] = {0, 0, 0, 0, 0};
    for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
         i += size_t(gridDim.x) * blockDim.x) {
        const double x = v[i], y = r[i];
        a[0] += 1.0;
        a[1] += x;
        a[2] += x * x;
        a[3] += y;
        a[4] += x * y;
    }
    for (int q = 0; q < 5; ++q) red[threadIdx.x][q] = a[q];
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (int(threadIdx.x) < s)
            for (int q = 0; q < 5; ++q) red[threadIdx.x][q] += red[threadIdx.x + s][q];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int q = 0; q < 5; ++q) sums[blockIdx.x * 5 + q] = red[0][q];
}

// One block: L ~ a v + b from the block sums, the exact LoG of each
// candidate from its chunk partials, the true range over the candidates;
// coef = (a / range, (b - min) / range) so that out = coef0 v + coef1.
__global__ __launch_bounds__(256) void calib_fit_kernel(const double* __restrict__ sums,
                                                        const double* __restrict__ partial,
                                                        uint32_t chunks, uint32_t k,
                                                        float* __restrict__ coef) {
    __shared__ double red[256][5];
    __shared__ double lval[2 * kMaxExtremes];
    const uint32_t t = threadIdx.x;
    for (int q = 0; q < 5; ++q) red[t][q] = t < uint32_t(kFitBlocks) ? sums[t * 5 + q] : 0.0;
    if (t < 2 * k) {
        double l = 0.0;
        for (uint32_t c = 0; c < chunks; ++c) l += partial[size_t(t) * chunks + c];
        lval[t] = l;
    }
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (int(t) < s)
            for (int q = 0; q < 5; ++q) red[t][q] += red[t + s][q];
        __syncthreads();
    }
    if (t != 0) return;
    const double* S = red[0];
    const double det = S[0] * S[2] - S[1] * S[1];
    const double a = det > 0.0 ? (S[0] * S[4] - S[1] * S[3]) / det : 0.0;
    const double b = S[0] > 0.0 ? (S[3] - a * S[1]) / S[0] : 0.0;
    double mn = lval[0], mx = lval[k];
    for (uint32_t i = 1; i < k; ++i) mn = fmin(mn, lval[i]);
    for (uint32_t i = k + 1; i < 2 * k; ++i) mx = fmax(mx, lval[i]);
    const double range = mx - mn;
    // safeRCP of the range, as the min-max normalisation (FB/Laser.cu:317-339)
    const double inv = (range > 1e-5 || range < -1e-5) ? 1.0 / range : 3.402823466e+38;
    coef[0] = float(a * inv);
    coef[1] = float((b - mn) * inv);
}

__global__ void calib_apply_kernel(float* __restrict__ v, size_t n, const float* __restrict__ coef) {
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) v[i] = fmaf(coef[0], v[i], coef[1]);
}

PointArgs point_args(const fbp_reconstruction_info& info, const float* I, const fbp_vec3* laser) {
    PointArgs a{};
    a.I = I;
    a.laser = laser;
    a.L = info.num_laser_targets;
    a.T = info.num_time_bins;
    a.X = info.voxel_resolution.x;
    a.Y = info.voxel_resolution.y;
    a.Z = info.voxel_resolution.z;
    a.hmin[0] = info.hidden_volume_min.x;
    a.hmin[1] = info.hidden_volume_min.y;
    a.hmin[2] = info.hidden_volume_min.z;
    a.vs[0] = info.hidden_volume_voxel_size.x;
    a.vs[1] = info.hidden_volume_voxel_size.y;
    a.vs[2] = info.hidden_volume_voxel_size.z;
    a.t0 = info.time_offset;
    a.dt = info.time_step;
    a.discard = info.discard_first_last_bounces;
    a.lpos = info.laser_position;
    a.spos = info.sensor_position;
    return a;
}

fbp_status check_confocal_f32(fbp_context ctx, const fbp_reconstruction_info* info,
                              const fbp_transient* tr, const char* what) {
    if (info->num_laser_targets == 0 || info->num_time_bins == 0 || info->voxel_resolution.x == 0 ||
        info->voxel_resolution.y == 0 || info->voxel_resolution.z == 0)
        return fail(ctx, FBP_ERR_INVALID_ARGUMENT, "%s: zero-sized grid", what);
    if (!(info->time_step > 0.f) || !finite_info(*info))
        return fail(ctx, FBP_ERR_INVALID_ARGUMENT, "%s: time_step must be > 0, fields finite", what);
    if (tr->storage != FBP_STORAGE_F32 || tr->layout != FBP_LAYOUT_TIME_FASTEST)
        return fail(ctx, FBP_ERR_UNSUPPORTED, "%s: F32 time-fastest transients only", what);
    return FBP_OK;
}

size_t align256(size_t b) { return (b + 255) & ~size_t(255); }

}  // namespace

}  // namespace fbp

using namespace fbp;

extern "C" fbp_status fbp_backproject_confocal_points(fbp_context ctx,
                                                      const fbp_reconstruction_info* info,
                                                      const fbp_transient* tr,
                                                      const fbp_vec3* laser_targets,
                                                      const uint32_t* voxels, uint32_t count,
                                                      double* values, void* stream) {
    if (!ctx || !info || !tr || !tr->data || !laser_targets || (count && (!voxels || !values)))
        return fail(ctx, FBP_ERR_INVALID_ARGUMENT, "points: null argument");
    FBP_CHECK(check_confocal_f32(ctx, info, tr, "points"));
    FBP_CHECK(bind(ctx));
    if (count == 0) return FBP_OK;
    const PointArgs a = point_args(*info, static_cast<const float*>(tr->data), laser_targets);
    hipLaunchKernelGGL(bp_points_kernel, dim3((count + 3) / 4), dim3(256), 0, as_stream(stream), a,
                       voxels, count, values);
    FBP_CHECK_HIP(ctx, hipGetLastError());
    return FBP_OK;
}

extern "C" fbp_status fbp_learned_calibrate(fbp_context ctx, const fbp_reconstruction_info* info,
                                            const fbp_transient* tr,
                                            const fbp_vec3* laser_targets, float* volume,
                                            const float* reference, int32_t kernel_size,
                                            float sigma, uint32_t extremes, void* stream) {
    if (!ctx || !info || !tr || !tr->data || !laser_targets || !volume || !reference)
        return fail(ctx, FBP_ERR_INVALID_ARGUMENT, "calibrate: null argument");
    FBP_CHECK(check_confocal_f32(ctx, info, tr, "calibrate"));
    if (kernel_size != 5)
        return fail(ctx, FBP_ERR_UNSUPPORTED, "calibrate: LoG kernel size 5 only");
    if (extremes == 0 || extremes > uint32_t(kMaxExtremes))
        return fail(ctx, FBP_ERR_INVALID_ARGUMENT, "calibrate: 1 <= extremes <= %d", kMaxExtremes);
    FBP_CHECK(bind(ctx));
    hipStream_t s = as_stream(stream);

    Weights w{};
    FBP_CHECK(fbp_log_kernel_host(kernel_size, sigma, w.w));
    const uint32_t X = info->voxel_resolution.x, Y = info->voxel_resolution.y,
                   Z = info->voxel_resolution.z;
    const size_t n = size_t(X) * Y * Z;
    if (n < extremes || n >= (size_t(1) << 31))
        return fail(ctx, FBP_ERR_UNSUPPORTED, "calibrate: volume size out of range");
    const uint32_t k = extremes;

    // Scan-point chunks per candidate: enough blocks to fill the device.
    const uint32_t L = info->num_laser_targets;
    const uint32_t target = 8u * uint32_t(ctx->num_cus > 0 ? ctx->num_cus : 256);
    uint32_t chunks = std::max(1u, std::min(L, (target + 2 * k - 1) / (2 * k)));
    const uint32_t chunk_len = (L + chunks - 1) / chunks;
    chunks = (L + chunk_len - 1) / chunk_len;

    const uint32_t w0 = uint32_t((n + 255) / 256);  // level-0 waves (4 entries per lane)
    const size_t cand = size_t(w0) * k;
    const size_t bytes = 4 * 2 * align256(cand * 4) + align256(size_t(2) * k * chunks * 8) +
                         align256(size_t(kFitBlocks) * 5 * 8) + 256;
    FBP_CHECK(ensure_scratch(ctx, bytes));
    char* cur = static_cast<char*>(ctx->scratch);
    auto take = [&](size_t b) {
        char* p = cur;
        cur += align256(b);
        return p;
    };
    float* lo_v[2] = {reinterpret_cast<float*>(take(cand * 4)), reinterpret_cast<float*>(take(cand * 4))};
    uint32_t* lo_i[2] = {reinterpret_cast<uint32_t*>(take(cand * 4)),
                         reinterpret_cast<uint32_t*>(take(cand * 4))};
    float* hi_v[2] = {reinterpret_cast<float*>(take(cand * 4)), reinterpret_cast<float*>(take(cand * 4))};
    uint32_t* hi_i[2] = {reinterpret_cast<uint32_t*>(take(cand * 4)),
                         reinterpret_cast<uint32_t*>(take(cand * 4))};
    double* partial = reinterpret_cast<double*>(take(size_t(2) * k * chunks * 8));
    double* sums = reinterpret_cast<double*>(take(size_t(kFitBlocks) * 5 * 8));
    float* coef = reinterpret_cast<float*>(take(256));

    // Extreme candidates: per-wave selection (4 entries per lane: short
    // rounds, many waves), repeated until one wave holds the k of each side.
    auto blocks = [](uint32_t waves) { return dim3((waves + 3) / 4, 2); };
    hipLaunchKernelGGL(wave_extremes_kernel<4>, blocks(w0), dim3(256), 0, s, volume, nullptr,
                       volume, nullptr, uint32_t(n), k, lo_v[0], lo_i[0], hi_v[0], hi_i[0]);
    uint32_t m = w0 * k;
    int cb = 0;
    while (m > k) {
        const uint32_t wv = (m + 255) / 256;
        hipLaunchKernelGGL(wave_extremes_kernel<4>, blocks(wv), dim3(256), 0, s, lo_v[cb],
                           lo_i[cb], hi_v[cb], hi_i[cb], m, k, lo_v[cb ^ 1], lo_i[cb ^ 1],
                           hi_v[cb ^ 1], hi_i[cb ^ 1]);
        m = wv * k;
        cb ^= 1;
    }

    const PointArgs a = point_args(*info, static_cast<const float*>(tr->data), laser_targets);
    hipLaunchKernelGGL(calib_neigh_kernel, dim3(2 * k, chunks), dim3(128), 0, s, a, w, lo_i[cb],
                       hi_i[cb], k, chunk_len, partial);
    hipLaunchKernelGGL(calib_sums_kernel, dim3(kFitBlocks), dim3(256), 0, s, volume, reference, n,
                       sums);
    hipLaunchKernelGGL(calib_fit_kernel, dim3(1), dim3(256), 0, s, sums, partial, chunks, k, coef);
    hipLaunchKernelGGL(calib_apply_kernel, dim3(unsigned((n + 255) / 256)), dim3(256), 0, s, volume,
                       n, coef);
    FBP_CHECK_HIP(ctx, hipGetLastError());
    return FBP_OK;
}
