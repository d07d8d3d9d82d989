// Post-processing filters on [Z][Y][X] float volumes.
//
// Laplacian (FB/PostprocessingFilters.cu:81-98, FB/transient_postprocessing.cuh:21-50):
//   out = k^3 v - sum_{k^3 box, in-bounds only} v
// computed as three 1D box passes (zero outside the volume) fused with the
// final subtraction.
//
// LoG (FB/PostprocessingFilters.cu:12-32, 102-175; transient_postprocessing.cuh:52-83):
// the reference convolves with a dense k^3 kernel, clamp-to-edge per axis.
// Its kernel is w = G(x)G(y)G(z) (r^2 - 3 s^2) / s^4 - c with G the
// normalised Gaussian, so when no Gaussian tap is thresholded away it
// factors exactly into rank-1 terms:
//   LoG(v) = (Hx Gy Gz + Gx Hy Gz + Gx Gy Hz)(v) / s^4 - 3/s^2 Gx Gy Gz (v) - c Bx By Bz (v)
// with g = e/sum(e), h = x^2 g, b = 1 (clamp-to-edge 1D passes).  That is
// three HBM passes instead of k^3 gathers per voxel; otherwise the dense path
// runs.
#include <cfloat>
#include <cmath>
#include <vector>

#include "chain.h"
#include "fbp_internal.h"

namespace fbp {

constexpr int kMaxTaps = 15;  // kernel_size <= 15

struct Taps {
    int half;
    float g[kMaxTaps], h[kMaxTaps];
};

// Index along `axis` of element v, and the stride of that axis.
struct AxisInfo {
    uint32_t X, Y, Z;
};

__device__ __forceinline__ void decompose(size_t v, const AxisInfo& a, uint32_t& x, uint32_t& y,
                                          uint32_t& z) {
    x = uint32_t(v % a.X);
    y = uint32_t((v / a.X) % a.Y);
    z = uint32_t(v / (size_t(a.X) * a.Y));
}

template <int AXIS>
__device__ __forceinline__ size_t shifted(uint32_t x, uint32_t y, uint32_t z, int c,
                                          const AxisInfo& a) {
    if (AXIS == 0) x = uint32_t(c);
    if (AXIS == 1) y = uint32_t(c);
    if (AXIS == 2) z = uint32_t(c);
    return (size_t(z) * a.Y + y) * a.X + x;
}

template <int AXIS>
__device__ __forceinline__ uint32_t coord(uint32_t x, uint32_t y, uint32_t z) {
    return AXIS == 0 ? x : (AXIS == 1 ? y : z);
}
template <int AXIS>
__device__ __forceinline__ uint32_t extent(const AxisInfo& a) {
    return AXIS == 0 ? a.X : (AXIS == 1 ? a.Y : a.Z);
}

// ---- Laplacian: zero-outside box sums; final pass subtracts from k^3 v.
template <int AXIS>
__global__ void box_pass_kernel(const float* __restrict__ in, float* __restrict__ out, AxisInfo a,
                                int half) {
    const size_t n = size_t(a.X) * a.Y * a.Z;
    const size_t v = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (v >= n) return;
    uint32_t x, y, z;
    decompose(v, a, x, y, z);
    const int c = int(coord<AXIS>(x, y, z)), e = int(extent<AXIS>(a));
    float s = 0.f;
    for (int d = -half; d <= half; ++d) {
        const int cc = c + d;
        if (cc < 0 || cc >= e) continue;
        s += in[shifted<AXIS>(x, y, z, cc, a)];
    }
    out[v] = s;
}

__global__ void laplacian_final_kernel(const float* __restrict__ vol, const float* __restrict__ bxy,
                                       float* __restrict__ out, AxisInfo a, int half, float k3) {
    const size_t n = size_t(a.X) * a.Y * a.Z;
    const size_t v = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (v >= n) return;
    uint32_t x, y, z;
    decompose(v, a, x, y, z);
    float s = 0.f;
    for (int d = -half; d <= half; ++d) {
        const int cc = int(z) + d;
        if (cc < 0 || cc >= int(a.Z)) continue;
        s += bxy[shifted<2>(x, y, z, cc, a)];
    }
    out[v] = k3 * vol[v] - s;
}

// ---- LoG, separable form.
// Pass X: in -> (Gx v, Hx v, Bx v)
__global__ void log_pass_x(const float* __restrict__ in, float* __restrict__ g, float* __restrict__ h,
                           float* __restrict__ b, AxisInfo a, Taps t) {
    const size_t n = size_t(a.X) * a.Y * a.Z;
    const size_t v = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (v >= n) return;
    uint32_t x, y, z;
    decompose(v, a, x, y, z);
    float sg = 0.f, sh = 0.f, sb = 0.f;
    for (int d = -t.half; d <= t.half; ++d) {
        const int cc = min(max(int(x) + d, 0), int(a.X) - 1);
        const float val = in[shifted<0>(x, y, z, cc, a)];
        sg = fmaf(t.g[d + t.half], val, sg);
        sh = fmaf(t.h[d + t.half], val, sh);
        sb += val;
    }
    g[v] = sg;
    h[v] = sh;
    b[v] = sb;
}

// Pass Y: (gx, hx, bx) -> (Gy gx, Gy hx, Hy gx, By bx)
__global__ void log_pass_y(const float* __restrict__ gx, const float* __restrict__ hx,
                           const float* __restrict__ bx, float* __restrict__ gg,
                           float* __restrict__ hg, float* __restrict__ gh, float* __restrict__ bb,
                           AxisInfo a, Taps t) {
    const size_t n = size_t(a.X) * a.Y * a.Z;
    const size_t v = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (v >= n) return;
    uint32_t x, y, z;
    decompose(v, a, x, y, z);
    float s_gg = 0.f, s_hg = 0.f, s_gh = 0.f, s_bb = 0.f;
    for (int d = -t.half; d <= t.half; ++d) {
        const int cc = min(max(int(y) + d, 0), int(a.Y) - 1);
        const size_t i = shifted<1>(x, y, z, cc, a);
        const float vg = gx[i], vh = hx[i];
        s_gg = fmaf(t.g[d + t.half], vg, s_gg);
        s_hg = fmaf(t.g[d + t.half], vh, s_hg);
        s_gh = fmaf(t.h[d + t.half], vg, s_gh);
        s_bb += bx[i];
    }
    gg[v] = s_gg;
    hg[v] = s_hg;
    gh[v] = s_gh;
    bb[v] = s_bb;
}

// Pass Z: out = (Gz hg + Gz gh + Hz gg) * a4 - Gz gg * a2 - c * Bz bb
__global__ void log_pass_z(const float* __restrict__ gg, const float* __restrict__ hg,
                           const float* __restrict__ gh, const float* __restrict__ bb,
                           float* __restrict__ out, AxisInfo a, Taps t, float a4, float a2,
                           float c) {
    const size_t n = size_t(a.X) * a.Y * a.Z;
    const size_t v = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (v >= n) return;
    uint32_t x, y, z;
    decompose(v, a, x, y, z);
    float s_r2 = 0.f, s_g = 0.f, s_b = 0.f;
    for (int d = -t.half; d <= t.half; ++d) {
        const int cc = min(max(int(z) + d, 0), int(a.Z) - 1);
        const size_t i = shifted<2>(x, y, z, cc, a);
        const float gz = t.g[d + t.half], hz = t.h[d + t.half];
        s_r2 = fmaf(gz, hg[i] + gh[i], fmaf(hz, gg[i], s_r2));
        s_g = fmaf(gz, gg[i], s_g);
        s_b += bb[i];
    }
    out[v] = (s_r2 * a4 - s_g * a2) - c * s_b;
}

// ---- LoG, dense form (kernel with thresholded taps): the reference's loop.
__global__ void log_dense_kernel(const float* __restrict__ in, float* __restrict__ out,
                                 const float* __restrict__ w, AxisInfo a, int k, int half) {
    const size_t n = size_t(a.X) * a.Y * a.Z;
    const size_t v = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (v >= n) return;
    uint32_t x, y, z;
    decompose(v, a, x, y, z);
    float sum = 0.f;
    for (int kz = -half; kz <= half; ++kz) {
        const int iz = min(max(int(z) + kz, 0), int(a.Z) - 1);
        for (int ky = -half; ky <= half; ++ky) {
            const int iy = min(max(int(y) + ky, 0), int(a.Y) - 1);
            for (int kx = -half; kx <= half; ++kx) {
                const int ix = min(max(int(x) + kx, 0), int(a.X) - 1);
                sum += in[(size_t(iz) * a.Y + iy) * a.X + ix] *
                       w[(kz + half) * k * k + (ky + half) * k + (kx + half)];
            }
        }
    }
    out[v] = sum;
}

// ---- min / max and normalisation (deterministic two-stage reductions).
constexpr int kRedBlocks = 1024;

__global__ void minmax_partial_kernel(const float* __restrict__ v, size_t n,
                                      float* __restrict__ part) {
    float mn = FLT_MAX, mx = -FLT_MAX;
    for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
         i += size_t(gridDim.x) * blockDim.x) {
        const float x = v[i];
        mn = fminf(mn, x);
        mx = fmaxf(mx, x);
    }
    __shared__ float smn[256], smx[256];
    smn[threadIdx.x] = mn;
    smx[threadIdx.x] = mx;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if (int(threadIdx.x) < w) {
            smn[threadIdx.x] = fminf(smn[threadIdx.x], smn[threadIdx.x + w]);
            smx[threadIdx.x] = fmaxf(smx[threadIdx.x], smx[threadIdx.x + w]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = smn[0];
        part[2 * blockIdx.x + 1] = smx[0];
    }
}

__global__ void minmax_final_kernel(float* __restrict__ part, int nparts) {
    float mn = FLT_MAX, mx = -FLT_MAX;
    for (int i = threadIdx.x; i < nparts; i += blockDim.x) {
        mn = fminf(mn, part[2 * i]);
        mx = fmaxf(mx, part[2 * i + 1]);
    }
    __shared__ float smn[256], smx[256];
    smn[threadIdx.x] = mn;
    smx[threadIdx.x] = mx;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if (int(threadIdx.x) < w) {
            smn[threadIdx.x] = fminf(smn[threadIdx.x], smn[threadIdx.x + w]);
            smx[threadIdx.x] = fmaxf(smx[threadIdx.x], smx[threadIdx.x + w]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[2 * kRedBlocks] = smn[0];
        part[2 * kRedBlocks + 1] = smx[0];
    }
}

// v = (v - min) * safeRCP(max - min)  (FB/transient_postprocessing.cuh:10-17)
__global__ void normalize_apply_kernel(float* __restrict__ v, size_t n, const float* __restrict__ mm) {
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float mn = mm[0], mx = mm[1];
    v[i] = rn_mul(rn_sub(v[i], mn), safe_rcp(rn_sub(mx, mn)));
}

static fbp_status run_minmax(fbp_context ctx, const float* v, size_t n, hipStream_t s,
                             float** mm_out) {
    FBP_CHECK(ensure_scratch(ctx, (2 * kRedBlocks + 2) * sizeof(float)));
    float* part = static_cast<float*>(ctx->scratch);
    size_t blocks = (n + 255) / 256;
    if (blocks > kRedBlocks) blocks = kRedBlocks;
    hipLaunchKernelGGL(minmax_partial_kernel, dim3(unsigned(blocks)), dim3(256), 0, s, v, n, part);
    hipLaunchKernelGGL(minmax_final_kernel, dim3(1), dim3(256), 0, s, part, int(blocks));
    FBP_CHECK_HIP(ctx, hipGetLastError());
    *mm_out = part + 2 * kRedBlocks;
    return FBP_OK;
}

static dim3 grid_for(size_t n) { return dim3(unsigned((n + 255) / 256)); }

}  // namespace fbp

using namespace fbp;

extern "C" {

fbp_status fbp_log_kernel_host(int32_t size, float std1, float* out) {
    // LoG::calculateLaplacianKernel (FB/PostprocessingFilters.cu:102-175),
    // same float operations in the same order.
    if (!out || size < 1 || size > kMaxTaps) return FBP_ERR_INVALID_ARGUMENT;
    const int lim = (size - 1) / 2;
    const float std2 = std1 * std1;
    const int n = size * size * size;
    std::vector<float> xs(n), ys(n), zs(n), w(n);
    float maxw = 0.0f;
    for (int kz = 0; kz < size; ++kz)
        for (int ky = 0; ky < size; ++ky)
            for (int kx = 0; kx < size; ++kx) {
                const float x = float(kx - lim), y = float(ky - lim), z = float(kz - lim);
                const int idx = kz * size * size + ky * size + kx;
                xs[idx] = x;
                ys[idx] = y;
                zs[idx] = z;
                w[idx] = expf(-(x * x + y * y + z * z) / (2.0f * std2));
                maxw = std::max(w[idx], maxw);
            }
    const float thr = FLT_EPSILON * maxw;
    float sum_w = 0.0f;
    for (float& v : w) {
        if (v < thr) v = 0.0f;
        sum_w += v;
    }
    if (sum_w != 0.0f)
        for (float& v : w) v /= sum_w;
    float sum_w1 = 0.0f;
    std::vector<float> w1(n);
    for (int i = 0; i < n; ++i) {
        const float x = xs[i], y = ys[i], z = zs[i];
        w1[i] = w[i] * (x * x + y * y + z * z - 3.0f * std2) / (std2 * std2);
        sum_w1 += w1[i];
    }
    const float adj = sum_w1 / float(n);
    for (int i = 0; i < n; ++i) out[i] = w1[i] - adj;
    return FBP_OK;
}

fbp_status fbp_filter_laplacian(fbp_context ctx, const float* in, float* out, fbp_uvec3 size,
                                int32_t kernel_size, void* stream) {
    if (!ctx || !in || !out || size.x == 0 || size.y == 0 || size.z == 0 || kernel_size < 1)
        return fail(ctx, FBP_ERR_INVALID_ARGUMENT, "laplacian: bad argument");
    if (in == out) return fail(ctx, FBP_ERR_INVALID_ARGUMENT, "laplacian: in and out must differ");
    FBP_CHECK(bind(ctx));
    const AxisInfo a{size.x, size.y, size.z};
    const size_t n = size_t(size.x) * size.y * size.z;
    FBP_CHECK(ensure_scratch(ctx, 2 * n * sizeof(float)));
    float* t0 = static_cast<float*>(ctx->scratch);
    float* t1 = t0 + n;
    // halfFilterSize = floor(filterSize / 2); multiplier filterSize^3 in float.
    const float fk = float(kernel_size);
    const int half = int(std::floor(fk / 2.0f));
    const float k3 = fk * fk * fk;
    hipStream_t s = as_stream(stream);
    hipLaunchKernelGGL(box_pass_kernel<0>, grid_for(n), dim3(256), 0, s, in, t0, a, half);
    hipLaunchKernelGGL(box_pass_kernel<1>, grid_for(n), dim3(256), 0, s, t0, t1, a, half);
    hipLaunchKernelGGL(laplacian_final_kernel, grid_for(n), dim3(256), 0, s, in, t1, out, a, half, k3);
    FBP_CHECK_HIP(ctx, hipGetLastError());
    return FBP_OK;
}

fbp_status fbp_filter_log(fbp_context ctx, const float* in, float* out, fbp_uvec3 size,
                          int32_t kernel_size, float sigma, void* stream) {
    if (!ctx || !in || !out || size.x == 0 || size.y == 0 || size.z == 0)
        return fail(ctx, FBP_ERR_INVALID_ARGUMENT, "log: bad argument");
    if (in == out) return fail(ctx, FBP_ERR_INVALID_ARGUMENT, "log: in and out must differ");
    if (kernel_size < 1 || kernel_size > kMaxTaps || !(sigma > 0.f))
        return fail(ctx, FBP_ERR_UNSUPPORTED, "log: kernel_size in [1, %d], sigma > 0", kMaxTaps);
    FBP_CHECK(bind(ctx));
    const AxisInfo a{size.x, size.y, size.z};
    const size_t n = size_t(size.x) * size.y * size.z;
    hipStream_t s = as_stream(stream);
    const int k = kernel_size, half = (k - 1) / 2;
    // Separable iff no Gaussian tap falls below FLT_EPSILON * max (the
    // corner tap is the smallest), and the kernel is odd-sized.
    const float std2 = sigma * sigma;
    const float corner = float(half);
    const float wc = expf(-(3.0f * corner * corner) / (2.0f * std2));
    const bool separable = (k % 2 == 1) && !(wc < FLT_EPSILON * 1.0f);
    if (separable) {
        Taps t{};
        t.half = half;
        double e[kMaxTaps], se = 0.0;
        for (int d = -half; d <= half; ++d) {
            e[d + half] = std::exp(-double(d) * d / (2.0 * double(std2)));
            se += e[d + half];
        }
        for (int d = -half; d <= half; ++d) {
            t.g[d + half] = float(e[d + half] / se);
            t.h[d + half] = float(double(d) * d * e[d + half] / se);
        }
        // c = mean of the un-centred kernel w1 = ghat^3 (r^2 - 3 s^2) / s^4
        //   = 3 (m2 - s^2) / (s^4 k^3),  m2 = sum_x x^2 ghat(x).
        double m2 = 0.0;
        for (int d = -half; d <= half; ++d) m2 += double(d) * d * e[d + half] / se;
        const double s2 = double(std2);
        const float c = float(3.0 * (m2 - s2) / (s2 * s2 * double(k) * k * k));
        FBP_CHECK(ensure_scratch(ctx, 7 * n * sizeof(float)));
        float* buf = static_cast<float*>(ctx->scratch);
        float *gx = buf, *hx = buf + n, *bx = buf + 2 * n;
        float *gg = buf + 3 * n, *hg = buf + 4 * n, *gh = buf + 5 * n, *bb = buf + 6 * n;
        hipLaunchKernelGGL(log_pass_x, grid_for(n), dim3(256), 0, s, in, gx, hx, bx, a, t);
        hipLaunchKernelGGL(log_pass_y, grid_for(n), dim3(256), 0, s, gx, hx, bx, gg, hg, gh, bb, a, t);
        const float a4 = float(1.0 / (double(std2) * std2));
        const float a2 = float(3.0 / double(std2));
        hipLaunchKernelGGL(log_pass_z, grid_for(n), dim3(256), 0, s, gg, hg, gh, bb, out, a, t, a4,
                           a2, c);
    } else {
        std::vector<float> w(size_t(k) * k * k);
        fbp_log_kernel_host(k, sigma, w.data());
        FBP_CHECK(ensure_scratch(ctx, w.size() * sizeof(float)));
        float* dw = static_cast<float*>(ctx->scratch);
        FBP_CHECK_HIP(ctx, hipMemcpyAsync(dw, w.data(), w.size() * sizeof(float),
                                          hipMemcpyHostToDevice, s));
        hipLaunchKernelGGL(log_dense_kernel, grid_for(n), dim3(256), 0, s, in, out, dw, a, k, half);
        // w is a host vector consumed by the async copy.
        FBP_CHECK_HIP(ctx, hipStreamSynchronize(s));
    }
    FBP_CHECK_HIP(ctx, hipGetLastError());
    return FBP_OK;
}

fbp_status fbp_normalize_minmax(fbp_context ctx, float* v, uint64_t n, void* stream) {
    if (!ctx || !v || n == 0) return fail(ctx, FBP_ERR_INVALID_ARGUMENT, "normalize: bad argument");
    FBP_CHECK(bind(ctx));
    hipStream_t s = as_stream(stream);
    float* mm = nullptr;
    FBP_CHECK(run_minmax(ctx, v, n, s, &mm));
    hipLaunchKernelGGL(normalize_apply_kernel, grid_for(n), dim3(256), 0, s, v, n, mm);
    FBP_CHECK_HIP(ctx, hipGetLastError());
    return FBP_OK;
}

fbp_status fbp_minmax(fbp_context ctx, const float* v, uint64_t n, float* host_min,
                      float* host_max, void* stream) {
    if (!ctx || !v || n == 0 || !host_min || !host_max)
        return fail(ctx, FBP_ERR_INVALID_ARGUMENT, "minmax: bad argument");
    FBP_CHECK(bind(ctx));
    hipStream_t s = as_stream(stream);
    float* mm = nullptr;
    FBP_CHECK(run_minmax(ctx, v, n, s, &mm));
    FBP_CHECK_HIP(ctx, hipMemcpyAsync(ctx->pinned, mm, 2 * sizeof(float), hipMemcpyDeviceToHost, s));
    FBP_CHECK_HIP(ctx, hipStreamSynchronize(s));
    *host_min = ctx->pinned[0];
    *host_max = ctx->pinned[1];
    return FBP_OK;
}

}  // extern "C"
