// Post-processing filters on [Z][Y][X] float volumes.
//
// Laplacian (FB/PostprocessingFilters.cu:81-98, FB/transient_postprocessing.cuh:21-50):
//   out = k^3 v - sum_{k^3 box, in-bounds only} v
// computed as 1D box sums along x, y and z (zero outside the volume) and the
// final subtraction, in one pass over the volume (boxes up to 15 wide;
// wider boxes take three passes through scratch).
//
// LoG (FB/PostprocessingFilters.cu:12-32, 102-175; transient_postprocessing.cuh:52-83):
// the reference convolves with a dense k^3 kernel, clamp-to-edge per axis.
// Its kernel is w = G(x)G(y)G(z) (r^2 - 3 s^2) / s^4 - c with G the
// normalised Gaussian, so when no Gaussian tap is thresholded away it
// factors exactly into rank-1 terms:
//   LoG(v) = (Hx Gy Gz + Gx Hy Gz + Gx Gy Hz)(v) / s^4 - 3/s^2 Gx Gy Gz (v) - c Bx By Bz (v)
// with g = e/sum(e), h = x^2 g, b = 1 (clamp-to-edge 1D passes), evaluated
// in one pass over the volume (x and y passes in LDS per z-plane, the z pass
// over a register ring of planes) instead of k^3 gathers per voxel;
// otherwise the dense path runs.
#include <algorithm>
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

// ---- Fused forms: one workgroup owns a 16 x 16 (x, y) tile and walks a
// chunk of z-planes.  Each input plane is read once into LDS with its x/y
// halo; the x and y passes run in LDS, the z pass over a register ring of
// the last 2H + 1 planes' y-pass results.  Every output takes the same
// operations in the same order as a separable evaluation in three passes
// over the whole volume (x, then y, then z), with one read of the volume
// (plus halos from L2) and one write instead of the intermediate volumes.
constexpr int kFT = 16;  // tile edge (x and y); 256 threads

// Optional epilogue of the fused LoG: the min and max of the output, one
// pair per block in part[2 b], part[2 b + 1] (reduced by the consumer:
// block_minmax in fbp_internal.h).
template <int H, bool MM>
__global__ __launch_bounds__(256) void log_fused_kernel(const float* __restrict__ in,
                                                        float* __restrict__ out, AxisInfo a, Taps t,
                                                        float a4, float a2, float c, int zchunk,
                                                        float* __restrict__ mm_part) {
    constexpr int K = 2 * H + 1, PW = kFT + 2 * H, PH = kFT + 2 * H;
    __shared__ float pl[PH][PW + 1];
    __shared__ float sg[PH][kFT], sh[PH][kFT], sb[PH][kFT];
    const int tid = threadIdx.x, tx = tid & (kFT - 1), ty = tid / kFT;
    const int X = int(a.X), Y = int(a.Y), Z = int(a.Z);
    const int x0 = int(blockIdx.x) * kFT, y0 = int(blockIdx.y) * kFT;
    const int zb = int(blockIdx.z) * zchunk, ze = min(Z, zb + zchunk);
    float tg[K], th[K];
#pragma unroll
    for (int d = 0; d < K; ++d) {
        tg[d] = t.g[d];
        th[d] = t.h[d];
    }
    float lmn = FLT_MAX, lmx = -FLT_MAX;
    float rgg[K], rhg[K], rgh[K], rbb[K];
#pragma unroll
    for (int d = 0; d < K; ++d) rgg[d] = rhg[d] = rgh[d] = rbb[d] = 0.f;
    // Plane clamp(z) with clamp-to-edge x/y halo, loaded into registers one
    // plane ahead of its use.
    constexpr int NPF = (PH * PW + 255) / 256;
    float pf[NPF];
    auto load = [&](int z) {
        const int zc = min(max(z, 0), Z - 1);
#pragma unroll
        for (int u = 0; u < NPF; ++u) {
            const int i = min(tid + u * 256, PH * PW - 1);
            const int r = i / PW, q = i - r * PW;
            const int gy = min(max(y0 + r - H, 0), Y - 1), gx = min(max(x0 + q - H, 0), X - 1);
            pf[u] = in[(size_t(zc) * Y + gy) * X + gx];
        }
    };
    load(zb - H);
    for (int zi = zb - H; zi < ze + H; ++zi) {
#pragma unroll
        for (int u = 0; u < NPF; ++u) {
            const int i = tid + u * 256;
            if (i < PH * PW) pl[i / PW][i % PW] = pf[u];
        }
        __syncthreads();
        if (zi + 1 < ze + H) load(zi + 1);
        // Pass X over every halo row.
        for (int i = tid; i < PH * kFT; i += 256) {
            const int r = i / kFT, q = i - r * kFT;
            float g = 0.f, h = 0.f, b = 0.f;
#pragma unroll
            for (int d = 0; d < K; ++d) {
                const float v = pl[r][q + d];
                g = fmaf(tg[d], v, g);
                h = fmaf(th[d], v, h);
                b += v;
            }
            sg[r][q] = g;
            sh[r][q] = h;
            sb[r][q] = b;
        }
        __syncthreads();
        // Pass Y for this thread's (x, y), pushed into the ring.
        float gg = 0.f, hg = 0.f, gh = 0.f, bb = 0.f;
#pragma unroll
        for (int d = 0; d < K; ++d) {
            const float vg = sg[ty + d][tx], vh = sh[ty + d][tx];
            gg = fmaf(tg[d], vg, gg);
            hg = fmaf(tg[d], vh, hg);
            gh = fmaf(th[d], vg, gh);
            bb += sb[ty + d][tx];
        }
#pragma unroll
        for (int d = 0; d + 1 < K; ++d) {
            rgg[d] = rgg[d + 1];
            rhg[d] = rhg[d + 1];
            rgh[d] = rgh[d + 1];
            rbb[d] = rbb[d + 1];
        }
        rgg[K - 1] = gg;
        rhg[K - 1] = hg;
        rgh[K - 1] = gh;
        rbb[K - 1] = bb;
        // Pass Z: the ring holds planes zo - H .. zo + H.
        const int zo = zi - H;
        const int ox = x0 + tx, oy = y0 + ty;
        if (zo >= zb && ox < X && oy < Y) {
            float s_r2 = 0.f, s_g = 0.f, s_b = 0.f;
#pragma unroll
            for (int d = 0; d < K; ++d) {
                s_r2 = fmaf(tg[d], rhg[d] + rgh[d], fmaf(th[d], rgg[d], s_r2));
                s_g = fmaf(tg[d], rgg[d], s_g);
                s_b += rbb[d];
            }
            const float val = (s_r2 * a4 - s_g * a2) - c * s_b;
            out[(size_t(zo) * Y + oy) * X + ox] = val;
            if constexpr (MM) {
                lmn = fminf(lmn, val);
                lmx = fmaxf(lmx, val);
            }
        }
        __syncthreads();
    }
    if constexpr (MM) {
        __shared__ float smn[256], smx[256];
        smn[tid] = lmn;
        smx[tid] = lmx;
        __syncthreads();
        for (int w = 128; w > 0; w >>= 1) {
            if (tid < w) {
                smn[tid] = fminf(smn[tid], smn[tid + w]);
                smx[tid] = fmaxf(smx[tid], smx[tid + w]);
            }
            __syncthreads();
        }
        const uint32_t bid = (blockIdx.z * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x;
        if (tid == 0) {
            mm_part[2 * bid] = smn[0];
            mm_part[2 * bid + 1] = smx[0];
        }
    }
}

// Laplacian, zero outside the volume: k^3 v - (box sum), box sums by
// x, then y, then z in order (adding the zeros of the padding leaves each
// partial sum unchanged).
template <int H>
__global__ __launch_bounds__(256) void laplacian_fused_kernel(const float* __restrict__ in,
                                                              float* __restrict__ out, AxisInfo a,
                                                              float k3, int zchunk) {
    constexpr int K = 2 * H + 1, PW = kFT + 2 * H, PH = kFT + 2 * H;
    __shared__ float pl[PH][PW + 1];
    __shared__ float sb[PH][kFT];
    const int tid = threadIdx.x, tx = tid & (kFT - 1), ty = tid / kFT;
    const int X = int(a.X), Y = int(a.Y), Z = int(a.Z);
    const int x0 = int(blockIdx.x) * kFT, y0 = int(blockIdx.y) * kFT;
    const int zb = int(blockIdx.z) * zchunk, ze = min(Z, zb + zchunk);
    float rb[K], rv[K];
#pragma unroll
    for (int d = 0; d < K; ++d) rb[d] = rv[d] = 0.f;
    // Plane z (zero outside the volume), loaded into registers one plane
    // ahead of its use; every load is issued (outside ones read element 0).
    constexpr int NPF = (PH * PW + 255) / 256;
    float pf[NPF];
    auto load = [&](int z) {
        const bool zin = z >= 0 && z < Z;
#pragma unroll
        for (int u = 0; u < NPF; ++u) {
            const int i = min(tid + u * 256, PH * PW - 1);
            const int r = i / PW, q = i - r * PW;
            const int gy = y0 + r - H, gx = x0 + q - H;
            const bool ok = zin && gy >= 0 && gy < Y && gx >= 0 && gx < X;
            const float v = in[ok ? (size_t(z) * Y + gy) * X + gx : 0];
            pf[u] = ok ? v : 0.f;
        }
    };
    load(zb - H);
    for (int zi = zb - H; zi < ze + H; ++zi) {
#pragma unroll
        for (int u = 0; u < NPF; ++u) {
            const int i = tid + u * 256;
            if (i < PH * PW) pl[i / PW][i % PW] = pf[u];
        }
        __syncthreads();
        if (zi + 1 < ze + H) load(zi + 1);
        for (int i = tid; i < PH * kFT; i += 256) {
            const int r = i / kFT, q = i - r * kFT;
            float b = 0.f;
#pragma unroll
            for (int d = 0; d < K; ++d) b += pl[r][q + d];
            sb[r][q] = b;
        }
        __syncthreads();
        float bxy = 0.f;
#pragma unroll
        for (int d = 0; d < K; ++d) bxy += sb[ty + d][tx];
#pragma unroll
        for (int d = 0; d + 1 < K; ++d) {
            rb[d] = rb[d + 1];
            rv[d] = rv[d + 1];
        }
        rb[K - 1] = bxy;
        rv[K - 1] = pl[ty + H][tx + H];
        const int zo = zi - H;
        const int ox = x0 + tx, oy = y0 + ty;
        if (zo >= zb && ox < X && oy < Y) {
            float sz = 0.f;
#pragma unroll
            for (int d = 0; d < K; ++d) sz += rb[d];
            out[(size_t(zo) * Y + oy) * X + ox] = k3 * rv[H] - sz;
        }
        __syncthreads();
    }
}

// Grid of the fused kernels: z chunks so that about 4 workgroups per CU
// run, each chunk at least 2H + 2 planes long (the halo planes are read
// again by the neighbouring chunk).
inline dim3 fused_grid(fbp_context ctx, const AxisInfo& a, int H, int& zchunk) {
    const uint32_t tiles = ((a.X + kFT - 1) / kFT) * ((a.Y + kFT - 1) / kFT);
    const uint32_t want = 4u * uint32_t(ctx->num_cus > 0 ? ctx->num_cus : 256);
    const uint32_t nz = std::max<uint32_t>(1, (want + tiles - 1) / tiles);
    zchunk = std::max<int>(2 * H + 2, int((a.Z + nz - 1) / nz));
    return dim3((a.X + kFT - 1) / kFT, (a.Y + kFT - 1) / kFT,
                unsigned((a.Z + uint32_t(zchunk) - 1) / uint32_t(zchunk)));
}

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

// v = (v - min) * safeRCP(max - min) with min / max reduced from per-block
// pairs (every block reduces them itself).
__global__ __launch_bounds__(256) void normalize_partials_kernel(float* __restrict__ v, size_t n,
                                                                 const float* __restrict__ part,
                                                                 uint32_t count) {
    float mn, mx;
    block_minmax(part, count, mn, mx);
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) v[i] = rn_mul(rn_sub(v[i], mn), safe_rcp(rn_sub(mx, mn)));
}

// LoG of in into out; with mm, also per-block min / max pairs of the output
// (mm->part in ctx->scratch, mm->count pairs), from the fused kernel's
// epilogue or, on the dense path, a min/max pass.
fbp_status log_impl(fbp_context ctx, const float* in, float* out, fbp_uvec3 size,
                    int32_t kernel_size, float sigma, hipStream_t s, MinMaxPartials* mm) {
    if (!ctx || !in || !out || size.x == 0 || size.y == 0 || size.z == 0)
        return fail(ctx, FBP_ERR_INVALID_ARGUMENT, "log: bad argument");
    if (in == out) return fail(ctx, FBP_ERR_INVALID_ARGUMENT, "log: in and out must differ");
    if (kernel_size < 1 || kernel_size > kMaxTaps || !(sigma > 0.f))
        return fail(ctx, FBP_ERR_UNSUPPORTED, "log: kernel_size in [1, %d], sigma > 0", kMaxTaps);
    FBP_CHECK(bind(ctx));
    const AxisInfo a{size.x, size.y, size.z};
    const size_t n = size_t(size.x) * size.y * size.z;
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
        const float a4 = float(1.0 / (double(std2) * std2));
        const float a2 = float(3.0 / double(std2));
        int zchunk = 0;
        const dim3 grid = fused_grid(ctx, a, half, zchunk);
        float* part = nullptr;
        if (mm) {
            mm->count = grid.x * grid.y * grid.z;
            FBP_CHECK(ensure_scratch(ctx, size_t(2) * mm->count * sizeof(float)));
            part = static_cast<float*>(ctx->scratch);
            mm->part = part;
        }
        switch (half) {  // k <= kMaxTaps = 15: half <= 7
#define FBP_LOG(HH)                                                                                \
    case HH:                                                                                       \
        if (mm)                                                                                    \
            hipLaunchKernelGGL((log_fused_kernel<HH, true>), grid, dim3(256), 0, s, in, out, a, t, \
                               a4, a2, c, zchunk, part);                                           \
        else                                                                                       \
            hipLaunchKernelGGL((log_fused_kernel<HH, false>), grid, dim3(256), 0, s, in, out, a,   \
                               t, a4, a2, c, zchunk, part);                                        \
        break;
            FBP_LOG(0) FBP_LOG(1) FBP_LOG(2) FBP_LOG(3) FBP_LOG(4) FBP_LOG(5) FBP_LOG(6) FBP_LOG(7)
#undef FBP_LOG
        }
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
        if (mm) {
            size_t blocks = (n + 255) / 256;
            if (blocks > kRedBlocks) blocks = kRedBlocks;
            FBP_CHECK(ensure_scratch(ctx, 2 * blocks * sizeof(float)));
            mm->part = static_cast<float*>(ctx->scratch);
            mm->count = uint32_t(blocks);
            hipLaunchKernelGGL(minmax_partial_kernel, dim3(unsigned(blocks)), dim3(256), 0, s, out,
                               n, mm->part);
        }
    }
    FBP_CHECK_HIP(ctx, hipGetLastError());
    return FBP_OK;
}

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
    // halfFilterSize = floor(filterSize / 2); multiplier filterSize^3 in float.
    const float fk = float(kernel_size);
    const int half = int(std::floor(fk / 2.0f));
    const float k3 = fk * fk * fk;
    hipStream_t s = as_stream(stream);
    if (half <= 7) {
        int zchunk = 0;
        const dim3 grid = fused_grid(ctx, a, half, zchunk);
        switch (half) {
#define FBP_LAP(HH)                                                                              \
    case HH:                                                                                     \
        hipLaunchKernelGGL(laplacian_fused_kernel<HH>, grid, dim3(256), 0, s, in, out, a, k3,   \
                           zchunk);                                                              \
        break;
            FBP_LAP(0) FBP_LAP(1) FBP_LAP(2) FBP_LAP(3) FBP_LAP(4) FBP_LAP(5) FBP_LAP(6) FBP_LAP(7)
#undef FBP_LAP
        }
        FBP_CHECK_HIP(ctx, hipGetLastError());
        return FBP_OK;
    }
    // Boxes wider than 15: three passes through scratch.
    FBP_CHECK(ensure_scratch(ctx, 2 * n * sizeof(float)));
    float* t0 = static_cast<float*>(ctx->scratch);
    float* t1 = t0 + n;
    hipLaunchKernelGGL(box_pass_kernel<0>, grid_for(n), dim3(256), 0, s, in, t0, a, half);
    hipLaunchKernelGGL(box_pass_kernel<1>, grid_for(n), dim3(256), 0, s, t0, t1, a, half);
    hipLaunchKernelGGL(laplacian_final_kernel, grid_for(n), dim3(256), 0, s, in, t1, out, a, half, k3);
    FBP_CHECK_HIP(ctx, hipGetLastError());
    return FBP_OK;
}

fbp_status fbp_filter_log(fbp_context ctx, const float* in, float* out, fbp_uvec3 size,
                          int32_t kernel_size, float sigma, void* stream) {
    return log_impl(ctx, in, out, size, kernel_size, sigma, as_stream(stream), nullptr);
}

fbp_status fbp_filter_log_normalize(fbp_context ctx, const float* in, float* out,
                                    fbp_uvec3 size, int32_t kernel_size, float sigma,
                                    void* stream) {
    if (!ctx || size.x == 0 || size.y == 0 || size.z == 0)
        return fail(ctx, FBP_ERR_INVALID_ARGUMENT, "log_normalize: bad argument");
    FBP_CHECK(bind(ctx));
    hipStream_t s = as_stream(stream);
    const size_t n = size_t(size.x) * size.y * size.z;
    MinMaxPartials mm{};
    FBP_CHECK(log_impl(ctx, in, out, size, kernel_size, sigma, s, &mm));
    hipLaunchKernelGGL(normalize_partials_kernel, grid_for(n), dim3(256), 0, s, out, n, mm.part,
                       mm.count);
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
