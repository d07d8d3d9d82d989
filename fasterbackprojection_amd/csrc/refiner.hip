// Learned refiner: a 3-level 3D U-Net (16/32/64 channels) that maps the
// LoG-filtered, normalised MIS-sampled backprojection to the plain
// backprojection + LoG volume (min-max normalised).  No reference
// counterpart.  x is the input divided by its largest magnitude.
//
//   e1 = relu(conv3(x; 1->16))            @ S        (VALU)
//   e2 = relu(conv3(pool(e1); 16->32))    @ S/2      (MFMA)
//   e3 = relu(conv3(pool(e2); 32->64))    @ S/4      (MFMA)
//   d2 = relu(conv3(up(e3); 64->32)) + e2 @ S/2      (MFMA)
//   d1 = relu(conv1(up(d2); 32->16)) + e1 @ S        (MFMA, 1x1x1)
//   y  = conv3(d1; 16->1) + x             @ S        (VALU)
//
// Forward convolutions with 16+ input and output channels run as BF16
// implicit GEMMs on MFMA with FP32 accumulation; activations are BF16.
// Training (backward, L1 + SSIM loss, Adam) runs on FP32 VALU kernels and
// keeps FP32 master weights.  Parameters are one flat FP32 array in the
// order w1 b1 w2 b2 ... w6 b6, weights as [tap][Cout][Cin].
#include <atomic>
#include <cmath>
#include <cstring>
#include <vector>

#include "fbp_internal.h"
#include "refiner_kernels.h"
#include "refiner_fused.h"

using namespace fbp::nn;

namespace fbp {
namespace {

constexpr int kLayers = 6;
// Cin, Cout, kernel size per layer.
constexpr int kCin[kLayers] = {1, 16, 32, 64, 32, 16};
constexpr int kCout[kLayers] = {16, 32, 64, 32, 16, 1};
constexpr int kKs[kLayers] = {3, 3, 3, 3, 1, 3};

size_t wcount(int l) { return size_t(kKs[l]) * kKs[l] * kKs[l] * kCin[l] * kCout[l]; }
size_t param_count() {
    size_t n = 0;
    for (int l = 0; l < kLayers; ++l) n += wcount(l) + kCout[l];
    return n;
}
size_t woff(int l) {
    size_t n = 0;
    for (int i = 0; i < l; ++i) n += wcount(i) + kCout[i];
    return n;
}
size_t boff(int l) { return woff(l) + wcount(l); }

inline dim3 g1(size_t n, unsigned bs = 256) { return dim3(unsigned((n + bs - 1) / bs)); }

// ------------------------------------------------------------ VALU kernels

// Generic direct convolution, FP32 math.  In: x (bf16 or f32) [V][CIN];
// out y: f32 [V][COUT] (pre-activation) and/or bf16 (after relu).
template <int CIN, int COUT, int KS, typename XT>
__global__ void conv_valu_fwd(const XT* __restrict__ x, const float* __restrict__ w,
                              const float* __restrict__ bias, float* __restrict__ pre,
                              bf16* __restrict__ out, Dims d, int relu) {
    constexpr int R = KS / 2;
    const size_t v = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (v >= d.vox()) return;
    int z, y, xx;
    unflat(v, d, z, y, xx);
    float acc[COUT];
#pragma unroll
    for (int c = 0; c < COUT; ++c) acc[c] = bias[c];
    for (int dz = -R; dz <= R; ++dz)
        for (int dy = -R; dy <= R; ++dy)
            for (int dx = -R; dx <= R; ++dx) {
                const int zz = z + dz, yv = y + dy, xv = xx + dx;
                if (zz < 0 || zz >= d.D || yv < 0 || yv >= d.H || xv < 0 || xv >= d.W) continue;
                const int tap = ((dz + R) * KS + (dy + R)) * KS + (dx + R);
                const size_t src = ((size_t(zz) * d.H + yv) * d.W + xv) * CIN;
                for (int ci = 0; ci < CIN; ++ci) {
                    float xi;
                    if constexpr (sizeof(XT) == 2)
                        xi = bf2f(x[src + ci]);
                    else
                        xi = x[src + ci];
#pragma unroll
                    for (int co = 0; co < COUT; ++co)
                        acc[co] = fmaf(w[(size_t(tap) * COUT + co) * CIN + ci], xi, acc[co]);
                }
            }
#pragma unroll
    for (int co = 0; co < COUT; ++co) {
        if (pre) pre[v * COUT + co] = acc[co];
        if (out) out[v * COUT + co] = f2bf(relu ? fmaxf(acc[co], 0.f) : acc[co]);
    }
}

// dx[v][ci] = sum_tap sum_co g[v - off(tap)][co] w[tap][co][ci]
template <int CIN, int COUT, int KS>
__global__ void conv_valu_dgrad(const float* __restrict__ g, const float* __restrict__ w,
                                float* __restrict__ dx, Dims d) {
    constexpr int R = KS / 2;
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= d.vox() * CIN) return;
    const size_t v = i / CIN;
    const int ci = int(i % CIN);
    int z, y, xx;
    unflat(v, d, z, y, xx);
    float acc = 0.f;
    for (int dz = -R; dz <= R; ++dz)
        for (int dy = -R; dy <= R; ++dy)
            for (int dx_ = -R; dx_ <= R; ++dx_) {
                const int zz = z - dz, yv = y - dy, xv = xx - dx_;
                if (zz < 0 || zz >= d.D || yv < 0 || yv >= d.H || xv < 0 || xv >= d.W) continue;
                const int tap = ((dz + R) * KS + (dy + R)) * KS + (dx_ + R);
                const float* gr = g + ((size_t(zz) * d.H + yv) * d.W + xv) * COUT;
                const float* wr = w + size_t(tap) * COUT * CIN + ci;
                for (int co = 0; co < COUT; ++co) acc = fmaf(gr[co], wr[size_t(co) * CIN], acc);
            }
    dx[i] = acc;
}

// dw[tap][co][ci] += sum_v g[v][co] x[v + off(tap)][ci]; db[co] += sum_v g[v][co].
// Block: one tap, a chunk of 64 voxels staged in LDS; FP32 atomics.
template <int CIN, int COUT, int KS, typename XT>
__global__ void conv_valu_wgrad(const float* __restrict__ g, const XT* __restrict__ x,
                                float* __restrict__ dw, float* __restrict__ db, Dims d) {
    constexpr int R = KS / 2;
    constexpr int CH = 64;
    __shared__ float sg[CH][COUT];
    __shared__ float sx[CH][CIN];
    const int tap = blockIdx.y;
    const int tz = tap / (KS * KS) - R, ty = (tap / KS) % KS - R, tx = tap % KS - R;
    const size_t v0 = size_t(blockIdx.x) * CH;
    for (int i = threadIdx.x; i < CH * COUT; i += blockDim.x) {
        const size_t v = v0 + i / COUT;
        sg[i / COUT][i % COUT] = v < d.vox() ? g[v * COUT + i % COUT] : 0.f;
    }
    for (int i = threadIdx.x; i < CH * CIN; i += blockDim.x) {
        const size_t v = v0 + i / CIN;
        float val = 0.f;
        if (v < d.vox()) {
            int z, y, xx;
            unflat(v, d, z, y, xx);
            const int zz = z + tz, yv = y + ty, xv = xx + tx;
            if (zz >= 0 && zz < d.D && yv >= 0 && yv < d.H && xv >= 0 && xv < d.W) {
                const size_t s = ((size_t(zz) * d.H + yv) * d.W + xv) * CIN + i % CIN;
                if constexpr (sizeof(XT) == 2)
                    val = bf2f(x[s]);
                else
                    val = x[s];
            }
        }
        sx[i / CIN][i % CIN] = val;
    }
    __syncthreads();
    for (int p = threadIdx.x; p < COUT * CIN; p += blockDim.x) {
        const int co = p / CIN, ci = p % CIN;
        float acc = 0.f;
        for (int k = 0; k < CH; ++k) acc = fmaf(sg[k][co], sx[k][ci], acc);
        atomicAdd(&dw[(size_t(tap) * COUT + co) * CIN + ci], acc);
    }
    if (tap == 0)
        for (int co = threadIdx.x; co < COUT; co += blockDim.x) {
            float acc = 0.f;
            for (int k = 0; k < CH; ++k) acc += sg[k][co];
            atomicAdd(&db[co], acc);
        }
}

__global__ void pool_fwd(const bf16* __restrict__ x, bf16* __restrict__ y, Dims dout, int C) {
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= dout.vox() * C) return;
    const size_t v = i / C;
    const int c = int(i % C);
    int z, yy, xx;
    unflat(v, dout, z, yy, xx);
    const int H = dout.H * 2, W = dout.W * 2;
    float s = 0.f;
    for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b)
            for (int e = 0; e < 2; ++e)
                s += bf2f(x[((size_t(2 * z + a) * H + 2 * yy + b) * W + 2 * xx + e) * C + c]);
    y[i] = f2bf(s * 0.125f);
}

// g_in (fine) += g_out (coarse parent) / 8
__global__ void pool_bwd(const float* __restrict__ gout, float* __restrict__ gin, Dims din, int C) {
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= din.vox() * C) return;
    const size_t v = i / C;
    const int c = int(i % C);
    int z, y, x;
    unflat(v, din, z, y, x);
    const size_t p = ((size_t(z / 2) * (din.H / 2) + y / 2) * (din.W / 2) + x / 2) * C + c;
    gin[i] += gout[p] * 0.125f;
}

__global__ void up_fwd(const bf16* __restrict__ x, bf16* __restrict__ y, Dims dout, int C) {
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= dout.vox() * C) return;
    const size_t v = i / C;
    const int c = int(i % C);
    int z, yy, xx;
    unflat(v, dout, z, yy, xx);
    y[i] = x[((size_t(z / 2) * (dout.H / 2) + yy / 2) * (dout.W / 2) + xx / 2) * C + c];
}

// g_coarse = sum of the 8 children
__global__ void up_bwd(const float* __restrict__ gfine, float* __restrict__ gcoarse, Dims dc, int C) {
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= dc.vox() * C) return;
    const size_t v = i / C;
    const int c = int(i % C);
    int z, y, x;
    unflat(v, dc, z, y, x);
    const int H = dc.H * 2, W = dc.W * 2;
    float s = 0.f;
    for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b)
            for (int e = 0; e < 2; ++e)
                s += gfine[((size_t(2 * z + a) * H + 2 * y + b) * W + 2 * x + e) * C + c];
    gcoarse[i] = s;
}

// g_pre = g * [pre > 0]
__global__ void relu_bwd(const float* __restrict__ g, const float* __restrict__ pre,
                         float* __restrict__ out, size_t n) {
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) out[i] = pre[i] > 0.f ? g[i] : 0.f;
}

__global__ void add_into(float* __restrict__ a, const float* __restrict__ b, size_t n) {
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) a[i] += b[i];
}

__global__ void zero_f(float* __restrict__ p, size_t n) {
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.f;
}

__global__ void to_bf16(const float* __restrict__ a, bf16* __restrict__ b, size_t n) {
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) b[i] = f2bf(a[i]);
}

// x / scale  (input normalisation by the volume's max |value|)
__global__ void scale_f(const float* __restrict__ a, float* __restrict__ b, size_t n,
                        const float* __restrict__ mm) {
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float m = fmaxf(fabsf(mm[0]), fabsf(mm[1]));
    b[i] = m > 0.f ? a[i] / m : 0.f;
}

// Largest |v|: per-block partial maxima, then one block over the partials.
constexpr int kAbsBlocks = 256;
__global__ void absmax_partial_kernel(const float* __restrict__ v, size_t n,
                                      float* __restrict__ part) {
    __shared__ float s[256];
    float m = 0.f;
    for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
         i += size_t(gridDim.x) * blockDim.x)
        m = fmaxf(m, fabsf(v[i]));
    s[threadIdx.x] = m;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (int(threadIdx.x) < w) s[threadIdx.x] = fmaxf(s[threadIdx.x], s[threadIdx.x + w]);
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = s[0];
}

__global__ void absmax_final_kernel(const float* __restrict__ part, int nparts,
                                    float* __restrict__ mm, XForm* __restrict__ xf) {
    __shared__ float s[256];
    float m = 0.f;
    for (int i = threadIdx.x; i < nparts; i += blockDim.x) m = fmaxf(m, part[i]);
    s[threadIdx.x] = m;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (int(threadIdx.x) < w) s[threadIdx.x] = fmaxf(s[threadIdx.x], s[threadIdx.x + w]);
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        mm[0] = s[0];
        mm[1] = 0.f;
        if (xf) *xf = XForm{0.f, 1.f, s[0], 0.f};  // x = in / max|in|
    }
}

// ---------------------------------------------------------------- L1 + SSIM
// SSIM with a (2r+1)^3 zero-padded box window normalised by the full window
// size (a symmetric operator, so its adjoint is itself), C1 = 1e-4, C2 = 9e-4.
__global__ void box_axis(const float* __restrict__ in, float* __restrict__ out, Dims d, int axis,
                         int r, int nch) {
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= d.vox() * nch) return;
    const size_t v = i / nch;
    const int c = int(i % nch);
    int z, y, x;
    unflat(v, d, z, y, x);
    float s = 0.f;
    for (int k = -r; k <= r; ++k) {
        int zz = z, yy = y, xx = x;
        if (axis == 0) xx += k;
        if (axis == 1) yy += k;
        if (axis == 2) zz += k;
        if (zz < 0 || zz >= d.D || yy < 0 || yy >= d.H || xx < 0 || xx >= d.W) continue;
        s += in[((size_t(zz) * d.H + yy) * d.W + xx) * nch + c];
    }
    out[i] = s / float(2 * r + 1);
}

// channels: 0 x, 1 y, 2 x^2, 3 y^2, 4 xy
__global__ void ssim_moments_in(const float* __restrict__ x, const float* __restrict__ t,
                                float* __restrict__ m, size_t n) {
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float a = x[i], b = t[i];
    m[i * 5 + 0] = a;
    m[i * 5 + 1] = b;
    m[i * 5 + 2] = a * a;
    m[i * 5 + 3] = b * b;
    m[i * 5 + 4] = a * b;
}

// Per voxel: SSIM map value and its partials w.r.t. (mu_x, E[x^2], E[xy])
// (the window means), written as 3 channels for the adjoint box pass.
__global__ void ssim_map(const float* __restrict__ mom, float* __restrict__ parts,
                         float* __restrict__ ssim_sum, size_t n) {
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    float s = 0.f;
    if (i < n) {
        const float C1 = 1e-4f, C2 = 9e-4f;
        const float mx = mom[i * 5 + 0], my = mom[i * 5 + 1];
        const float exx = mom[i * 5 + 2], eyy = mom[i * 5 + 3], exy = mom[i * 5 + 4];
        const float vx = exx - mx * mx, vy = eyy - my * my, cxy = exy - mx * my;
        const float A1 = 2.f * mx * my + C1, A2 = 2.f * cxy + C2;
        const float B1 = mx * mx + my * my + C1, B2 = vx + vy + C2;
        s = (A1 * A2) / (B1 * B2);
        // dS/dmx (holding vx, cxy) ; dS/dvx ; dS/dcxy
        const float dmx = (2.f * my * A2) / (B1 * B2) - s * (2.f * mx) / B1;
        const float dvx = -s / B2;
        const float dcxy = (2.f * A1) / (B1 * B2);
        // vx = exx - mx^2, cxy = exy - mx my: chain to (mx, exx, exy)
        parts[i * 3 + 0] = dmx - 2.f * mx * dvx - my * dcxy;
        parts[i * 3 + 1] = dvx;
        parts[i * 3 + 2] = dcxy;
    }
    __shared__ float red[256];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (int(threadIdx.x) < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicAdd(ssim_sum, red[0]);
}

// dL/dx = sign(x - t) / n  -  lambda / n * (B p0 + 2 x B p1 + t B p2)
__global__ void loss_grad(const float* __restrict__ x, const float* __restrict__ t,
                          const float* __restrict__ bp, float* __restrict__ g, size_t n,
                          float lambda, float* __restrict__ l1_sum) {
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    float a = 0.f;
    if (i < n) {
        const float diff = x[i] - t[i];
        a = fabsf(diff);
        const float sgn = diff > 0.f ? 1.f : (diff < 0.f ? -1.f : 0.f);
        const float dssim = bp[i * 3 + 0] + 2.f * x[i] * bp[i * 3 + 1] + t[i] * bp[i * 3 + 2];
        g[i] = (sgn - lambda * dssim) / float(n);
    }
    __shared__ float red[256];
    red[threadIdx.x] = a;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (int(threadIdx.x) < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicAdd(l1_sum, red[0]);
}

__global__ void adam_kernel(float* __restrict__ p, const float* __restrict__ g,
                            float* __restrict__ m, float* __restrict__ v, size_t n, float lr,
                            float b1, float b2, float c1, float c2, float eps, float gscale) {
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i] * gscale;
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// The last layer's FP32 weights (k = tap * 16 + cin) as two BF16 rows, hi
// and lo = bf16(w - hi), zero-padded to kD0K (the inference kernel's
// weight operand).
__global__ void split_last_layer(const float* __restrict__ w, bf16* __restrict__ ws) {
    const int k = int(blockIdx.x * blockDim.x + threadIdx.x);
    if (k >= kD0K) return;
    const float v = k < 27 * 16 ? w[k] : 0.f;
    const bf16 hi = f2bf(v);
    ws[k] = hi;
    ws[kD0K + k] = f2bf(v - bf2f(hi));
}

// Weights [tap][Cout][Cin] FP32 -> BF16 copies used by the MFMA kernels.
void cast_weights(const float* p, bf16* wb, bf16* w5s, hipStream_t s) {
    const size_t n = param_count();
    hipLaunchKernelGGL(to_bf16, g1(n), dim3(256), 0, s, p, wb, n);
    hipLaunchKernelGGL(split_last_layer, dim3((kD0K + 255) / 256), dim3(256), 0, s, p + woff(5),
                       w5s);
}

}  // namespace
}  // namespace fbp

// ------------------------------------------------------------------- C++ side

struct fbp_refiner_s {
    fbp_context ctx = nullptr;
    int S = 0;               // volume edge (multiple of 16)
    float* params = nullptr; // FP32 master
    bf16* wbf = nullptr;     // BF16 copy (same offsets)
    float* grad = nullptr;
    float* adam_m = nullptr;
    float* adam_v = nullptr;
    int step = 0;
    // activations
    float *x = nullptr, *y = nullptr, *g = nullptr;  // [V]
    bf16 *e1 = nullptr, *p1 = nullptr, *e2 = nullptr, *p2 = nullptr, *e3 = nullptr;
    bf16 *u2 = nullptr, *d2 = nullptr, *u1 = nullptr, *d1 = nullptr;
    float *pre1 = nullptr, *pre2 = nullptr, *pre3 = nullptr, *pre4 = nullptr, *pre5 = nullptr;
    // gradients of activations (FP32)
    float *gd1 = nullptr, *gu1 = nullptr, *gd2 = nullptr, *gu2 = nullptr, *ge3 = nullptr;
    float *gp2 = nullptr, *ge2 = nullptr, *gp1 = nullptr, *ge1 = nullptr, *tmp = nullptr;
    float* ssim = nullptr;   // moments (5V), parts (3V), box buffers (5V x 2), scalars
    float* c1 = nullptr;     // inference: relu(conv1(d2)) on the S/2 grid, FP32 [V/8][16]
    fbp::nn::XForm* xf = nullptr; // inference: input transform
    bf16* w5s = nullptr;     // inference: last layer's weights as BF16 hi and lo rows
    std::vector<void*> owned;
};

namespace fbp {
namespace {

template <typename T>
fbp_status alloc(fbp_refiner_s* r, T** p, size_t n) {
    FBP_CHECK_HIP(r->ctx, hipMalloc((void**)p, std::max<size_t>(n, 1) * sizeof(T)));
    r->owned.push_back(*p);
    return FBP_OK;
}

fbp_status refiner_alloc(fbp_refiner_s* r) {
    const size_t V = size_t(r->S) * r->S * r->S, V2 = V / 8, V4 = V / 64;
    const size_t P = param_count();
    FBP_CHECK(alloc(r, &r->params, P));
    FBP_CHECK(alloc(r, &r->wbf, P));
    FBP_CHECK(alloc(r, &r->grad, P));
    FBP_CHECK(alloc(r, &r->adam_m, P));
    FBP_CHECK(alloc(r, &r->adam_v, P));
    FBP_CHECK(alloc(r, &r->x, V));
    FBP_CHECK(alloc(r, &r->y, V));
    FBP_CHECK(alloc(r, &r->g, V));
    FBP_CHECK(alloc(r, &r->e1, V * 16));
    FBP_CHECK(alloc(r, &r->p1, V2 * 16));
    FBP_CHECK(alloc(r, &r->e2, V2 * 32));
    FBP_CHECK(alloc(r, &r->p2, V4 * 32));
    FBP_CHECK(alloc(r, &r->e3, V4 * 64));
    FBP_CHECK(alloc(r, &r->u2, V2 * 64));
    FBP_CHECK(alloc(r, &r->d2, V2 * 32));
    FBP_CHECK(alloc(r, &r->u1, V * 32));
    FBP_CHECK(alloc(r, &r->d1, V * 16));
    FBP_CHECK(alloc(r, &r->pre1, V * 16));
    FBP_CHECK(alloc(r, &r->pre2, V2 * 32));
    FBP_CHECK(alloc(r, &r->pre3, V4 * 64));
    FBP_CHECK(alloc(r, &r->pre4, V2 * 32));
    FBP_CHECK(alloc(r, &r->pre5, V * 16));
    FBP_CHECK(alloc(r, &r->gd1, V * 16));
    FBP_CHECK(alloc(r, &r->gu1, V * 32));
    FBP_CHECK(alloc(r, &r->gd2, V2 * 32));
    FBP_CHECK(alloc(r, &r->gu2, V2 * 64));
    FBP_CHECK(alloc(r, &r->ge3, V4 * 64));
    FBP_CHECK(alloc(r, &r->gp2, V4 * 32));
    FBP_CHECK(alloc(r, &r->ge2, V2 * 32));
    FBP_CHECK(alloc(r, &r->gp1, V2 * 16));
    FBP_CHECK(alloc(r, &r->ge1, V * 16));
    FBP_CHECK(alloc(r, &r->tmp, V * 32));
    FBP_CHECK(alloc(r, &r->ssim, V * 18 + 64 + kAbsBlocks));
    FBP_CHECK(alloc(r, &r->c1, V2 * 16));
    FBP_CHECK(alloc(r, &r->xf, 1));
    FBP_CHECK(alloc(r, &r->w5s, 2 * kD0K));
    return FBP_OK;
}

// Forward pass on r->x (normalised input, f32 [V]); output r->y (f32 [V]).
// train: keep pre-activations for the backward pass.
fbp_status refiner_forward_impl(fbp_refiner_s* r, hipStream_t s, bool train) {
    const int S = r->S;
    const Dims d1{S, S, S}, d2{S / 2, S / 2, S / 2}, d4{S / 4, S / 4, S / 4};
    const size_t V = d1.vox(), V2 = d2.vox(), V4 = d4.vox();
    const float* P = r->params;
    const bf16* Wb = r->wbf;
    // e1 = relu(conv3(x; 1->16))
    hipLaunchKernelGGL((conv_valu_fwd<1, 16, 3, float>), g1(V), dim3(256), 0, s, r->x, P + woff(0),
                       P + boff(0), train ? r->pre1 : nullptr, r->e1, d1, 1);
    hipLaunchKernelGGL(pool_fwd, g1(V2 * 16), dim3(256), 0, s, r->e1, r->p1, d2, 16);
    // Waves of the MFMA convolutions: MT tiles of 16 voxels each.
    auto waves = [](size_t vox, int mt) { return (vox / 16 + mt - 1) / mt * 64; };
    // e2 = relu(conv3(p1; 16->32))
    hipLaunchKernelGGL((conv_mfma_fwd<3, 16, 32, false, 2>), g1(waves(V2, 2)), dim3(256), 0, s,
                       r->p1, Wb + woff(1), P + boff(1), r->e2, nullptr,
                       train ? r->pre2 : nullptr, d2, 1);
    hipLaunchKernelGGL(pool_fwd, g1(V4 * 32), dim3(256), 0, s, r->e2, r->p2, d4, 32);
    // e3 = relu(conv3(p2; 32->64))
    hipLaunchKernelGGL((conv_mfma_fwd<3, 32, 64, false, 1>), g1(waves(V4, 1)), dim3(256), 0, s,
                       r->p2, Wb + woff(2), P + boff(2), r->e3, nullptr,
                       train ? r->pre3 : nullptr, d4, 1);
    // d2 = relu(conv3(up(e3); 64->32)) + e2  (upsampling fused into the input reads)
    hipLaunchKernelGGL((conv_mfma_fwd<3, 64, 32, true, 2>), g1(waves(V2, 2)), dim3(256), 0, s,
                       r->e3, Wb + woff(3), P + boff(3), r->d2, r->e2,
                       train ? r->pre4 : nullptr, d2, 1);
    // d1 = relu(conv1(up(d2); 32->16)) + e1
    hipLaunchKernelGGL((conv_mfma_fwd<1, 32, 16, true, 2>), g1(waves(V, 2)), dim3(256), 0, s,
                       r->d2, Wb + woff(4), P + boff(4), r->d1, r->e1,
                       train ? r->pre5 : nullptr, d1, 1);
    if (train) {
        // The backward pass takes the upsampled tensors as inputs of its
        // weight gradients.
        hipLaunchKernelGGL(up_fwd, g1(V2 * 64), dim3(256), 0, s, r->e3, r->u2, d2, 64);
        hipLaunchKernelGGL(up_fwd, g1(V * 32), dim3(256), 0, s, r->d2, r->u1, d1, 32);
    }
    // y = conv3(d1; 16->1) + x: the network predicts a correction of its
    // (normalised) input; the input needs no gradient, so the backward pass
    // starts from dL/dy unchanged.
    hipLaunchKernelGGL((conv_valu_fwd<16, 1, 3, bf16>), g1(V), dim3(256), 0, s, r->d1, P + woff(5),
                       P + boff(5), r->y, nullptr, d1, 0);
    hipLaunchKernelGGL(add_into, g1(V), dim3(256), 0, s, r->y, r->x, V);
    FBP_CHECK_HIP(r->ctx, hipGetLastError());
    (void)V2;
    return FBP_OK;
}

// Loss L = mean|y - t| + lambda (1 - mean SSIM(y, t)); r->g = dL/dy.
fbp_status refiner_loss(fbp_refiner_s* r, const float* target, float lambda, hipStream_t s) {
    const int S = r->S;
    const Dims d{S, S, S};
    const size_t V = d.vox();
    float* mom = r->ssim;            // 5V
    float* tmp5 = mom + 5 * V;       // 5V
    float* parts = tmp5 + 5 * V;     // 3V
    float* bparts = parts + 3 * V;   // 3V (+2V spare)
    float* scal = r->ssim + 18 * V;  // [0] ssim sum, [1] l1 sum
    const int rad = 2;
    hipLaunchKernelGGL(zero_f, dim3(1), dim3(64), 0, s, scal, size_t(2));
    hipLaunchKernelGGL(ssim_moments_in, g1(V), dim3(256), 0, s, r->y, target, tmp5, V);
    hipLaunchKernelGGL(box_axis, g1(V * 5), dim3(256), 0, s, tmp5, mom, d, 0, rad, 5);
    hipLaunchKernelGGL(box_axis, g1(V * 5), dim3(256), 0, s, mom, tmp5, d, 1, rad, 5);
    hipLaunchKernelGGL(box_axis, g1(V * 5), dim3(256), 0, s, tmp5, mom, d, 2, rad, 5);
    hipLaunchKernelGGL(ssim_map, g1(V), dim3(256), 0, s, mom, parts, scal, V);
    hipLaunchKernelGGL(box_axis, g1(V * 3), dim3(256), 0, s, parts, bparts, d, 0, rad, 3);
    hipLaunchKernelGGL(box_axis, g1(V * 3), dim3(256), 0, s, bparts, parts, d, 1, rad, 3);
    hipLaunchKernelGGL(box_axis, g1(V * 3), dim3(256), 0, s, parts, bparts, d, 2, rad, 3);
    hipLaunchKernelGGL(loss_grad, g1(V), dim3(256), 0, s, r->y, target, bparts, r->g, V, lambda,
                       scal + 1);
    FBP_CHECK_HIP(r->ctx, hipGetLastError());
    return FBP_OK;
}

// Backward from r->g into r->grad (accumulated).
fbp_status refiner_backward(fbp_refiner_s* r, hipStream_t s) {
    const int S = r->S;
    const Dims d1{S, S, S}, d2{S / 2, S / 2, S / 2}, d4{S / 4, S / 4, S / 4};
    const size_t V = d1.vox(), V2 = d2.vox(), V4 = d4.vox();
    float* G = r->grad;
    const float* P = r->params;
    auto wg_blocks = [](size_t vox, int taps) { return dim3(unsigned((vox + 63) / 64), unsigned(taps)); };
    // layer 6: y = conv3(d1; 16->1)
    hipLaunchKernelGGL((conv_valu_wgrad<16, 1, 3, bf16>), wg_blocks(V, 27), dim3(256), 0, s, r->g,
                       r->d1, G + woff(5), G + boff(5), d1);
    hipLaunchKernelGGL((conv_valu_dgrad<16, 1, 3>), g1(V * 16), dim3(256), 0, s, r->g, P + woff(5),
                       r->gd1, d1);
    // d1 = relu(pre5) + e1
    hipLaunchKernelGGL(relu_bwd, g1(V * 16), dim3(256), 0, s, r->gd1, r->pre5, r->tmp, V * 16);
    hipMemcpyAsync(r->ge1, r->gd1, V * 16 * sizeof(float), hipMemcpyDeviceToDevice, s);
    // layer 5: pre5 = conv1(u1; 32->16)
    hipLaunchKernelGGL((conv_valu_wgrad<32, 16, 1, bf16>), wg_blocks(V, 1), dim3(256), 0, s, r->tmp,
                       r->u1, G + woff(4), G + boff(4), d1);
    hipLaunchKernelGGL((conv_valu_dgrad<32, 16, 1>), g1(V * 32), dim3(256), 0, s, r->tmp,
                       P + woff(4), r->gu1, d1);
    // u1 = up(d2)
    hipLaunchKernelGGL(up_bwd, g1(V2 * 32), dim3(256), 0, s, r->gu1, r->gd2, d2, 32);
    // d2 = relu(pre4) + e2
    hipLaunchKernelGGL(relu_bwd, g1(V2 * 32), dim3(256), 0, s, r->gd2, r->pre4, r->tmp, V2 * 32);
    hipMemcpyAsync(r->ge2, r->gd2, V2 * 32 * sizeof(float), hipMemcpyDeviceToDevice, s);
    // layer 4: pre4 = conv3(u2; 64->32)
    hipLaunchKernelGGL((conv_valu_wgrad<64, 32, 3, bf16>), wg_blocks(V2, 27), dim3(256), 0, s,
                       r->tmp, r->u2, G + woff(3), G + boff(3), d2);
    hipLaunchKernelGGL((conv_valu_dgrad<64, 32, 3>), g1(V2 * 64), dim3(256), 0, s, r->tmp,
                       P + woff(3), r->gu2, d2);
    // u2 = up(e3)
    hipLaunchKernelGGL(up_bwd, g1(V4 * 64), dim3(256), 0, s, r->gu2, r->ge3, d4, 64);
    // e3 = relu(pre3)
    hipLaunchKernelGGL(relu_bwd, g1(V4 * 64), dim3(256), 0, s, r->ge3, r->pre3, r->tmp, V4 * 64);
    // layer 3: pre3 = conv3(p2; 32->64)
    hipLaunchKernelGGL((conv_valu_wgrad<32, 64, 3, bf16>), wg_blocks(V4, 27), dim3(256), 0, s,
                       r->tmp, r->p2, G + woff(2), G + boff(2), d4);
    hipLaunchKernelGGL((conv_valu_dgrad<32, 64, 3>), g1(V4 * 32), dim3(256), 0, s, r->tmp,
                       P + woff(2), r->gp2, d4);
    // p2 = pool(e2)
    hipLaunchKernelGGL(pool_bwd, g1(V2 * 32), dim3(256), 0, s, r->gp2, r->ge2, d2, 32);
    // e2 = relu(pre2)
    hipLaunchKernelGGL(relu_bwd, g1(V2 * 32), dim3(256), 0, s, r->ge2, r->pre2, r->tmp, V2 * 32);
    // layer 2: pre2 = conv3(p1; 16->32)
    hipLaunchKernelGGL((conv_valu_wgrad<16, 32, 3, bf16>), wg_blocks(V2, 27), dim3(256), 0, s,
                       r->tmp, r->p1, G + woff(1), G + boff(1), d2);
    hipLaunchKernelGGL((conv_valu_dgrad<16, 32, 3>), g1(V2 * 16), dim3(256), 0, s, r->tmp,
                       P + woff(1), r->gp1, d2);
    // p1 = pool(e1)
    hipLaunchKernelGGL(pool_bwd, g1(V * 16), dim3(256), 0, s, r->gp1, r->ge1, d1, 16);
    // e1 = relu(pre1)
    hipLaunchKernelGGL(relu_bwd, g1(V * 16), dim3(256), 0, s, r->ge1, r->pre1, r->tmp, V * 16);
    // layer 1: pre1 = conv3(x; 1->16)
    hipLaunchKernelGGL((conv_valu_wgrad<1, 16, 3, float>), wg_blocks(V, 27), dim3(256), 0, s,
                       r->tmp, r->x, G + woff(0), G + boff(0), d1);
    FBP_CHECK_HIP(r->ctx, hipGetLastError());
    return FBP_OK;
}

fbp_status normalize_input(fbp_refiner_s* r, const float* in, hipStream_t s) {
    const size_t V = size_t(r->S) * r->S * r->S;
    float* mm = r->ssim + 18 * V + 8;
    float* part = r->ssim + 18 * V + 64;
    hipLaunchKernelGGL(absmax_partial_kernel, dim3(kAbsBlocks), dim3(256), 0, s, in, V, part);
    hipLaunchKernelGGL(absmax_final_kernel, dim3(1), dim3(256), 0, s, part, kAbsBlocks, mm,
                       nullptr);
    hipLaunchKernelGGL(scale_f, g1(V), dim3(256), 0, s, in, r->x, V, mm);
    FBP_CHECK_HIP(r->ctx, hipGetLastError());
    return FBP_OK;
}

template <int CIN, int COUT, bool UP, int NBZ, int NBY, int NBX, int NWT, int MODE>
fbp_status launch_conv3(fbp_context ctx, const bf16* x, const bf16* w, const float* b,
                        const bf16* skip, bf16* y, bf16* pool, const bf16* cw, const float* cb,
                        float* cy, Dims d, hipStream_t s) {
    using Cf = Conv3Cfg<CIN, COUT, UP, NBZ, NBY, NBX, NWT, MODE>;
    if (d.D % Cf::TZ || d.H % Cf::TY || d.W % Cf::TX)
        return fail(ctx, FBP_ERR_UNSUPPORTED, "refiner: grid %dx%dx%d is not a multiple of the tile",
                    d.D, d.H, d.W);
    if (Cf::LDS > size_t(ctx->max_lds_per_block))
        return fail(ctx, FBP_ERR_UNSUPPORTED, "refiner: %zu B of LDS per block needed", Cf::LDS);
    auto kern = conv3_mfma_kernel<CIN, COUT, UP, NBZ, NBY, NBX, NWT, MODE>;
    // The attribute is set once per device (a host call on every launch
    // would stall the enqueue of the following kernels).
    static std::atomic<uint64_t> attr_set{0};
    const uint64_t bit = uint64_t(1) << (ctx->device & 63);
    if (!(attr_set.load(std::memory_order_acquire) & bit)) {
        FBP_CHECK_HIP(ctx, hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                                               hipFuncAttributeMaxDynamicSharedMemorySize,
                                               int(Cf::LDS)));
        attr_set.fetch_or(bit, std::memory_order_release);
    }
    const unsigned blocks = unsigned(d.vox() / (size_t(Cf::TZ) * Cf::TY * Cf::TX));
    hipLaunchKernelGGL(kern, dim3(blocks), dim3(256), Cf::LDS, s, x, w, b, skip, y, pool, cw, cb,
                       cy, d);
    FBP_CHECK_HIP(ctx, hipGetLastError());
    return FBP_OK;
}

// Inference forward in five fused kernels (refiner_fused.h): in is the raw
// input, xf its affine normalisation (device), or, when mm.part is set, the
// min-max normalisation from in's per-block (min, max) pairs; out the
// refined volume.
fbp_status refiner_infer(fbp_refiner_s* r, const float* in, const XForm* xf,
                         const MinMaxPartials& mm, float* out, hipStream_t s) {
    const int S = r->S;
    const Dims d1{S, S, S}, d2{S / 2, S / 2, S / 2}, d4{S / 4, S / 4, S / 4};
    const float* P = r->params;
    const bf16* Wb = r->wbf;
    fbp_context ctx = r->ctx;
    hipLaunchKernelGGL(enc0_kernel, dim3(unsigned(d1.vox() / (kE0Z * kE0Y * kE0X))), dim3(256), 0,
                       s, in, xf, mm.part, mm.count, P + woff(0), P + boff(0), r->e1, r->p1, d1);
    FBP_CHECK_HIP(ctx, hipGetLastError());
    FBP_CHECK((launch_conv3<16, 32, false, 2, 2, 2, 2, kModePool>(
        ctx, r->p1, Wb + woff(1), P + boff(1), nullptr, r->e2, r->p2, nullptr, nullptr, nullptr,
        d2, s)));
    FBP_CHECK((launch_conv3<32, 64, false, 1, 1, 1, 1, kModePlain>(
        ctx, r->p2, Wb + woff(2), P + boff(2), nullptr, r->e3, nullptr, nullptr, nullptr, nullptr,
        d4, s)));
    FBP_CHECK((launch_conv3<64, 32, true, 2, 2, 2, 2, kModeChain>(
        ctx, r->e3, Wb + woff(3), P + boff(3), r->e2, nullptr, nullptr, Wb + woff(4),
        P + boff(4), r->c1, d2, s)));
    hipLaunchKernelGGL(dec0_kernel, dim3(unsigned(d1.vox() / (kD0Z * kD0Y * kD0X))), dim3(256), 0,
                       s, r->c1, r->e1, r->w5s, P + boff(5), in, xf, mm.part, mm.count, out, d1);
    FBP_CHECK_HIP(ctx, hipGetLastError());
    return FBP_OK;
}

}  // namespace
}  // namespace fbp

using namespace fbp;

extern "C" {

size_t fbp_refiner_param_count(void) { return param_count(); }

fbp_status fbp_refiner_create(fbp_context ctx, uint32_t size, uint64_t seed, fbp_refiner* out) {
    if (!ctx || !out) return fail(ctx, FBP_ERR_INVALID_ARGUMENT, "refiner: null argument");
    *out = nullptr;
    if (size < 16 || size % 16 != 0)
        return fail(ctx, FBP_ERR_UNSUPPORTED, "refiner: volume edge must be a multiple of 16");
    FBP_CHECK(bind(ctx));
    fbp_refiner r = new fbp_refiner_s();
    r->ctx = ctx;
    r->S = int(size);
    fbp_status st = refiner_alloc(r);
    if (st != FBP_OK) {
        fbp_refiner_destroy(r);
        return st;
    }
    // He-normal initialisation from a 64-bit LCG + Box-Muller (host, seeded).
    std::vector<float> p(param_count(), 0.f);
    uint64_t state = seed * 6364136223846793005ull + 1442695040888963407ull;
    auto u01 = [&]() {
        state = state * 6364136223846793005ull + 1442695040888963407ull;
        return (double((state >> 11) & ((1ull << 53) - 1)) + 0.5) / double(1ull << 53);
    };
    for (int l = 0; l < kLayers; ++l) {
        const double fan_in = double(kKs[l]) * kKs[l] * kKs[l] * kCin[l];
        const double sd = std::sqrt(2.0 / fan_in) * (l == kLayers - 1 ? 0.1 : 1.0);
        for (size_t i = 0; i < wcount(l); ++i) {
            const double g = std::sqrt(-2.0 * std::log(u01())) * std::cos(6.283185307179586 * u01());
            p[woff(l) + i] = float(g * sd);
        }
    }
    st = fbp_refiner_load(r, p.data(), p.size());
    if (st != FBP_OK) {
        fbp_refiner_destroy(r);
        return st;
    }
    *out = r;
    return FBP_OK;
}

fbp_status fbp_refiner_destroy(fbp_refiner r) {
    if (!r) return FBP_ERR_INVALID_ARGUMENT;
    hipSetDevice(r->ctx->device);
    hipDeviceSynchronize();
    for (void* p : r->owned) hipFree(p);
    delete r;
    return FBP_OK;
}

fbp_status fbp_refiner_load(fbp_refiner r, const float* host_params, uint64_t count) {
    if (!r || !host_params) return FBP_ERR_INVALID_ARGUMENT;
    if (count != param_count())
        return fail(r->ctx, FBP_ERR_INVALID_ARGUMENT, "refiner: expected %zu parameters", param_count());
    FBP_CHECK(bind(r->ctx));
    FBP_CHECK_HIP(r->ctx, hipMemcpy(r->params, host_params, count * 4, hipMemcpyHostToDevice));
    FBP_CHECK_HIP(r->ctx, hipMemset(r->adam_m, 0, count * 4));
    FBP_CHECK_HIP(r->ctx, hipMemset(r->adam_v, 0, count * 4));
    r->step = 0;
    cast_weights(r->params, r->wbf, r->w5s, nullptr);
    FBP_CHECK_HIP(r->ctx, hipDeviceSynchronize());
    return FBP_OK;
}

fbp_status fbp_refiner_save(fbp_refiner r, float* host_params, uint64_t count) {
    if (!r || !host_params) return FBP_ERR_INVALID_ARGUMENT;
    if (count != param_count()) return FBP_ERR_INVALID_ARGUMENT;
    FBP_CHECK(bind(r->ctx));
    FBP_CHECK_HIP(r->ctx, hipDeviceSynchronize());
    FBP_CHECK_HIP(r->ctx, hipMemcpy(host_params, r->params, count * 4, hipMemcpyDeviceToHost));
    return FBP_OK;
}

fbp_status fbp_refiner_forward(fbp_refiner r, const float* input, float* output, void* stream) {
    if (!r || !input || !output) return FBP_ERR_INVALID_ARGUMENT;
    FBP_CHECK(bind(r->ctx));
    hipStream_t s = as_stream(stream);
    const size_t V = size_t(r->S) * r->S * r->S;
    float* mm = r->ssim + 18 * V + 8;
    float* part = r->ssim + 18 * V + 64;
    hipLaunchKernelGGL(absmax_partial_kernel, dim3(kAbsBlocks), dim3(256), 0, s, input, V, part);
    hipLaunchKernelGGL(absmax_final_kernel, dim3(1), dim3(256), 0, s, part, kAbsBlocks, mm, r->xf);
    FBP_CHECK_HIP(r->ctx, hipGetLastError());
    return refiner_infer(r, input, r->xf, MinMaxPartials{}, output, s);
}

fbp_status fbp_learned_reconstruct(fbp_refiner r, const fbp_reconstruction_info* info,
                                   const fbp_transient* tr, const fbp_vec3* laser_targets,
                                   uint32_t num_samples, uint64_t seed, int32_t kernel_size,
                                   float sigma, uint32_t extremes, float* volume, void* stream) {
    if (!r || !info || !tr || !laser_targets || !volume) return FBP_ERR_INVALID_ARGUMENT;
    fbp_context ctx = r->ctx;
    const uint32_t S = uint32_t(r->S);
    if (info->voxel_resolution.x != S || info->voxel_resolution.y != S ||
        info->voxel_resolution.z != S)
        return fail(ctx, FBP_ERR_INVALID_ARGUMENT, "learned: the voxel grid must be the refiner's %u^3",
                    S);
    if (kernel_size != 5)
        return fail(ctx, FBP_ERR_UNSUPPORTED, "learned: LoG kernel size 5 only");
    FBP_CHECK(bind(ctx));
    hipStream_t s = as_stream(stream);
    // The MIS volume and its LoG live in the refiner's training buffers x, y.
    fbp_mis_params p{};
    p.num_samples = num_samples;
    p.mode = FBP_MIS_TILED;
    p.noise = nullptr;
    p.noise_size = S * S * S;
    p.seed = seed;
    FBP_CHECK(fbp_backproject_confocal_mis(ctx, info, tr, laser_targets, &p, r->x, stream));
    MinMaxPartials mm{};
    FBP_CHECK(log_impl(ctx, r->x, r->y, fbp_uvec3{S, S, S}, kernel_size, sigma, s, &mm));
    FBP_CHECK(refiner_infer(r, r->y, nullptr, mm, volume, s));
    return learned_calibrate_impl(ctx, info, tr, laser_targets, volume, r->y, kernel_size, sigma,
                                  extremes, s);
}

fbp_status fbp_refiner_train_step(fbp_refiner r, const float* inputs, const float* targets,
                                  uint32_t batch, float lr, float ssim_weight, float* host_loss,
                                  void* stream) {
    if (!r || !inputs || !targets || batch == 0) return FBP_ERR_INVALID_ARGUMENT;
    FBP_CHECK(bind(r->ctx));
    hipStream_t s = as_stream(stream);
    const size_t V = size_t(r->S) * r->S * r->S, P = param_count();
    hipLaunchKernelGGL(zero_f, g1(P), dim3(256), 0, s, r->grad, P);
    double loss = 0.0;
    float* scal = r->ssim + 18 * V;
    for (uint32_t b = 0; b < batch; ++b) {
        FBP_CHECK(normalize_input(r, inputs + b * V, s));
        FBP_CHECK(refiner_forward_impl(r, s, true));
        FBP_CHECK(refiner_loss(r, targets + b * V, ssim_weight, s));
        FBP_CHECK(refiner_backward(r, s));
        if (host_loss) {
            float sc[2];
            FBP_CHECK_HIP(r->ctx, hipMemcpyAsync(sc, scal, 8, hipMemcpyDeviceToHost, s));
            FBP_CHECK_HIP(r->ctx, hipStreamSynchronize(s));
            loss += double(sc[1]) / double(V) + double(ssim_weight) * (1.0 - double(sc[0]) / double(V));
        }
    }
    r->step += 1;
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.f - std::pow(b1, float(r->step)), c2 = 1.f - std::pow(b2, float(r->step));
    hipLaunchKernelGGL(adam_kernel, g1(P), dim3(256), 0, s, r->params, r->grad, r->adam_m, r->adam_v,
                       P, lr, b1, b2, c1, c2, 1e-8f, 1.f / float(batch));
    cast_weights(r->params, r->wbf, r->w5s, s);
    FBP_CHECK_HIP(r->ctx, hipGetLastError());
    if (host_loss) *host_loss = float(loss / batch);
    return FBP_OK;
}

fbp_status fbp_refiner_gradient(fbp_refiner r, const float* input, const float* target,
                                float ssim_weight, float* host_grad, float* host_loss) {
    if (!r || !input || !target || !host_grad) return FBP_ERR_INVALID_ARGUMENT;
    FBP_CHECK(bind(r->ctx));
    const size_t V = size_t(r->S) * r->S * r->S, P = param_count();
    hipLaunchKernelGGL(zero_f, g1(P), dim3(256), 0, nullptr, r->grad, P);
    FBP_CHECK(normalize_input(r, input, nullptr));
    FBP_CHECK(refiner_forward_impl(r, nullptr, true));
    FBP_CHECK(refiner_loss(r, target, ssim_weight, nullptr));
    FBP_CHECK(refiner_backward(r, nullptr));
    FBP_CHECK_HIP(r->ctx, hipDeviceSynchronize());
    FBP_CHECK_HIP(r->ctx, hipMemcpy(host_grad, r->grad, P * 4, hipMemcpyDeviceToHost));
    if (host_loss) {
        float sc[2];
        FBP_CHECK_HIP(r->ctx, hipMemcpy(sc, r->ssim + 18 * V, 8, hipMemcpyDeviceToHost));
        *host_loss = float(double(sc[1]) / V + double(ssim_weight) * (1.0 - double(sc[0]) / V));
    }
    return FBP_OK;
}

}  // extern "C"
