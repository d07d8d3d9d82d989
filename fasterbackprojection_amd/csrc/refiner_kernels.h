// Kernels of the learned refiner (3D U-Net, channel-last [D][H][W][C]).
// No reference counterpart: the learned form maps a cheap MIS-sampled
// backprojection to the plain backprojection + LoG volume.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace fbp {
namespace nn {

typedef uint16_t bf16;
typedef short s4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ float bf2f(bf16 v) { return __uint_as_float(uint32_t(v) << 16); }
__device__ __forceinline__ bf16 f2bf(float f) {
    uint32_t b = __float_as_uint(f);
    if ((b & 0x7f800000u) == 0x7f800000u) return bf16(b >> 16);  // inf / nan
    b += 0x7fffu + ((b >> 16) & 1u);                              // round to nearest even
    return bf16(b >> 16);
}

struct Dims {
    int D, H, W;
    __host__ __device__ __forceinline__ size_t vox() const { return size_t(D) * H * W; }
};

__device__ __forceinline__ void unflat(size_t v, const Dims& d, int& z, int& y, int& x) {
    x = int(v % d.W);
    y = int((v / d.W) % d.H);
    z = int(v / (size_t(d.W) * d.H));
}

// Forward convolution, Cin and Cout multiples of 16, kernel KS^3 (1 or 3),
// zero padding KS/2, BF16 in, FP32 accumulation on MFMA 16x16x16, bias,
// optional ReLU, then optional skip addition; BF16 out.  Weights
// [tap][Cout][Cin] BF16.  One wave: MT tiles of 16 consecutive voxels times
// all Cout; every weight fragment it loads feeds the MT tiles.  Lane l: row
// (voxel) l % 16, k-group l / 16 (4 input channels).  UP: the input is the
// 2x coarser grid, read at the halved coordinates (fused nearest-neighbour
// upsampling).  Needs the voxel count to be a multiple of 16.
template <int KS, int CIN, int COUT, bool UP = false, int MT = 1>
__global__ __launch_bounds__(256) void conv_mfma_fwd(const bf16* __restrict__ x,
                                                     const bf16* __restrict__ w,
                                                     const float* __restrict__ bias,
                                                     bf16* __restrict__ y,
                                                     const bf16* __restrict__ skip,
                                                     float* __restrict__ pre_relu, Dims d,
                                                     int relu) {
    constexpr int NT = COUT / 16;
    constexpr int R = KS / 2;
    const int lane = threadIdx.x & 63;
    const size_t wave = (size_t(blockIdx.x) * blockDim.x + threadIdx.x) >> 6;
    const size_t tiles = d.vox() / 16;
    const size_t t0 = wave * MT;
    if (t0 >= tiles) return;
    const int row = lane & 15, kg = lane >> 4;
    const Dims din = UP ? Dims{d.D / 2, d.H / 2, d.W / 2} : d;
    // The lane's own voxel in each tile (a tile may span rows when W < 16).
    int z[MT], yy[MT], xr[MT];
    bool live[MT];
#pragma unroll
    for (int m = 0; m < MT; ++m) {
        live[m] = t0 + m < tiles;
        unflat((t0 + m) * 16 + row, d, z[m], yy[m], xr[m]);
    }
    f4 acc[MT][NT];
#pragma unroll
    for (int m = 0; m < MT; ++m)
#pragma unroll
        for (int n = 0; n < NT; ++n) acc[m][n] = f4{0.f, 0.f, 0.f, 0.f};
    for (int dz = -R; dz <= R; ++dz) {
        for (int dy = -R; dy <= R; ++dy) {
            for (int dx = -R; dx <= R; ++dx) {
                const int tap = ((dz + R) * KS + (dy + R)) * KS + (dx + R);
                size_t src[MT];
                bool in[MT];
#pragma unroll
                for (int m = 0; m < MT; ++m) {
                    const int zz = z[m] + dz, yv = yy[m] + dy, xv = xr[m] + dx;
                    in[m] = live[m] && zz >= 0 && zz < d.D && yv >= 0 && yv < d.H && xv >= 0 &&
                            xv < d.W;
                    const int sz = UP ? zz >> 1 : zz, sy = UP ? yv >> 1 : yv,
                              sx = UP ? xv >> 1 : xv;
                    src[m] = ((size_t(sz) * din.H + sy) * din.W + sx) * CIN;
                }
#pragma unroll
                for (int c0 = 0; c0 < CIN; c0 += 16) {
                    s4 a[MT];
#pragma unroll
                    for (int m = 0; m < MT; ++m) {
                        a[m] = s4{0, 0, 0, 0};
                        if (in[m]) a[m] = *reinterpret_cast<const s4*>(x + src[m] + c0 + 4 * kg);
                    }
#pragma unroll
                    for (int n = 0; n < NT; ++n) {
                        const s4 b = *reinterpret_cast<const s4*>(
                            w + (size_t(tap) * COUT + n * 16 + row) * CIN + c0 + 4 * kg);
#pragma unroll
                        for (int m = 0; m < MT; ++m)
                            acc[m][n] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[m], b, acc[m][n],
                                                                                  0, 0, 0);
                    }
                }
            }
        }
    }
    // D[i][j]: lane holds rows 4*kg + e, column row (= output channel offset)
#pragma unroll
    for (int m = 0; m < MT; ++m) {
        if (!live[m]) continue;
        const size_t v0 = (t0 + m) * 16;
#pragma unroll
        for (int n = 0; n < NT; ++n) {
            const int co = n * 16 + row;
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const size_t v = v0 + 4 * kg + e;
                float val = acc[m][n][e] + bias[co];
                if (pre_relu) pre_relu[v * COUT + co] = val;
                if (relu) val = fmaxf(val, 0.f);
                if (skip) val += bf2f(skip[v * COUT + co]);
                y[v * COUT + co] = f2bf(val);
            }
        }
    }
}

}  // namespace nn
}  // namespace fbp
