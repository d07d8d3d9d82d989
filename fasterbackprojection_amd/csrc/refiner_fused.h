// Inference forward of the learned refiner in five fused kernels.
//
//   enc0_kernel   e1 = relu(conv3(x; 1->16)) @ S and p1 = avgpool2(e1)      (VALU)
//   conv3_mfma    e2 = relu(conv3(p1; 16->32)) @ S/2 and p2 = avgpool2(e2)  (MFMA)
//   conv3_mfma    e3 = relu(conv3(p2; 32->64)) @ S/4                        (MFMA)
//   conv3_mfma    d2 = relu(conv3(up(e3); 64->32)) + e2 @ S/2, then
//                 c = relu(conv1(d2; 32->16)) @ S/2 from registers         (MFMA)
//   dec0_kernel   y = conv3(bf16(up(c) + e1); 16->1) + x @ S                (VALU)
//
// The 1x1x1 convolution commutes with nearest-neighbour upsampling, so it
// runs on the coarse grid and its output c (FP32) is upsampled when the last
// kernel stages its input: d1 = bf16(c(parent) + e1) is what the layer-wise
// forward stores.  x is the refiner input after its affine normalisation
// x = ((in - off) * mul) / max(|a|, |b|) (0 when that is 0), evaluated from
// the raw input by the first and last kernels, so no normalised copy exists.
//
// Every kernel stages its input tile (with the 3^3 halo) in LDS once.  The
// MFMA kernels are implicit GEMMs in the transposed form D[cout][voxel] =
// W^T[cout][k] * P[k][voxel] on v_mfma_f32_16x16x32_bf16 (k = tap * Cin +
// cin): lane l supplies weight row l % 16 and patch column (voxel) l % 16,
// 8 consecutive k of group l / 16; its accumulator holds 4 consecutive
// output channels of voxel l % 16, which is the layout the next layer's
// patch operand needs, so the 1x1x1 layer chains in registers.  A wave's
// 16 voxels form a 2 x 2 x 4 (z, y, x) brick.  The whole weight tensor of
// the layer sits in LDS next to the input tile (up to 124 KB per block).
#pragma once

#include "chain.h"
#include "refiner_kernels.h"

namespace fbp {
namespace nn {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));

// Input transform of the refiner: x = ((in - off) * mul) / max(|a|, |b|).
struct XForm {
    float off, mul, a, b;
};

__device__ __forceinline__ float xform_apply(float v, const XForm& t) {
    const float m = fmaxf(fabsf(t.a), fabsf(t.b));
    return m > 0.f ? rn_mul(rn_sub(v, t.off), t.mul) / m : 0.f;
}

// The input transform: *xf, or, with mm_part, the min-max normalisation's
// from per-block (min, max) pairs of the input: off = min, mul = m =
// safeRCP(max - min), a = (max - min) m (the largest normalised value), so
// x = ((v - min) m) / ((max - min) m).  Called by the whole block.
__device__ __forceinline__ XForm load_xform(const XForm* __restrict__ xf,
                                            const float* __restrict__ mm_part, uint32_t mm_count) {
    if (!mm_part) return *xf;
    float mn, mx;
    block_minmax(mm_part, mm_count, mn, mx);
    const float m = safe_rcp(rn_sub(mx, mn));
    return XForm{mn, m, rn_mul(rn_sub(mx, mn), m), 0.f};
}

__device__ __forceinline__ bf16x8 frag_at(const bf16* p) {
    return (bf16x8)(*reinterpret_cast<const u32x4*>(p));
}

// ---------------------------------------------------------------- bricks
// Lane r16's voxel in brick b of an (NBZ, NBY, NBX)-brick tile.
template <int NBY, int NBX>
__device__ __forceinline__ void brick_voxel(int b, int r16, int& vz, int& vy, int& vx) {
    const int bbx = b % NBX, bby = (b / NBX) % NBY, bbz = b / (NBX * NBY);
    vz = 2 * bbz + (r16 >> 3);
    vy = 2 * bby + ((r16 >> 2) & 1);
    vx = 4 * bbx + (r16 & 3);
}

// ---------------------------------------------------------------- enc0
// Block: a (4, 4, 8) (z, y, x) tile = 8 bricks, two per wave.  The 1 -> 16
// layer is a K = 27 (+1 zero tap) GEMM on v_mfma_f32_16x16x4_f32 (FP32
// operands, as the layer-wise FP32 convolution): lane (r16, g) supplies
// weight w[tap = 4 s + g][cout = r16] and x at that tap of voxel r16.
// Requires D % 4 == H % 4 == W % 8 == 0.
constexpr int kE0Z = 4, kE0Y = 4, kE0X = 8;

__global__ __launch_bounds__(256) void enc0_kernel(const float* __restrict__ in,
                                                   const XForm* __restrict__ xf,
                                                   const float* __restrict__ mm_part,
                                                   uint32_t mm_count,
                                                   const float* __restrict__ w,
                                                   const float* __restrict__ bias,
                                                   bf16* __restrict__ e1, bf16* __restrict__ p1,
                                                   Dims d) {
    constexpr int HZ = kE0Z + 2, HY = kE0Y + 2, HX = kE0X + 2;
    __shared__ float xt[HZ * HY * HX];
    __shared__ __align__(16) bf16 et[kE0Z * kE0Y * kE0X * 16];
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int r16 = lane & 15, g = lane >> 4;
    const int bx = int(blockIdx.x) % (d.W / kE0X);
    const int by = (int(blockIdx.x) / (d.W / kE0X)) % (d.H / kE0Y);
    const int bz = int(blockIdx.x) / ((d.W / kE0X) * (d.H / kE0Y));
    const int z0 = bz * kE0Z, y0 = by * kE0Y, x0 = bx * kE0X;
    // Weight operands of the 7 k-steps (L1/L2-resident), issued first.
    float a[7];
#pragma unroll
    for (int s = 0; s < 7; ++s) {
        const int tap = 4 * s + g;
        a[s] = w[min(tap, 26) * 16 + r16];
        if (tap >= 27) a[s] = 0.f;
    }
    // Every load of the tile is issued before the first LDS store (and
    // before the input transform's own loads and reduction), so the block
    // waits for memory once.  Branch-free: the address is clamped into the
    // volume, the value masked.
    constexpr int NH = HZ * HY * HX, PH = (NH + 255) / 256;
    float hv[PH];
    bool hin[PH];
#pragma unroll
    for (int j = 0; j < PH; ++j) {
        const int i = t + j * 256;
        const int hx = i % HX, hy = (i / HX) % HY, hz = i / (HX * HY);
        const int gz = z0 + hz - 1, gy = y0 + hy - 1, gx = x0 + hx - 1;
        hin[j] = i < NH && gz >= 0 && gz < d.D && gy >= 0 && gy < d.H && gx >= 0 && gx < d.W;
        const int cz = min(max(gz, 0), d.D - 1), cy = min(max(gy, 0), d.H - 1),
                  cx = min(max(gx, 0), d.W - 1);
        hv[j] = in[(size_t(cz) * d.H + cy) * d.W + cx];
    }
    const XForm tf = load_xform(xf, mm_part, mm_count);
#pragma unroll
    for (int j = 0; j < PH; ++j) {
        const int i = t + j * 256;
        if (i < NH) xt[i] = hin[j] ? xform_apply(hv[j], tf) : 0.f;
    }
    __syncthreads();
    int vz[2], vy[2], vx[2];
    f4 acc[2];
#pragma unroll
    for (int i = 0; i < 2; ++i) {
        brick_voxel<2, 2>(2 * wave + i, r16, vz[i], vy[i], vx[i]);
        acc[i] = f4{0.f, 0.f, 0.f, 0.f};
    }
#pragma unroll
    for (int s = 0; s < 7; ++s) {
        const int tap = min(4 * s + g, 26);
        const int dz = tap / 9, dy = (tap / 3) % 3, dx = tap % 3;
#pragma unroll
        for (int i = 0; i < 2; ++i) {
            const float b = xt[((vz[i] + dz) * HY + vy[i] + dy) * HX + vx[i] + dx];
            acc[i] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[s], b, acc[i], 0, 0, 0);
        }
    }
    // Lane (r16, g) holds channels 4 g + e of voxel r16.
#pragma unroll
    for (int i = 0; i < 2; ++i) {
        bf16 o[4];
#pragma unroll
        for (int e = 0; e < 4; ++e) o[e] = f2bf(fmaxf(acc[i][e] + bias[4 * g + e], 0.f));
        u32x2 pk;
        pk.x = uint32_t(o[0]) | (uint32_t(o[1]) << 16);
        pk.y = uint32_t(o[2]) | (uint32_t(o[3]) << 16);
        const size_t v = (size_t(z0 + vz[i]) * d.H + y0 + vy[i]) * d.W + x0 + vx[i];
        *reinterpret_cast<u32x2*>(e1 + v * 16 + 4 * g) = pk;
        *reinterpret_cast<u32x2*>(et + ((vz[i] * kE0Y + vy[i]) * kE0X + vx[i]) * 16 + 4 * g) = pk;
    }
    __syncthreads();
    // Average pool: (2, 2, 4) parents x 16 channels, one per thread, the 8
    // children summed in (z, y, x) order as the layer-wise pooling does.
    {
        const int c = t & 15, pv = t >> 4;
        const int px = pv & 3, py = (pv >> 2) & 1, pz = pv >> 3;
        float sum = 0.f;
#pragma unroll
        for (int a2 = 0; a2 < 2; ++a2)
#pragma unroll
            for (int b2 = 0; b2 < 2; ++b2)
#pragma unroll
                for (int e2 = 0; e2 < 2; ++e2)
                    sum += bf2f(et[(((2 * pz + a2) * kE0Y + 2 * py + b2) * kE0X + 2 * px + e2) * 16 + c]);
        const Dims dp{d.D / 2, d.H / 2, d.W / 2};
        const size_t pvox = (size_t(z0 / 2 + pz) * dp.H + y0 / 2 + py) * dp.W + x0 / 2 + px;
        p1[pvox * 16 + c] = f2bf(sum * 0.125f);
    }
}

// ---------------------------------------------------------------- conv3_mfma
enum { kModePool = 0, kModePlain = 1, kModeChain = 2 };

// CIN -> COUT 3^3 convolution, zero padding, bias + ReLU.  UP: the input is
// the 2x coarser grid, read at the halved coordinates.  Block: 4 waves over
// an (NBZ, NBY, NBX)-brick tile; NWT 16-channel output tiles per wave.
//   kModePool:  y = bf16(relu) and pool_out = avgpool2(y)
//   kModePlain: y = bf16(relu)
//   kModeChain: d = bf16(relu + skip) kept in registers, then
//               chain_out = relu(conv1(d; COUT -> 16) + chain_b) in FP32.
// Dims d: output grid; it must be a multiple of the tile.
template <int CIN, int COUT, bool UP, int NBZ, int NBY, int NBX, int NWT, int MODE>
struct Conv3Cfg {
    static constexpr int TZ = 2 * NBZ, TY = 2 * NBY, TX = 4 * NBX;
    static constexpr int NB = NBZ * NBY * NBX;
    static constexpr int NT = COUT / 16;
    static constexpr int NG = NT / NWT;      // wave groups over output channels
    static constexpr int BG = 4 / NG;        // wave groups over bricks
    static constexpr int BPW = NB / BG;      // bricks per wave
    static constexpr int IZ = UP ? TZ / 2 + 2 : TZ + 2;
    static constexpr int IY = UP ? TY / 2 + 2 : TY + 2;
    static constexpr int IX = UP ? TX / 2 + 2 : TX + 2;
    static constexpr int CSTR = CIN + 8;     // bf16 per staged voxel (padded)
    static constexpr int IN_ELEMS = IZ * IY * IX * CSTR;
    static constexpr int W_ELEMS = 27 * COUT * CIN;
    static constexpr int KSTEPS = (27 * CIN + 31) / 32;
    static constexpr size_t LDS = size_t(IN_ELEMS + W_ELEMS) * sizeof(bf16);
    static_assert(NT % NWT == 0 && 4 % NG == 0 && NB % BG == 0, "wave split");
    static_assert(CIN % 8 == 0 && (CIN & (CIN - 1)) == 0, "CIN");
    static_assert(MODE != kModePool || size_t(TZ * TY * TX * COUT) <= size_t(IN_ELEMS), "pool tile");
    static_assert(MODE != kModeChain || (COUT == 32 && NWT == 2), "chain needs all 32 channels");
};

template <int CIN, int COUT, bool UP, int NBZ, int NBY, int NBX, int NWT, int MODE>
__global__ __launch_bounds__(256) void conv3_mfma_kernel(
    const bf16* __restrict__ x, const bf16* __restrict__ w, const float* __restrict__ bias,
    const bf16* __restrict__ skip, bf16* __restrict__ y, bf16* __restrict__ pool_out,
    const bf16* __restrict__ chain_w, const float* __restrict__ chain_b,
    float* __restrict__ chain_out, Dims d) {
    using C = Conv3Cfg<CIN, COUT, UP, NBZ, NBY, NBX, NWT, MODE>;
    extern __shared__ __align__(16) unsigned char lds_raw[];
    bf16* sin = reinterpret_cast<bf16*>(lds_raw);
    bf16* sw = sin + C::IN_ELEMS;
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int r16 = lane & 15, g = lane >> 4;
    const int nbx = d.W / C::TX, nby = d.H / C::TY;
    const int bx = int(blockIdx.x) % nbx, by = (int(blockIdx.x) / nbx) % nby;
    const int bz = int(blockIdx.x) / (nbx * nby);
    const int z0 = bz * C::TZ, y0 = by * C::TY, x0 = bx * C::TX;
    const Dims din = UP ? Dims{d.D / 2, d.H / 2, d.W / 2} : d;
    // Input tile origin in input coordinates (the halo starts one voxel
    // before the tile; with UP, one coarse voxel before the parent).
    const int iz0 = (UP ? z0 / 2 : z0) - 1, iy0 = (UP ? y0 / 2 : y0) - 1,
              ix0 = (UP ? x0 / 2 : x0) - 1;

    // Stage the input tile (zero outside the volume) and the weights, every
    // load of a batch issued before its LDS stores.
    constexpr int PARTS = CIN / 8;
    constexpr int NIN = C::IZ * C::IY * C::IX * PARTS, PIN = (NIN + 255) / 256;
    {
        u32x4 buf[PIN];
#pragma unroll
        for (int j = 0; j < PIN; ++j) {
            const int i = t + j * 256;
            const int part = i % PARTS, q = i / PARTS;
            const int hx = q % C::IX, hy = (q / C::IX) % C::IY, hz = q / (C::IX * C::IY);
            const int gz = iz0 + hz, gy = iy0 + hy, gx = ix0 + hx;
            const bool inb =
                i < NIN && gz >= 0 && gz < din.D && gy >= 0 && gy < din.H && gx >= 0 && gx < din.W;
            // Branch-free: the address is clamped into the volume, the value masked.
            const int cz = min(max(gz, 0), din.D - 1), cy = min(max(gy, 0), din.H - 1),
                      cx = min(max(gx, 0), din.W - 1);
            buf[j] = *reinterpret_cast<const u32x4*>(
                x + ((size_t(cz) * din.H + cy) * din.W + cx) * CIN + part * 8);
            if (!inb) buf[j] = u32x4{0u, 0u, 0u, 0u};
        }
#pragma unroll
        for (int j = 0; j < PIN; ++j) {
            const int i = t + j * 256;
            if (i < NIN)
                *reinterpret_cast<u32x4*>(sin + (i / PARTS) * C::CSTR + (i % PARTS) * 8) = buf[j];
        }
    }
    constexpr int NW16 = C::W_ELEMS / 8, PW = (NW16 + 255) / 256, WB = 9;  // batches of 9
#pragma unroll
    for (int j0 = 0; j0 < PW; j0 += WB) {
        u32x4 buf[WB];
#pragma unroll
        for (int j = 0; j < WB; ++j) {
            const int i = t + (j0 + j) * 256;
            if (j0 + j < PW && i < NW16) buf[j] = reinterpret_cast<const u32x4*>(w)[i];
        }
#pragma unroll
        for (int j = 0; j < WB; ++j) {
            const int i = t + (j0 + j) * 256;
            if (j0 + j < PW && i < NW16) reinterpret_cast<u32x4*>(sw)[i] = buf[j];
        }
    }
    __syncthreads();

    const int ng = wave % C::NG, bgi = wave / C::NG;
    // This lane's voxel in each of its bricks (tile coordinates), and the
    // staged-tile offsets of its 3 taps per axis (d = tap offset + 1).
    int vz[C::BPW], vy[C::BPW], vx[C::BPW];
    int oz[C::BPW][3], oy[C::BPW][3], ox[C::BPW][3];
#pragma unroll
    for (int i = 0; i < C::BPW; ++i) {
        brick_voxel<NBY, NBX>(bgi * C::BPW + i, r16, vz[i], vy[i], vx[i]);
#pragma unroll
        for (int dd = 0; dd < 3; ++dd) {
            oz[i][dd] = (UP ? ((vz[i] + dd - 1) >> 1) + 1 : vz[i] + dd) * (C::IY * C::IX);
            oy[i][dd] = (UP ? ((vy[i] + dd - 1) >> 1) + 1 : vy[i] + dd) * C::IX;
            ox[i][dd] = UP ? ((vx[i] + dd - 1) >> 1) + 1 : vx[i] + dd;
        }
    }
    f4 acc[C::BPW][NWT];
#pragma unroll
    for (int i = 0; i < C::BPW; ++i)
#pragma unroll
        for (int n = 0; n < NWT; ++n) acc[i][n] = f4{0.f, 0.f, 0.f, 0.f};

    // Operands of k-step s: k = 32 s + 8 g + j, tap = k / CIN, cin = k % CIN.
    auto fetch = [&](int s, bf16x8 (&bf)[C::BPW], bf16x8 (&af)[NWT]) {
        const int k0 = 32 * s + 8 * g;
        const bool valid = k0 < 27 * CIN;  // the last step of CIN = 16 has a dummy tap
        const int tap = valid ? k0 / CIN : 26, cin0 = k0 % CIN;
        const int dz = tap / 9, dy = (tap / 3) % 3, dx = tap % 3;
#pragma unroll
        for (int i = 0; i < C::BPW; ++i) {
            bf[i] = frag_at(sin + (oz[i][dz] + oy[i][dy] + ox[i][dx]) * C::CSTR + cin0);
            if (!valid) bf[i] = bf16x8{};
        }
#pragma unroll
        for (int n = 0; n < NWT; ++n) {
            const int co = (ng * NWT + n) * 16 + r16;
            af[n] = frag_at(sw + (size_t(tap) * COUT + co) * CIN + cin0);
            if (!valid) af[n] = bf16x8{};
        }
    };
    // Fully unrolled, operands fetched one step ahead of their MFMAs.
    bf16x8 bcur[C::BPW], acur[NWT];
    fetch(0, bcur, acur);
#pragma unroll
    for (int s = 0; s < C::KSTEPS; ++s) {
        bf16x8 bnxt[C::BPW], anxt[NWT];
        if (s + 1 < C::KSTEPS) fetch(s + 1, bnxt, anxt);
#pragma unroll
        for (int i = 0; i < C::BPW; ++i)
#pragma unroll
            for (int n = 0; n < NWT; ++n)
                acc[i][n] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(acur[n], bcur[i], acc[i][n], 0, 0, 0);
        if (s + 1 < C::KSTEPS) {
#pragma unroll
            for (int i = 0; i < C::BPW; ++i) bcur[i] = bnxt[i];
#pragma unroll
            for (int n = 0; n < NWT; ++n) acur[n] = anxt[n];
        }
    }

    // Epilogue.  Lane (r16, g) holds channels 16 n' + 4 g + e of voxel r16
    // of each brick (n' = ng * NWT + n).
    if constexpr (MODE == kModePool) __syncthreads();  // the input tile becomes the output tile
#pragma unroll
    for (int i = 0; i < C::BPW; ++i) {
        const size_t v = (size_t(z0 + vz[i]) * d.H + y0 + vy[i]) * d.W + x0 + vx[i];
        u32x2 packed[NWT];
#pragma unroll
        for (int n = 0; n < NWT; ++n) {
            const int c0 = (ng * NWT + n) * 16 + 4 * g;
            u32x2 sk = {0u, 0u};
            if constexpr (MODE == kModeChain)
                sk = *reinterpret_cast<const u32x2*>(skip + v * COUT + c0);
            bf16 o[4];
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                float r = fmaxf(acc[i][n][e] + bias[c0 + e], 0.f);
                if constexpr (MODE == kModeChain) {
                    const uint32_t word = e < 2 ? sk.x : sk.y;
                    r += bf2f(bf16((e & 1) ? (word >> 16) : (word & 0xffffu)));
                }
                o[e] = f2bf(r);
            }
            packed[n].x = uint32_t(o[0]) | (uint32_t(o[1]) << 16);
            packed[n].y = uint32_t(o[2]) | (uint32_t(o[3]) << 16);
            if constexpr (MODE != kModeChain)
                *reinterpret_cast<u32x2*>(y + v * COUT + c0) = packed[n];
            if constexpr (MODE == kModePool)
                *reinterpret_cast<u32x2*>(
                    sin + ((vz[i] * C::TY + vy[i]) * C::TX + vx[i]) * COUT + c0) = packed[n];
        }
        if constexpr (MODE == kModeChain) {
            // d = bf16(relu + skip) as the 1x1x1 layer's patch operand:
            // element j < 4 is channel 4 g + j, j >= 4 channel 16 + 4 g + j - 4;
            // the weight operand (row r16 = output channel) uses the same
            // channel order, so every k pairs the same channel.
            const bf16x8 bop = (bf16x8)(u32x4{packed[0].x, packed[0].y, packed[1].x, packed[1].y});
            const bf16* wr = chain_w + r16 * COUT;
            const u32x2 wlo = *reinterpret_cast<const u32x2*>(wr + 4 * g);
            const u32x2 whi = *reinterpret_cast<const u32x2*>(wr + 16 + 4 * g);
            const bf16x8 aop = (bf16x8)(u32x4{wlo.x, wlo.y, whi.x, whi.y});
            const f4 c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(aop, bop, f4{0.f, 0.f, 0.f, 0.f},
                                                                 0, 0, 0);
            f4 o;
#pragma unroll
            for (int e = 0; e < 4; ++e) o[e] = fmaxf(c[e] + chain_b[4 * g + e], 0.f);
            *reinterpret_cast<f4*>(chain_out + v * 16 + 4 * g) = o;
        }
    }
    if constexpr (MODE == kModePool) {
        __syncthreads();
        constexpr int PZ = C::TZ / 2, PY = C::TY / 2, PX = C::TX / 2;
        const Dims dp{d.D / 2, d.H / 2, d.W / 2};
        for (int i = t; i < PZ * PY * PX * COUT; i += 256) {
            const int c = i % COUT, pv = i / COUT;
            const int px = pv % PX, py = (pv / PX) % PY, pz = pv / (PX * PY);
            float s = 0.f;
#pragma unroll
            for (int a = 0; a < 2; ++a)
#pragma unroll
                for (int b = 0; b < 2; ++b)
#pragma unroll
                    for (int e = 0; e < 2; ++e)
                        s += bf2f(sin[(((2 * pz + a) * C::TY + 2 * py + b) * C::TX + 2 * px + e) *
                                          COUT + c]);
            const size_t pvox =
                (size_t(z0 / 2 + pz) * dp.H + y0 / 2 + py) * dp.W + x0 / 2 + px;
            pool_out[pvox * COUT + c] = f2bf(s * 0.125f);
        }
    }
}

// ---------------------------------------------------------------- dec0
// Block: a (4, 4, 8) tile = 8 bricks, two per wave.  The staged input is
// d1 = bf16(c(parent) + e1) with its halo (zero outside the volume).  The
// 16 -> 1 layer is a K = 432 (+16 zero) GEMM on v_mfma_f32_16x16x32_bf16
// whose weight operand has two live rows: row 0 the BF16 rounding hi of the
// FP32 weights, row 1 the BF16 rounding of w - hi, so out = D[0] + D[1]
// carries the FP32 weights to 2^-17 (d1 is BF16-exact, so every product is
// exact).  ws: [2][448] BF16 (hi row, lo row; k = tap * 16 + cin).
// y = (D[0] + D[1] + bias) + x.  Requires D % 4 == H % 4 == W % 8 == 0.
constexpr int kD0Z = 4, kD0Y = 4, kD0X = 8;
constexpr int kD0K = 448;

__global__ __launch_bounds__(256) void dec0_kernel(const float* __restrict__ c1,
                                                   const bf16* __restrict__ e1,
                                                   const bf16* __restrict__ ws,
                                                   const float* __restrict__ bias,
                                                   const float* __restrict__ in,
                                                   const XForm* __restrict__ xf,
                                                   const float* __restrict__ mm_part,
                                                   uint32_t mm_count, float* __restrict__ out,
                                                   Dims d) {
    constexpr int HZ = kD0Z + 2, HY = kD0Y + 2, HX = kD0X + 2;
    constexpr int CS = 16 + 8;  // bf16 per staged voxel (padded)
    __shared__ __align__(16) bf16 dt[HZ * HY * HX * CS];
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int r16 = lane & 15, g = lane >> 4;
    const int bx = int(blockIdx.x) % (d.W / kD0X);
    const int by = (int(blockIdx.x) / (d.W / kD0X)) % (d.H / kD0Y);
    const int bz = int(blockIdx.x) / ((d.W / kD0X) * (d.H / kD0Y));
    const int z0 = bz * kD0Z, y0 = by * kD0Y, x0 = bx * kD0X;
    const Dims dc{d.D / 2, d.H / 2, d.W / 2};
    constexpr int KS = kD0K / 32;
    __shared__ __align__(16) bf16 wl[2 * kD0K];
    for (int i = t; i < 2 * kD0K / 8; i += 256)
        reinterpret_cast<u32x4*>(wl)[i] = reinterpret_cast<const u32x4*>(ws)[i];
    int vz[2], vy[2], vx[2];
#pragma unroll
    for (int i = 0; i < 2; ++i) brick_voxel<2, 2>(2 * wave + i, r16, vz[i], vy[i], vx[i]);
    // The raw input at this lane's two voxels (for y = ... + x), issued early.
    float xin[2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
        xin[i] = in[(size_t(z0 + vz[i]) * d.H + y0 + vy[i]) * d.W + x0 + vx[i]];
    // Stage d1; all loads of the tile issued before the first LDS store.
    // Branch-free: the address is clamped into the volume, the value masked.
    constexpr int NS = HZ * HY * HX * 2, PS = (NS + 255) / 256;
    u32x4 ev[PS];
    f4 ca[PS], cb[PS];
    bool inb[PS];
#pragma unroll
    for (int j = 0; j < PS; ++j) {
        const int i = t + j * 256;
        const int half = i & 1, q = i >> 1;
        const int hx = q % HX, hy = (q / HX) % HY, hz = q / (HX * HY);
        const int gz = z0 + hz - 1, gy = y0 + hy - 1, gx = x0 + hx - 1;
        inb[j] = i < NS && gz >= 0 && gz < d.D && gy >= 0 && gy < d.H && gx >= 0 && gx < d.W;
        const int cz = min(max(gz, 0), d.D - 1), cy = min(max(gy, 0), d.H - 1),
                  cx = min(max(gx, 0), d.W - 1);
        const size_t v = (size_t(cz) * d.H + cy) * d.W + cx;
        const size_t p = (size_t(cz >> 1) * dc.H + (cy >> 1)) * dc.W + (cx >> 1);
        ev[j] = *reinterpret_cast<const u32x4*>(e1 + v * 16 + half * 8);
        ca[j] = *reinterpret_cast<const f4*>(c1 + p * 16 + half * 8);
        cb[j] = *reinterpret_cast<const f4*>(c1 + p * 16 + half * 8 + 4);
    }
#pragma unroll
    for (int j = 0; j < PS; ++j) {
        const int i = t + j * 256;
        if (i >= NS) continue;
        u32x4 pk = {0u, 0u, 0u, 0u};
        if (inb[j]) {
            uint32_t o[4];
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const uint32_t word = ev[j][k];
                const float lo = (k < 2 ? ca[j][2 * k] : cb[j][2 * k - 4]) + bf2f(bf16(word & 0xffffu));
                const float hi = (k < 2 ? ca[j][2 * k + 1] : cb[j][2 * k - 3]) + bf2f(bf16(word >> 16));
                o[k] = uint32_t(f2bf(lo)) | (uint32_t(f2bf(hi)) << 16);
            }
            pk = u32x4{o[0], o[1], o[2], o[3]};
        }
        *reinterpret_cast<u32x4*>(dt + (i >> 1) * CS + (i & 1) * 8) = pk;
    }
    __syncthreads();
    const XForm tf = load_xform(xf, mm_part, mm_count);
    f4 acc[2] = {f4{0.f, 0.f, 0.f, 0.f}, f4{0.f, 0.f, 0.f, 0.f}};
#pragma unroll
    for (int s = 0; s < KS; ++s) {
        const int tap = min(2 * s + (g >> 1), 26), ci0 = (g & 1) * 8;
        const int dz = tap / 9, dy = (tap / 3) % 3, dx = tap % 3;
        bf16x8 a = frag_at(wl + min(r16, 1) * kD0K + 32 * s + 8 * g);
        if (r16 >= 2) a = bf16x8{};
#pragma unroll
        for (int i = 0; i < 2; ++i) {
            const bf16x8 b =
                frag_at(dt + (((vz[i] + dz) * HY + vy[i] + dy) * HX + vx[i] + dx) * CS + ci0);
            acc[i] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc[i], 0, 0, 0);
        }
    }
    if (g != 0) return;  // rows 0 and 1 live in lanes 0..15
#pragma unroll
    for (int i = 0; i < 2; ++i) {
        const size_t v = (size_t(z0 + vz[i]) * d.H + y0 + vy[i]) * d.W + x0 + vx[i];
        out[v] = ((acc[i][0] + acc[i][1]) + bias[0]) + xform_apply(xin[i], tf);
    }
}

}  // namespace nn
}  // namespace fbp
