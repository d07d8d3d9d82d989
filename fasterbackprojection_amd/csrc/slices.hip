// Depth-plane slice images and their PNG files.
//
// Reference: Laser::reconstructDepthConfocal / Exhaustive write one PNG per
// depth plane (FB/Laser.cu:410-424, 463-476) through TransientImage::save
// (FB/TransientImage.cpp:27-45), which uses Image<uchar>::flipImageVertically,
// resize and interpolate (FB/Image.h:45-67, 93-137, 161-177):
//   g(p)  = clamp(uint(a[p * D + d] * 255), 0, 255)     pixel p = r W + c
//   RGBA  = (g, g, g, 255), rows flipped (r -> H-1-r)
//   when the output size W' x H' differs from W x H, every channel is
//   resampled bilinearly: xs = float(x) * ((W-1)/(W'-1)), x0 = int(xs),
//   x1 = min(x0+1, W-1), dx = xs - x0 (same along y), and
//   out = uchar(((1-dx)(1-dy) p00 + dx(1-dy) p01) + (1-dx) dy p10) + dx dy p11)
//   in FP32, products and sums in this order (so the interpolated alpha can
//   read 254 where the weights round below 1).
// Values that are negative or NaN (min-max normalised activations are
// neither) map to 0.
//
// Two HBM passes: slice_bytes_kernel transposes the depth-fastest activation
// [P][D] into flipped byte planes [D][H][W] through LDS (coalesced reads of
// the depth rows, coalesced byte writes of the pixel rows); slice_rgba_kernel
// writes one RGBA word per output pixel from the byte planes.  The PNG
// encoder runs on the host (zlib), one file per plane.
#include <zlib.h>

#include <cstdio>
#include <cstring>
#include <vector>

#include "fbp_internal.h"

namespace fbp {

namespace {

constexpr int kTileP = 64;  // pixels per transpose tile
constexpr int kTileD = 64;  // depths per transpose tile

__device__ __forceinline__ uint8_t gray_byte(float v) {
    const float s = v * 255.0f;
    if (!(s > 0.f)) return 0;  // negative, zero or NaN
    if (s >= 255.f) return 255;
    return uint8_t(uint32_t(s));  // truncation toward zero
}

// Grid (ceil(P / 64), ceil(D / 64)), 256 threads.
__global__ __launch_bounds__(256) void slice_bytes_kernel(const float* __restrict__ act, uint32_t W,
                                                          uint32_t H, uint32_t D,
                                                          uint8_t* __restrict__ bytes) {
    __shared__ uint8_t tile[kTileD][kTileP + 4];
    const uint32_t P = W * H;
    const uint32_t p0 = blockIdx.x * kTileP, d0 = blockIdx.y * kTileD;
    const uint32_t t = threadIdx.x;
    // Read: 4 pixels per pass, 64 consecutive depths of one pixel per wave.
    for (uint32_t i = t / kTileD; i < kTileP; i += 256 / kTileD) {
        const uint32_t p = p0 + i, d = d0 + (t % kTileD);
        tile[t % kTileD][i] = (p < P && d < D) ? gray_byte(act[size_t(p) * D + d]) : uint8_t(0);
    }
    __syncthreads();
    // Write: 64 consecutive pixels of one depth per wave, rows flipped.
    for (uint32_t j = t / kTileP; j < kTileD; j += 256 / kTileP) {
        const uint32_t p = p0 + (t % kTileP), d = d0 + j;
        if (p >= P || d >= D) continue;
        const uint32_t r = p / W, c = p - r * W;
        bytes[size_t(d) * P + size_t(H - 1 - r) * W + c] = tile[j][t % kTileP];
    }
}

// Bilinear sample of Image<uchar>::interpolate (FB/Image.h:45-67).
__device__ __forceinline__ uint8_t interp(float p00, float p01, float p10, float p11, float dx,
                                          float dy) {
    const float v = (1.0f - dx) * (1.0f - dy) * p00 + dx * (1.0f - dy) * p01 +
                    (1.0f - dx) * dy * p10 + dx * dy * p11;
    return uint8_t(int(v));  // v in [0, 255]: truncation
}

// One thread per output pixel: grid (ceil(OW*OH / 256), D).
__global__ __launch_bounds__(256) void slice_rgba_kernel(const uint8_t* __restrict__ bytes,
                                                         uint32_t W, uint32_t H, uint32_t OW,
                                                         uint32_t OH, int resize, float xr,
                                                         float yr, uint32_t* __restrict__ rgba) {
    const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= OW * OH) return;
    const uint32_t d = blockIdx.y;
    const uint8_t* plane = bytes + size_t(d) * W * H;
    const uint32_t y = q / OW, x = q - y * OW;
    uint32_t g, a = 255u;
    if (!resize) {
        g = plane[q];
    } else {
        const float xs = float(int(x)) * xr, ys = float(int(y)) * yr;
        const int x0 = int(xs), y0 = int(ys);
        const int x1 = min(x0 + 1, int(W) - 1), y1 = min(y0 + 1, int(H) - 1);
        const float dx = xs - float(x0), dy = ys - float(y0);
        const float p00 = plane[y0 * W + x0], p01 = plane[y0 * W + x1];
        const float p10 = plane[y1 * W + x0], p11 = plane[y1 * W + x1];
        g = interp(p00, p01, p10, p11, dx, dy);
        a = interp(255.f, 255.f, 255.f, 255.f, dx, dy);
    }
    rgba[size_t(d) * OW * OH + q] = g | (g << 8) | (g << 16) | (a << 24);
}

void put_be32(std::vector<uint8_t>& out, uint32_t v) {
    out.push_back(uint8_t(v >> 24));
    out.push_back(uint8_t(v >> 16));
    out.push_back(uint8_t(v >> 8));
    out.push_back(uint8_t(v));
}

// One PNG chunk: length, type, data, CRC-32 of type + data.
void put_chunk(std::vector<uint8_t>& out, const char type[4], const uint8_t* data, size_t n) {
    put_be32(out, uint32_t(n));
    const size_t at = out.size();
    out.insert(out.end(), type, type + 4);
    if (n) out.insert(out.end(), data, data + n);
    const uLong crc = crc32(0L, out.data() + at, uInt(n + 4));
    put_be32(out, uint32_t(crc));
}

}  // namespace

}  // namespace fbp

using namespace fbp;

extern "C" {

fbp_status fbp_depth_slices_rgba(fbp_context ctx, const float* activation, uint32_t width,
                                 uint32_t height, uint32_t num_depths, uint32_t out_width,
                                 uint32_t out_height, uint8_t* rgba, void* stream) {
    if (!ctx || !activation || !rgba || width == 0 || height == 0 || num_depths == 0 ||
        out_width == 0 || out_height == 0)
        return fail(ctx, FBP_ERR_INVALID_ARGUMENT, "depth_slices_rgba: bad argument");
    // Image sizes are 16-bit in the reference (FB/Image.h:17).
    if (width > 65535u || height > 65535u || out_width > 65535u || out_height > 65535u)
        return fail(ctx, FBP_ERR_INVALID_ARGUMENT, "depth_slices_rgba: image edge above 65535");
    const bool resize = out_width != width || out_height != height;
    // The resampling ratio (W-1)/(W'-1) needs W' > 1 and H' > 1.
    if (resize && (out_width < 2 || out_height < 2))
        return fail(ctx, FBP_ERR_INVALID_ARGUMENT,
                    "depth_slices_rgba: resized images need at least 2 x 2 pixels");
    if (reinterpret_cast<uintptr_t>(rgba) % 4 != 0)
        return fail(ctx, FBP_ERR_INVALID_ARGUMENT, "depth_slices_rgba: rgba must be 4-byte aligned");
    if (num_depths > 65535u)
        return fail(ctx, FBP_ERR_INVALID_ARGUMENT, "depth_slices_rgba: at most 65535 planes");
    FBP_CHECK(bind(ctx));
    hipStream_t s = as_stream(stream);
    const size_t P = size_t(width) * height;
    const size_t nbytes = P * num_depths;
    FBP_CHECK(ensure_scratch(ctx, nbytes));
    uint8_t* bytes = static_cast<uint8_t*>(ctx->scratch);
    const dim3 g1(unsigned((P + kTileP - 1) / kTileP), (num_depths + kTileD - 1) / kTileD);
    hipLaunchKernelGGL(slice_bytes_kernel, g1, dim3(256), 0, s, activation, width, height,
                       num_depths, bytes);
    FBP_CHECK_HIP(ctx, hipGetLastError());
    const float xr = resize ? float(width - 1) / float(out_width - 1) : 1.f;
    const float yr = resize ? float(height - 1) / float(out_height - 1) : 1.f;
    const size_t Q = size_t(out_width) * out_height;
    const dim3 g2(unsigned((Q + 255) / 256), num_depths);
    hipLaunchKernelGGL(slice_rgba_kernel, g2, dim3(256), 0, s, bytes, width, height, out_width,
                       out_height, resize ? 1 : 0, xr, yr, reinterpret_cast<uint32_t*>(rgba));
    FBP_CHECK_HIP(ctx, hipGetLastError());
    return FBP_OK;
}

fbp_status fbp_write_png_rgba(const char* path, const uint8_t* rgba, uint32_t width,
                              uint32_t height) {
    if (!path || !rgba || width == 0 || height == 0 || width > 65535u || height > 65535u)
        return fail(nullptr, FBP_ERR_INVALID_ARGUMENT, "write_png_rgba: bad argument");
    // Scanlines with filter type 0 (none), 8-bit RGBA (colour type 6).
    const size_t row = size_t(width) * 4;
    std::vector<uint8_t> raw(height * (row + 1));
    for (uint32_t y = 0; y < height; ++y) {
        raw[y * (row + 1)] = 0;
        std::memcpy(&raw[y * (row + 1) + 1], rgba + y * row, row);
    }
    uLongf zn = compressBound(uLong(raw.size()));
    std::vector<uint8_t> z(zn);
    if (compress2(z.data(), &zn, raw.data(), uLong(raw.size()), 6) != Z_OK)
        return fail(nullptr, FBP_ERR_IO, "write_png_rgba: zlib compression failed");
    std::vector<uint8_t> out = {0x89, 'P', 'N', 'G', '\r', '\n', 0x1a, '\n'};
    uint8_t ihdr[13];
    const uint32_t wh[2] = {width, height};
    for (int i = 0; i < 2; ++i)
        for (int b = 0; b < 4; ++b) ihdr[4 * i + b] = uint8_t(wh[i] >> (24 - 8 * b));
    ihdr[8] = 8;   // bit depth
    ihdr[9] = 6;   // RGBA
    ihdr[10] = 0;  // deflate
    ihdr[11] = 0;  // adaptive filtering (per-row filter byte)
    ihdr[12] = 0;  // no interlace
    put_chunk(out, "IHDR", ihdr, sizeof(ihdr));
    put_chunk(out, "IDAT", z.data(), zn);
    put_chunk(out, "IEND", nullptr, 0);
    FILE* f = std::fopen(path, "wb");
    if (!f) return fail(nullptr, FBP_ERR_IO, "write_png_rgba: cannot open %s", path);
    const bool ok = std::fwrite(out.data(), 1, out.size(), f) == out.size();
    if (std::fclose(f) != 0 || !ok)
        return fail(nullptr, FBP_ERR_IO, "write_png_rgba: write failed for %s", path);
    return FBP_OK;
}

}  // extern "C"
