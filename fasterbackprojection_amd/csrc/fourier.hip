// Frequency-domain paths on rocFFT: temporal band-pass filter_H, LoGFFT and
// the confocal LCT.  Plans, the filter_H kernel spectrum, the LCT inverse
// filter and resampling operators are cached on the context.
#include <rocfft/rocfft.h>

#include <cfloat>
#include <cmath>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "chain.h"
#include "fbp_internal.h"

namespace fbp {

namespace {

struct FftPlan {
    rocfft_plan plan = nullptr;
    size_t work = 0;
    uint64_t last_use = 0;
};

// Plans kept per context; beyond this the least recently used one is
// destroyed (after a device synchronisation, so no transform of it is still
// queued), bounding the device memory plans hold for callers that vary
// their batch sizes.
constexpr size_t kMaxPlans = 24;

struct Csr {  // M x M sparse operator on the device
    int* rowptr = nullptr;
    int* cols = nullptr;
    float* vals = nullptr;
};

struct FftCache {
    std::map<std::string, FftPlan> plans;
    uint64_t plan_clock = 0;
    void* work = nullptr;
    size_t work_bytes = 0;
    rocfft_execution_info info = nullptr;
    // filter_H: half spectrum of Re(K) (complex float, N/2 + 1)
    std::string khat_key;
    float2* khat = nullptr;
    // LCT: half spectrum of the inverse filter (complex float), operators
    std::string inv_key;
    float2* invpsf = nullptr;
    uint32_t op_m = 0;
    Csr mtx, mtxi;
    // LCT z-pass twiddles exp(-2 pi i k / (2M)), k < M
    uint32_t tw_m = 0;
    float2* tw = nullptr;
    // The inverse filter in the 2048-point z-pass's load order (invt_tile_kernel)
    float2* invt = nullptr;
    // The PSF is mirror-symmetric in y (psf[z][2N-1-y][x] == psf[z][y][x]
    // bit for bit), so inv(kx, 2N - k) = exp(2 pi i k / 2N) inv(kx, k): the
    // 2048-point z-pass keeps the tiles of rows ky <= N only.
    int inv_sym = 0;
    // exp(-2 pi i k / 256), k < 256 (the 256-point y-pass)
    float2* tw256 = nullptr;
};

std::once_flag g_rocfft_once;
rocfft_status g_rocfft_setup = rocfft_status_success;

FftCache* cache_of(fbp_context ctx) {
    if (!ctx->fft_cache) ctx->fft_cache = new FftCache();
    return static_cast<FftCache*>(ctx->fft_cache);
}

fbp_status fft_fail(fbp_context ctx, rocfft_status st, const char* what) {
    return fail(ctx, FBP_ERR_FFT, "%s: rocfft status %d", what, int(st));
}

fbp_status get_plan(fbp_context ctx, rocfft_transform_type type, rocfft_precision prec,
                    std::vector<size_t> lengths, size_t batch, FftPlan** out) {
    std::call_once(g_rocfft_once, [] { g_rocfft_setup = rocfft_setup(); });
    if (g_rocfft_setup != rocfft_status_success) return fft_fail(ctx, g_rocfft_setup, "rocfft_setup");
    FftCache* c = cache_of(ctx);
    std::string key = std::to_string(int(type)) + "/" + std::to_string(int(prec)) + "/" +
                      std::to_string(batch);
    for (size_t l : lengths) key += "x" + std::to_string(l);
    auto it = c->plans.find(key);
    if (it == c->plans.end()) {
        if (c->plans.size() >= kMaxPlans) {
            auto lru = c->plans.begin();
            for (auto j = c->plans.begin(); j != c->plans.end(); ++j)
                if (j->second.last_use < lru->second.last_use) lru = j;
            FBP_CHECK_HIP(ctx, hipDeviceSynchronize());
            rocfft_plan_destroy(lru->second.plan);
            c->plans.erase(lru);
        }
        FftPlan p;
        rocfft_status st = rocfft_plan_create(&p.plan, rocfft_placement_notinplace, type, prec,
                                              lengths.size(), lengths.data(), batch, nullptr);
        if (st != rocfft_status_success) return fft_fail(ctx, st, "rocfft_plan_create");
        st = rocfft_plan_get_work_buffer_size(p.plan, &p.work);
        if (st != rocfft_status_success) return fft_fail(ctx, st, "rocfft_plan_get_work_buffer_size");
        it = c->plans.emplace(key, p).first;
    }
    it->second.last_use = ++c->plan_clock;
    *out = &it->second;
    return FBP_OK;
}

fbp_status run_plan(fbp_context ctx, FftPlan* p, void* in, void* out, hipStream_t s) {
    FftCache* c = cache_of(ctx);
    if (!c->info) {
        rocfft_status st = rocfft_execution_info_create(&c->info);
        if (st != rocfft_status_success) return fft_fail(ctx, st, "rocfft_execution_info_create");
    }
    if (p->work > c->work_bytes) {
        FBP_CHECK_HIP(ctx, hipStreamSynchronize(s));
        if (c->work) FBP_CHECK_HIP(ctx, hipFree(c->work));
        c->work = nullptr;
        FBP_CHECK_HIP(ctx, hipMalloc(&c->work, p->work));
        c->work_bytes = p->work;
    }
    rocfft_status st = rocfft_execution_info_set_stream(c->info, s);
    if (st != rocfft_status_success) return fft_fail(ctx, st, "rocfft_execution_info_set_stream");
    if (c->work_bytes > 0) {
        st = rocfft_execution_info_set_work_buffer(c->info, c->work, c->work_bytes);
        if (st != rocfft_status_success)
            return fft_fail(ctx, st, "rocfft_execution_info_set_work_buffer");
    }
    void* ib[1] = {in};
    void* ob[1] = {out};
    st = rocfft_execute(p->plan, ib, ob, c->info);
    if (st != rocfft_status_success) return fft_fail(ctx, st, "rocfft_execute");
    return FBP_OK;
}

inline dim3 grid1(size_t n, unsigned bs = 256) { return dim3(unsigned((n + bs - 1) / bs)); }

// ------------------------------------------------------------------ filter_H

// Laser::filter_H_cuda's length rules (FB/Laser.cu:128-137), float arithmetic:
// t6 = round(6 sigma / dt), bumped to even (so 1 becomes 2: nt_pad = nt + 2,
// one bin of padding each side).  Returns false when t6 is still below 2
// after the bump (t6 <= 0: no padding, nt_pad <= nt) or the padded length
// exceeds 2^24 samples.
bool filter_h_lengths(uint32_t nt, float dt, float wl_mean, float& wl_sigma, uint32_t& nt_pad,
                      uint32_t& pad) {
    if (std::fabs(wl_sigma - 0.0f) < FLT_EPSILON) wl_sigma = wl_mean / sqrtf(2.0f);
    const double t6f = std::round(6.0 * double(wl_sigma) / double(dt));
    nt_pad = pad = 0;
    // NaN, negative and huge ratios end here, before the int conversion.
    if (!(t6f >= 0.0) || t6f > double(1 << 24)) return false;
    int t6 = static_cast<int>(std::round(6 * wl_sigma / dt));
    if (t6 % 2 == 1) t6 += 1;
    if (t6 < 2 || uint64_t(nt) + 2u * uint64_t(t6 - 1) > (uint64_t(1) << 24)) return false;
    nt_pad = nt + 2u * uint32_t(t6 - 1);
    pad = (nt_pad - nt) / 2u;
    return true;
}

// Transform length for the N-point circular convolution of filter_H.  N
// itself when it factors into 2, 3, 5, 7, 11 and 13 (rocFFT's radices);
// otherwise (N = 1106 = 2 * 7 * 79 for 1024 bins) the smallest such M >=
// N + R: K's Gaussian envelope (sigma = wl_sigma) is below 1e-30 of its peak
// beyond R = 2 t6 bins, so on the cropped samples the M-point convolution
// with K's taps |lag| <= R equals the reference's N-point circular one to
// FP32 rounding, and no Bluestein transform is needed.
bool smooth_length(uint32_t n) {
    for (uint32_t p : {2u, 3u, 5u, 7u, 11u, 13u})
        while (n % p == 0) n /= p;
    return n == 1;
}
uint32_t filter_h_transform_length(uint32_t N, uint32_t R) {
    if (smooth_length(N)) return N;
    uint32_t M = N + R;
    while (!smooth_length(M)) ++M;
    return M;
}

// Rows of M >= N floats: the padded series in [0, N), zeros after it.
__global__ void fh_pad_kernel(const float* __restrict__ data, float* __restrict__ real, uint64_t batch,
                              uint32_t nt, uint32_t N, uint32_t M, uint32_t pad, int edge) {
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= batch * M) return;
    const size_t b = i / M;
    const uint32_t t = uint32_t(i % M);
    const float* row = data + b * nt;
    float v;
    if (t >= N)
        v = 0.f;
    else if (t < pad)
        v = edge ? row[0] : 0.f;
    else if (t >= pad + nt)
        v = edge ? row[nt - 1] : 0.f;
    else
        v = row[t - pad];
    real[i] = v;
}

// H[b][f] *= K[f]  (half spectra; multiplyHK, FB/fourier.cuh:17-31)
__global__ void fh_mul_kernel(float2* __restrict__ H, const float2* __restrict__ K, uint64_t batch,
                              uint32_t nh) {
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= batch * nh) return;
    const float2 h = H[i], k = K[i % nh];
    H[i] = make_float2(h.x * k.x - h.y * k.y, h.x * k.y + h.y * k.x);
}

// scale: N / M when the transform is longer than the padded series (1 when
// M = N: the reference's unnormalised N-point inverse).
__global__ void fh_unpad_kernel(const float* __restrict__ real, float* __restrict__ data,
                                uint64_t batch, uint32_t nt, uint32_t M, uint32_t pad,
                                float scale) {
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= batch * nt) return;
    const size_t b = i / nt;
    const float v = real[b * M + pad + (i % nt)];
    data[i] = scale == 1.0f ? v : v * scale;
}

// --------------------------------------------------------------------- LoGFFT
// buildLoGKernel3D + multiplyKernel (FB/transient_postprocessing.cuh:97-147),
// with x the halved (fastest) axis.
__global__ void logfft_mul_kernel(float2* __restrict__ C, uint32_t X, uint32_t Y, uint32_t Z,
                                  float sigma) {
    const uint32_t xh = X / 2 + 1;
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= size_t(xh) * Y * Z) return;
    const int x = int(i % xh), y = int((i / xh) % Y), z = int(i / (size_t(xh) * Y));
    const float PI = 3.14159265358979323846f;
    const float fx = (x <= int(X) / 2) ? float(x) : float(x - int(X));
    const float fy = (y <= int(Y) / 2) ? float(y) : float(y - int(Y));
    const float fz = (z <= int(Z) / 2) ? float(z) : float(z - int(Z));
    const float fxn = fx / float(X), fyn = fy / float(Y), fzn = fz / float(Z);
    const float k2 = fxn * fxn + fyn * fyn + fzn * fzn;
    const float lap = -4.0f * PI * PI * k2;
    const float gauss = expf(-2.0f * PI * PI * sigma * sigma * k2);
    const float val = lap * gauss;
    const float2 a = C[i];
    C[i] = make_float2(a.x * val - a.y * 0.0f, a.x * 0.0f + a.y * val);
}

// ------------------------------------------------------------------------ LCT
// python/lct.py:109-211.  Layouts: rect [x][y][t] (N, N, M); work volumes
// [t][y][x]; padded FFT volume [2M][2N][2N] real, half spectrum
// [2M][2N][N+1] complex.

// Sum of the 2^K values src[0 .. 2^K) as the balanced tree that K rounds of
// python/lct.py:103 (rect[:, :, 0::2] + rect[:, :, 1::2]) build: left
// subtree + right subtree at every level.  The stack holds the pending left
// subtree of each level (a binary carry over the element index).
__device__ inline float lct_tree_sum(const float* __restrict__ src, uint32_t K) {
    if (K == 0) return src[0];
    float stack[17];
    const uint32_t n = 1u << K;
    for (uint32_t i = 0; i < n; ++i) {
        float v = src[i];
        uint32_t lv = 0;
        for (; (i >> lv) & 1u; ++lv) v = stack[lv] + v;
        stack[lv] = v;
    }
    return stack[K];
}

// Tiled transpose rect[x][y][t_in] -> d[t][y][x] (t_in = t * 2^K .. (t+1) *
// 2^K - 1 summed by lct_tree_sum) with z-scaling and trimming; M is the
// downsampled bin count.
// inv_m1 = 1 / (M - 1) in double (the host's IEEE division, passed in so
// that no thread divides).
__global__ void lct_permute_kernel(const float* __restrict__ rect, float* __restrict__ d, uint32_t N,
                                   uint32_t M, uint32_t K, uint32_t z_trim, int power4,
                                   double inv_m1) {
    __shared__ float tile[32][33];
    const uint32_t y = blockIdx.z;
    const uint32_t t0 = blockIdx.x * 32, x0 = blockIdx.y * 32;
    const size_t m_in = size_t(M) << K;
    // load: threads along t (coalesced in rect for K = 0)
    for (int r = threadIdx.y; r < 32; r += blockDim.y) {
        const uint32_t x = x0 + r, t = t0 + threadIdx.x;
        float v = 0.f;
        if (x < N && t < M) v = lct_tree_sum(rect + (size_t(x) * N + y) * m_in + (size_t(t) << K), K);
        tile[r][threadIdx.x] = v;
    }
    __syncthreads();
    for (int r = threadIdx.y; r < 32; r += blockDim.y) {
        const uint32_t t = t0 + r, x = x0 + threadIdx.x;
        if (x < N && t < M) {
            float v = tile[threadIdx.x][r];
            if (t < z_trim) v = 0.f;
            // grid_z = linspace(0, 1, M)[t]
            const double gz = (M > 1) ? (t == M - 1 ? 1.0 : double(t) * inv_m1) : 0.0;
            const double g2 = gz * gz;
            v = float(double(v) * (power4 ? g2 * g2 : g2));
            d[(size_t(t) * N + y) * N + x] = v;
        }
    }
}

__global__ void fill_kernel(float* __restrict__ p, size_t n) {
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.f;
}

// out[r][p] = sum_j A[r][j] in[j][p] for pixel p < N*N; writes into the
// padded volume (row pitch 2N, plane pitch 4N^2) or a dense [M][N][N] array.
__global__ void lct_resample_kernel(const float* __restrict__ in, float* __restrict__ out,
                                    const int* __restrict__ rowptr, const int* __restrict__ cols,
                                    const float* __restrict__ vals, uint32_t N, uint32_t M,
                                    int to_padded) {
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    const size_t NN = size_t(N) * N;
    if (i >= NN * M) return;
    const uint32_t r = uint32_t(i / NN);
    const size_t p = i % NN;
    float acc = 0.f;
    for (int e = rowptr[r]; e < rowptr[r + 1]; ++e) acc = fmaf(vals[e], in[size_t(cols[e]) * NN + p], acc);
    if (to_padded) {
        const size_t y = p / N, x = p % N;
        out[(size_t(r) * 2 * N + y) * 2 * N + x] = acc;
    } else {
        out[i] = acc;
    }
}

// Padded planes z < M of the resampled volume: pad[z][y][x] (2N x 2N) =
// sum_j A[z][j] in[j][y][x] for y, x < N, zero elsewhere (no separate fill).
__global__ void lct_resample_planes_kernel(const float* __restrict__ in, float* __restrict__ pad,
                                           const int* __restrict__ rowptr,
                                           const int* __restrict__ cols,
                                           const float* __restrict__ vals, uint32_t N, uint32_t M) {
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    const size_t P = size_t(4) * N * N;
    if (i >= P * M) return;
    const uint32_t z = uint32_t(i / P);
    const uint32_t q = uint32_t(i % P), y = q / (2 * N), x = q % (2 * N);
    float acc = 0.f;
    if (y < N && x < N) {
        const size_t NN = size_t(N) * N, p = size_t(y) * N + x;
        for (int e = rowptr[z]; e < rowptr[z + 1]; ++e)
            acc = fmaf(vals[e], in[size_t(cols[e]) * NN + p], acc);
    }
    pad[i] = acc;
}

// lct_resample_planes_kernel for N a multiple of 4: four consecutive x per
// thread (16-byte loads and stores), plane z = blockIdx.y; the same sums in
// the same order.
__global__ __launch_bounds__(256) void lct_resample_planes4_kernel(
    const float* __restrict__ in, float* __restrict__ pad, const int* __restrict__ rowptr,
    const int* __restrict__ cols, const float* __restrict__ vals, uint32_t N) {
    const uint32_t z = blockIdx.y;
    const uint32_t w4 = N / 2;  // 16-byte groups per padded row of 2N
    const uint32_t q = blockIdx.x * 256u + threadIdx.x;
    if (q >= 2u * N * w4) return;
    const uint32_t y = q / w4, x = (q % w4) * 4u;
    float4 acc = make_float4(0.f, 0.f, 0.f, 0.f);
    const size_t NN = size_t(N) * N;
    if (y < N && x < N) {
        const size_t p = size_t(y) * N + x;
        const int e1 = rowptr[z + 1];
        for (int e = rowptr[z]; e < e1; ++e) {
            const float w = vals[e];
            const float4 v = *reinterpret_cast<const float4*>(in + size_t(cols[e]) * NN + p);
            acc.x = fmaf(w, v.x, acc.x);
            acc.y = fmaf(w, v.y, acc.y);
            acc.z = fmaf(w, v.z, acc.z);
            acc.w = fmaf(w, v.w, acc.w);
        }
    }
    *reinterpret_cast<float4*>(pad + size_t(z) * 4 * NN + size_t(q) * 4) = acc;
}

// ----------------------------------------------------------- LCT z-pass
// For every (ky, kx) column of the half spectrum S[z][ky][kx] of the
// planes z < M (2D transforms of the padded planes), in one workgroup:
//   X[k] = sum_{z < M} S[z] e^{-2 pi i z k / 2M}    (the upper M planes of
//          the padded volume are zero: pruned input)
//   Y[k] = X[k] * inv[k][ky][kx]                     (Wiener / inverse filter)
//   y[z] = sum_k Y[k] e^{+2 pi i z k / 2M}, z < M    (crop: only z < M kept)
// written back to S in place.  kLctCols adjacent kx columns per workgroup
// (64-byte rows per plane); the 2M x kLctCols block lives in LDS.  Forward:
// radix-2 decimation in time from bit-reversed input; inverse: decimation
// in frequency to bit-reversed output, read back through the bit reversal.
constexpr int kLctCols = 8;
constexpr int kLctThreads = 512;
// The 2048-point z-pass: 1024 threads, kZU = 128 rows of kLctCols columns.
constexpr int kZ2048Threads = 1024;
constexpr int kZU = kZ2048Threads / kLctCols;

__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

__device__ __forceinline__ uint32_t bitrev(uint32_t i, uint32_t lg) {
    return __brev(i) >> (32u - lg);
}

__global__ __launch_bounds__(kLctThreads) void lct_zpass_kernel(float2* __restrict__ S,
                                                                const float2* __restrict__ inv,
                                                                const float2* __restrict__ tw,
                                                                uint32_t N, uint32_t M,
                                                                uint32_t lg, uint32_t sp) {
    extern __shared__ float2 lds[];
    float2* twl = lds;           // M twiddles
    float2* buf = lds + M;       // [2M][kLctCols]
    const uint32_t tid = threadIdx.x;
    const uint32_t K = 2 * M, H = N + 1, Y2 = 2 * N;
    const uint32_t kx0 = blockIdx.x * kLctCols, ky = blockIdx.y;
    for (uint32_t k = tid; k < M; k += kLctThreads) twl[k] = tw[k];
    for (uint32_t idx = tid; idx < K * kLctCols; idx += kLctThreads) {
        const uint32_t i = idx / kLctCols, c = idx % kLctCols;
        const uint32_t z = bitrev(i, lg), kx = kx0 + c;
        float2 v = make_float2(0.f, 0.f);
        if (z < M && kx < H) v = S[(size_t(z) * Y2 + ky) * sp + kx];
        buf[idx] = v;
    }
    // Forward, decimation in time: butterflies of span h = 2^s.
    for (uint32_t s = 0; s < lg; ++s) {
        __syncthreads();
        for (uint32_t b = tid; b < M * kLctCols; b += kLctThreads) {
            const uint32_t j = b / kLctCols, c = b % kLctCols;
            const uint32_t pos = j & ((1u << s) - 1u);
            const uint32_t i0 = ((j >> s) << (s + 1)) | pos, i1 = i0 + (1u << s);
            const float2 w = twl[pos << (lg - 1u - s)];
            const float2 u = buf[i0 * kLctCols + c];
            const float2 v = cmul(buf[i1 * kLctCols + c], w);
            buf[i0 * kLctCols + c] = make_float2(u.x + v.x, u.y + v.y);
            buf[i1 * kLctCols + c] = make_float2(u.x - v.x, u.y - v.y);
        }
    }
    __syncthreads();
    for (uint32_t idx = tid; idx < K * kLctCols; idx += kLctThreads) {
        const uint32_t k = idx / kLctCols, kx = kx0 + idx % kLctCols;
        if (kx < H) buf[idx] = cmul(buf[idx], inv[(size_t(k) * Y2 + ky) * H + kx]);
    }
    // Inverse, decimation in frequency with conjugate twiddles.
    for (uint32_t s = lg; s-- > 0;) {
        __syncthreads();
        for (uint32_t b = tid; b < M * kLctCols; b += kLctThreads) {
            const uint32_t j = b / kLctCols, c = b % kLctCols;
            const uint32_t pos = j & ((1u << s) - 1u);
            const uint32_t i0 = ((j >> s) << (s + 1)) | pos, i1 = i0 + (1u << s);
            const float2 t = twl[pos << (lg - 1u - s)];
            const float2 w = make_float2(t.x, -t.y);
            const float2 u = buf[i0 * kLctCols + c];
            const float2 v = buf[i1 * kLctCols + c];
            buf[i0 * kLctCols + c] = make_float2(u.x + v.x, u.y + v.y);
            buf[i1 * kLctCols + c] = cmul(make_float2(u.x - v.x, u.y - v.y), w);
        }
    }
    __syncthreads();
    for (uint32_t idx = tid; idx < M * kLctCols; idx += kLctThreads) {
        const uint32_t z = idx / kLctCols, c = idx % kLctCols, kx = kx0 + c;
        if (kx < H) S[(size_t(z) * Y2 + ky) * sp + kx] = buf[bitrev(z, lg) * kLctCols + c];
    }
}

// 2M = 2048: the same z-pass with register DFTs.  Index digits (forward):
//   n = a + 128 b, k = k0 + 16 k'   (radix 16 over b; twiddle W^{a k0})
//   a = a' + 8 c,  k' = k1 + 16 k2  (radix 16 over c; twiddle W_128^{a' k1})
//   radix 8 over a'                 -> X[k0 + 16 k1 + 256 k2]
// and (inverse, digits of the held k first):
//   k = kl + 256 k2, z = z0 + 8 z'  (radix 8 over k2; twiddle W^{-kl z0})
//   kl = l0 + 16 l1, z' = z1 + 16 z2 (radix 16 over l1; twiddle W_256^{-l0 z1})
//   radix 16 over l0                -> y[z0 + 8 z1 + 128 z2], z2 < 8 kept.
// 64 threads per column; four LDS round trips (the first pass reads global
// memory, the last writes it).  LDS index pos * 8 + col + 4 (pos / 8): the
// padding spreads every pass's 64-lane pattern evenly over the banks.
constexpr float kC16[16] = {1.0f, 0.92387953251128674f, 0.70710678118654752f,
                            0.38268343236508977f, 0.0f, -0.38268343236508977f,
                            -0.70710678118654752f, -0.92387953251128674f, -1.0f,
                            -0.92387953251128674f, -0.70710678118654752f,
                            -0.38268343236508977f, 0.0f, 0.38268343236508977f,
                            0.70710678118654752f, 0.92387953251128674f};

// W_16^m = cos(2 pi m / 16) - i sin(2 pi m / 16) (conjugated when INV).
template <bool INV>
__device__ __forceinline__ float2 w16(int m) {
    const float c = kC16[m & 15], sn = kC16[(m + 12) & 15];  // sin = cos(x - pi/2)
    return make_float2(c, INV ? sn : -sn);
}

// a W16^m (W16 = exp(-2 pi i / 16); INV: the conjugate); the quarter turns
// m = 4, 8, 12 as the exact swaps and negations their products give.
template <bool INV>
__device__ __forceinline__ float2 mul_w16(float2 a, int m) {
    switch (m & 15) {
        case 0: return a;
        case 4: return INV ? make_float2(-a.y, a.x) : make_float2(a.y, -a.x);
        case 8: return make_float2(-a.x, -a.y);
        case 12: return INV ? make_float2(a.y, -a.x) : make_float2(-a.y, a.x);
        default: return cmul(a, w16<INV>(m));
    }
}

__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }

template <int LG>
constexpr int brev_c(int i) {
    int r = 0;
    for (int b = 0; b < LG; ++b) r |= ((i >> b) & 1) << (LG - 1 - b);
    return r;
}

// In-register DFT of N = 8 or 16 points, natural order in and out.
template <int N, bool INV>
__device__ __forceinline__ void dft_reg(float2 (&v)[N]) {
    constexpr int LG = N == 8 ? 3 : 4;
    float2 t[N];
#pragma unroll
    for (int i = 0; i < N; ++i) t[i] = v[brev_c<LG>(i)];
#pragma unroll
    for (int s = 0; s < LG; ++s) {
        const int h = 1 << s;
#pragma unroll
        for (int j = 0; j < N / 2; ++j) {
            const int pos = j & (h - 1);
            const int i0 = ((j >> s) << (s + 1)) | pos, i1 = i0 + h;
            const float2 b = mul_w16<INV>(t[i1], pos * (16 / (2 * h)));
            t[i1] = csub(t[i0], b);
            t[i0] = cadd(t[i0], b);
        }
    }
#pragma unroll
    for (int i = 0; i < N; ++i) v[i] = t[i];
}

// W_2048^m from the table of W_2048^k, k < 1024 (W^{k + 1024} = -W^k).
template <bool INV>
__device__ __forceinline__ float2 w2048(const float2* twl, uint32_t m) {
    const float2 t = twl[m & 1023u];
    const float sg = (m & 1024u) ? -1.f : 1.f;
    return make_float2(sg * t.x, INV ? -sg * t.y : sg * t.y);
}

__global__ __launch_bounds__(kZ2048Threads) void lct_zpass2048_kernel(float2* __restrict__ S,
                                                                    const float2* __restrict__ invt,
                                                                    const float2* __restrict__ tw,
                                                                    uint32_t N, uint32_t gx,
                                                                    uint32_t sp,
                                                                    const float2* __restrict__ rotw) {
    extern __shared__ float2 lds[];
    float2* twl = lds;          // 1024 twiddles
    float2* buf = lds + 1024;   // [2048][kLctCols] + padding
    const uint32_t tid = threadIdx.x;
    const uint32_t col = tid % kLctCols, u = tid / kLctCols;  // u in [0, kZU)
    const uint32_t H = N + 1, Y2 = 2 * N;
    // Blocks are dealt to the XCDs round-robin; each XCD takes a contiguous
    // range of (ky, column block) in order, so the two column blocks that
    // share a 128-byte line of a plane run one after the other on one XCD.
    const uint32_t nb = gridDim.x, nb8 = nb / 8u;
    const uint32_t bid = blockIdx.x < nb8 * 8u ? (blockIdx.x % 8u) * nb8 + blockIdx.x / 8u
                                               : blockIdx.x;
    const uint32_t bx = bid % gx, seq = bid / gx;
    // rotw != nullptr (2N = 256, mirror-symmetric PSF): rows ky and 2N - ky
    // share the filter tile of row min(ky, 2N - ky) (the second times
    // exp(2 pi i k / 2N), k = 2N - ky, = conj(rotw[k])); the sequence
    // 0, N, 1, 2N - 1, 2, 2N - 2, ... puts each pair next to each other.
    uint32_t ky = seq, kyt = seq;
    if (rotw) {
        ky = seq == 0 ? 0u : seq == 1 ? N : (seq % 2 == 0 ? seq / 2 : Y2 - (seq - 1) / 2);
        kyt = ky <= N ? ky : Y2 - ky;
    }
    const uint32_t kx0 = bx * kLctCols, kx = kx0 + col;
    const bool live = kx < H;
    auto at = [&](uint32_t pos) -> float2& { return buf[pos * kLctCols + col + 4u * (pos >> 3)]; };
    for (uint32_t k = tid; k < 1024; k += kZ2048Threads) twl[k] = tw[k];
    // The filter values this thread applies in pass 3 (the tiled copy: one
    // contiguous 8 KB block per (h, k2) for the workgroup), loaded now so
    // that their latency overlaps the forward passes.
    float2 fil[256 / kZU][8];
    const float2* ft = invt + (size_t(kyt) * gx + bx) * (256 / kZU) * 8 * kZ2048Threads;
#pragma unroll
    for (int h = 0; h < 256 / kZU; ++h)
#pragma unroll
        for (int k2 = 0; k2 < 8; ++k2) fil[h][k2] = ft[(h * 8 + k2) * kZ2048Threads + tid];
    if (ky != kyt) {
        const float2 w = rotw[kyt];
        const float2 rot = make_float2(w.x, -w.y);
#pragma unroll
        for (int h = 0; h < 256 / kZU; ++h)
#pragma unroll
            for (int k2 = 0; k2 < 8; ++k2) fil[h][k2] = cmul(fil[h][k2], rot);
    }
    __syncthreads();

    // Forward pass 1: a = u; x[a + 128 b] = 0 for b >= 8.
#pragma unroll
    for (int h = 0; h < 128 / kZU; ++h) {
        const uint32_t a = u + uint32_t(kZU) * h;
        float2 v[16];
#pragma unroll
        for (int b = 0; b < 16; ++b) {
            v[b] = make_float2(0.f, 0.f);
            if (b < 8 && live) v[b] = S[(size_t(a + 128u * b) * Y2 + ky) * sp + kx];
        }
        dft_reg<16, false>(v);
#pragma unroll
        for (int k0 = 0; k0 < 16; ++k0)
            at(k0 * 128u + a) = k0 == 0 ? v[0] : cmul(v[k0], w2048<false>(twl, a * k0));
    }
    __syncthreads();
    // Forward pass 2: q = k0 * 8 + a' = u.
    float2 r2[128 / kZU][16];
#pragma unroll
    for (int h = 0; h < 128 / kZU; ++h) {
        const uint32_t q = u + uint32_t(kZU) * h, k0 = q / 8u, ap = q % 8u;
#pragma unroll
        for (int c = 0; c < 16; ++c) r2[h][c] = at(k0 * 128u + ap + 8u * c);
    }
    __syncthreads();
#pragma unroll
    for (int h = 0; h < 128 / kZU; ++h) {
        const uint32_t q = u + uint32_t(kZU) * h, k0 = q / 8u, ap = q % 8u;
        dft_reg<16, false>(r2[h]);
#pragma unroll
        for (int k1 = 0; k1 < 16; ++k1)
            at(k0 * 128u + k1 * 8u + ap) =
                k1 == 0 ? r2[h][0] : cmul(r2[h][k1], w2048<false>(twl, 16u * ap * k1));
    }
    __syncthreads();
    // Forward pass 3 (radix 8 over a'), filter, inverse pass 1 (radix 8
    // over k2): r = k0 * 16 + k1 in {u, u + kZU}.
    float2 r3[256 / kZU][8];
#pragma unroll
    for (int h = 0; h < 256 / kZU; ++h) {
        const uint32_t r = u + uint32_t(kZU) * h, k0 = r / 16u, k1 = r % 16u;
#pragma unroll
        for (int ap = 0; ap < 8; ++ap) r3[h][ap] = at(k0 * 128u + k1 * 8u + ap);
    }
    __syncthreads();
#pragma unroll
    for (int h = 0; h < 256 / kZU; ++h) {
        const uint32_t r = u + uint32_t(kZU) * h, k0 = r / 16u, k1 = r % 16u, kl = k0 + 16u * k1;
        dft_reg<8, false>(r3[h]);
#pragma unroll
        for (int k2 = 0; k2 < 8; ++k2) r3[h][k2] = cmul(r3[h][k2], fil[h][k2]);
        dft_reg<8, true>(r3[h]);
#pragma unroll
        for (int z0 = 0; z0 < 8; ++z0)
            at(z0 * 256u + kl) = z0 == 0 ? r3[h][0] : cmul(r3[h][z0], w2048<true>(twl, kl * z0));
    }
    __syncthreads();
    // Inverse pass 2: q = z0 * 16 + l0 = u.
#pragma unroll
    for (int h = 0; h < 128 / kZU; ++h) {
        const uint32_t q = u + uint32_t(kZU) * h, z0 = q / 16u, l0 = q % 16u;
#pragma unroll
        for (int l1 = 0; l1 < 16; ++l1) r2[h][l1] = at(z0 * 256u + l0 + 16u * l1);
    }
    __syncthreads();
#pragma unroll
    for (int h = 0; h < 128 / kZU; ++h) {
        const uint32_t q = u + uint32_t(kZU) * h, z0 = q / 16u, l0 = q % 16u;
        dft_reg<16, true>(r2[h]);
#pragma unroll
        for (int z1 = 0; z1 < 16; ++z1)
            at(z0 * 256u + z1 * 16u + l0) =
                z1 == 0 ? r2[h][0] : cmul(r2[h][z1], w2048<true>(twl, 8u * l0 * z1));
    }
    __syncthreads();
    // Inverse pass 3 (radix 16 over l0): q = z0 * 16 + z1 = u;
    // outputs z = z0 + 8 z1 + 128 z2 for z2 < 8 (z < 1024).
#pragma unroll
    for (int h = 0; h < 128 / kZU; ++h) {
        const uint32_t q = u + uint32_t(kZU) * h, z0 = q / 16u, z1 = q % 16u;
#pragma unroll
        for (int l0 = 0; l0 < 16; ++l0) r2[h][l0] = at(z0 * 256u + z1 * 16u + l0);
        dft_reg<16, true>(r2[h]);
        if (live)
#pragma unroll
            for (int z2 = 0; z2 < 8; ++z2)
                S[(size_t(z0 + 8u * z1 + 128u * z2) * Y2 + ky) * sp + kx] = r2[h][z2];
    }
}

// ----------------------------------------------------------- LCT y-pass
// Length-256 DFT along ky of the half spectrum S[z][ky][kx] (kx < H, row
// pitch sp complex, planes of 256 rows), in place, for 2N = 256: one workgroup per
// (plane, 16 kx columns), 256 = 16 x 16 with x[n], n = 16 n1 + n2:
//   A[n2][k1] = sum_n1 x[16 n1 + n2] W16^(n1 k1)     (thread n2)
//   X[k1 + 16 k2] = sum_n2 A[n2][k1] W256^(n2 k1) W16^(n2 k2)   (thread k1)
// W = exp(-2 pi i / n); INV: the conjugate, unnormalised (as rocFFT's
// inverse).  The rows of 16 columns are 128-byte pieces of the planes.
constexpr int kYCols = 16;
// PRUNE: forward, rows ky >= 128 of the input are zero (not read);
// inverse, only output rows < 128 are written (the crop keeps y < N = 128).
template <bool INV, bool PRUNE>
__global__ __launch_bounds__(256) void lct_ypass256_kernel(float2* __restrict__ S,
                                                           const float2* __restrict__ tw256,
                                                           uint32_t H, uint32_t sp) {
    __shared__ float2 twl[256];
    __shared__ float2 buf[16][16 * kYCols + 4];  // [n2][k1 * 16 + column], padded rows
    const uint32_t t = threadIdx.x, col = t % kYCols, u = t / kYCols;
    const uint32_t kx = blockIdx.x * kYCols + col;
    const bool live = kx < H;
    float2* P = S + size_t(blockIdx.y) * 256 * sp;
    twl[t] = tw256[t];
    float2 v[16];
#pragma unroll
    for (int n1 = 0; n1 < 16; ++n1)
        v[n1] = live && !(PRUNE && !INV && n1 >= 8) ? P[size_t(16 * n1 + int(u)) * sp + kx]
                                                    : make_float2(0.f, 0.f);
    dft_reg<16, INV>(v);
    __syncthreads();  // twiddles
#pragma unroll
    for (int k1 = 0; k1 < 16; ++k1) {
        float2 w = twl[u * uint32_t(k1)];  // u k1 < 256
        if (INV) w.y = -w.y;
        buf[u][k1 * kYCols + col] = k1 == 0 ? v[0] : cmul(v[k1], w);
    }
    __syncthreads();
#pragma unroll
    for (int n2 = 0; n2 < 16; ++n2) v[n2] = buf[n2][u * kYCols + col];
    dft_reg<16, INV>(v);
    if (live)
#pragma unroll
        for (int k2 = 0; k2 < (PRUNE && INV ? 8 : 16); ++k2)
            P[size_t(u + 16u * uint32_t(k2)) * sp + kx] = v[k2];
}

// Forward transform along x of the resampled planes, fused with the
// resampling, for N = 128 (2N = 256): row y < 128 of plane z is
// x[n] = sum_e vals[e] in[cols[e]][y][n] for n < 128 (the resampling's sums
// in its order) and 0 for 128 <= n < 256; its 256-point DFT X[k], k <= 128,
// goes to spectrum row (z, y) (pitch sp).  Rows y >= 128 are zero and are
// not written (the pruned y-pass does not read them).  Sixteen rows per
// workgroup, sixteen threads per row: the y-pass's 16 x 16 scheme along x.
__global__ __launch_bounds__(256) void lct_xr2c256_kernel(const float* __restrict__ in,
                                                          float2* __restrict__ S,
                                                          const int* __restrict__ rowptr,
                                                          const int* __restrict__ cols,
                                                          const float* __restrict__ vals,
                                                          const float2* __restrict__ tw256,
                                                          uint32_t sp) {
    constexpr uint32_t N = 128;
    __shared__ float2 twl[256];
    __shared__ float2 buf[16][16 * 17];  // [row][n2 * 17 + k1]: conflict-free both ways
    const uint32_t t = threadIdx.x, r = t / 16u, u = t % 16u;
    const uint32_t y = blockIdx.x * 16u + r, z = blockIdx.y;
    twl[t] = tw256[t];
    const size_t NN = size_t(N) * N;
    const float* row = in + size_t(y) * N;
    float acc[8];
#pragma unroll
    for (int n1 = 0; n1 < 8; ++n1) acc[n1] = 0.f;
    const int e1 = rowptr[z + 1];
    for (int e = rowptr[z]; e < e1; ++e) {
        const float w = vals[e];
        const float* src = row + size_t(cols[e]) * NN + u;
#pragma unroll
        for (int n1 = 0; n1 < 8; ++n1) acc[n1] = fmaf(w, src[16 * n1], acc[n1]);
    }
    float2 v[16];
#pragma unroll
    for (int n1 = 0; n1 < 16; ++n1) v[n1] = make_float2(n1 < 8 ? acc[n1] : 0.f, 0.f);
    dft_reg<16, false>(v);
    __syncthreads();  // twiddles
#pragma unroll
    for (int k1 = 0; k1 < 16; ++k1)
        buf[r][u * 17u + uint32_t(k1)] = k1 == 0 ? v[0] : cmul(v[k1], twl[u * uint32_t(k1)]);
    __syncthreads();
#pragma unroll
    for (int n2 = 0; n2 < 16; ++n2) v[n2] = buf[r][uint32_t(n2) * 17u + u];
    dft_reg<16, false>(v);
    float2* dst = S + (size_t(z) * 256 + y) * sp;
#pragma unroll
    for (int k2 = 0; k2 <= 8; ++k2)
        if (u + 16u * uint32_t(k2) <= N) dst[u + 16u * uint32_t(k2)] = v[k2];
}

// Inverse transform along x for N = 128, cropped: spectrum row (z, y),
// y < 128, holds X[k], k <= 128, of a real row; its unnormalised 256-point
// inverse DFT (X[256 - k] = conj X[k]) at n < 128 goes to
// out[z][y][n] (planes of 128 x 128).
__global__ __launch_bounds__(256) void lct_xc2r256_kernel(const float2* __restrict__ S,
                                                          float* __restrict__ out,
                                                          const float2* __restrict__ tw256,
                                                          uint32_t sp) {
    constexpr uint32_t N = 128;
    __shared__ float2 twl[256];
    __shared__ float2 buf[16][16 * 17];  // [row][j2 * 17 + m1]
    const uint32_t t = threadIdx.x, r = t / 16u, u = t % 16u;
    const uint32_t y = blockIdx.x * 16u + r, z = blockIdx.y;
    twl[t] = tw256[t];
    const float2* src = S + (size_t(z) * 256 + y) * sp;
    float2 v[16];
#pragma unroll
    for (int j1 = 0; j1 < 16; ++j1) {
        const uint32_t k = 16u * uint32_t(j1) + u;
        if (k <= N) {
            v[j1] = src[k];
        } else {
            const float2 c = src[256u - k];
            v[j1] = make_float2(c.x, -c.y);
        }
    }
    dft_reg<16, true>(v);
    __syncthreads();  // twiddles
#pragma unroll
    for (int m1 = 0; m1 < 16; ++m1) {
        float2 w = twl[u * uint32_t(m1)];
        w.y = -w.y;
        buf[r][u * 17u + uint32_t(m1)] = m1 == 0 ? v[0] : cmul(v[m1], w);
    }
    __syncthreads();
#pragma unroll
    for (int j2 = 0; j2 < 16; ++j2) v[j2] = buf[r][uint32_t(j2) * 17u + u];
    dft_reg<16, true>(v);
    float* dst = out + (size_t(z) * N + y) * N;
#pragma unroll
    for (int m2 = 0; m2 < 8; ++m2) dst[u + 16u * uint32_t(m2)] = v[m2].x;
}

__global__ void lct_filter_kernel(float2* __restrict__ C, const float2* __restrict__ inv, size_t n) {
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float2 a = C[i], b = inv[i];
    C[i] = make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// crop [:M, :N, :N] of the padded real volume, times 1/(8 M N^2) (ifftn's
// normalisation), then mtxi along t and max(., 0).
__global__ void lct_final_kernel(const float* __restrict__ pad, float* __restrict__ out,
                                 const int* __restrict__ rowptr, const int* __restrict__ cols,
                                 const float* __restrict__ vals, uint32_t N, uint32_t M,
                                 uint32_t pitch, float norm) {
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    const size_t NN = size_t(N) * N;
    if (i >= NN * M) return;
    const uint32_t r = uint32_t(i / NN);
    const size_t p = i % NN;
    const size_t y = p / N, x = p % N;
    float acc = 0.f;
    for (int e = rowptr[r]; e < rowptr[r + 1]; ++e)
        acc = fmaf(vals[e], pad[(size_t(cols[e]) * pitch + y) * pitch + x], acc);
    out[i] = fmaxf(acc * norm, 0.f);
}

// lct_final_kernel for N a multiple of 4: four consecutive x per thread,
// output plane r = blockIdx.y, source planes of pitch x pitch (2N: the
// padded planes; N: planes already cropped); the same sums in the same
// order.
__global__ __launch_bounds__(256) void lct_final4_kernel(const float* __restrict__ pad,
                                                         float* __restrict__ out,
                                                         const int* __restrict__ rowptr,
                                                         const int* __restrict__ cols,
                                                         const float* __restrict__ vals, uint32_t N,
                                                         uint32_t pitch, float norm) {
    const uint32_t r = blockIdx.y;
    const uint32_t n4 = N / 4;
    const uint32_t q = blockIdx.x * 256u + threadIdx.x;
    if (q >= N * n4) return;
    const uint32_t y = q / n4, x = (q % n4) * 4u;
    float4 acc = make_float4(0.f, 0.f, 0.f, 0.f);
    const int e1 = rowptr[r + 1];
    for (int e = rowptr[r]; e < e1; ++e) {
        const float w = vals[e];
        const float4 v =
            *reinterpret_cast<const float4*>(pad + (size_t(cols[e]) * pitch + y) * pitch + x);
        acc.x = fmaf(w, v.x, acc.x);
        acc.y = fmaf(w, v.y, acc.y);
        acc.z = fmaf(w, v.z, acc.z);
        acc.w = fmaf(w, v.w, acc.w);
    }
    *reinterpret_cast<float4*>(out + size_t(r) * N * N + size_t(q) * 4) =
        make_float4(fmaxf(acc.x * norm, 0.f), fmaxf(acc.y * norm, 0.f), fmaxf(acc.z * norm, 0.f),
                    fmaxf(acc.w * norm, 0.f));
}

// psf (define_psf, python/lct.py:295-347) in FP64: for every (y, x) column the
// z with the smallest |(4 slope)^2 (x^2 + y^2) - z| (all ties) is 1; the
// result is divided by sqrt(count) (the two normalisations of the reference
// reduce to this) and rolled by (0, U, U).  Writes unnormalised 0/1 and
// counts ones.
__device__ __forceinline__ double linspace_at(double a, double b, uint32_t n, uint32_t i) {
    if (n == 1) return a;
    if (i == n - 1) return b;
    const double step = (b - a) / double(n - 1);
    return double(i) * step + a;
}

__global__ void psf_kernel(double* __restrict__ psf, uint32_t U, uint32_t V, double slope,
                           unsigned long long* __restrict__ count) {
    const uint32_t n2u = 2 * U, n2v = 2 * V;
    const size_t col = size_t(blockIdx.x) * blockDim.x + threadIdx.x;  // (y, x) column
    if (col >= size_t(n2u) * n2u) return;
    const uint32_t yi = uint32_t(col / n2u), xi = uint32_t(col % n2u);
    const double x = linspace_at(-1.0, 1.0, n2u, xi), y = linspace_at(-1.0, 1.0, n2u, yi);
    const double s4 = (4.0 * slope) * (4.0 * slope);
    const double c = s4 * (x * x + y * y);
    double best = INFINITY;
    for (uint32_t z = 0; z < n2v; ++z) best = fmin(best, fabs(c - linspace_at(0.0, 2.0, n2v, z)));
    // rolled destination of (y, x): ((y + U) mod 2U, (x + U) mod 2U)
    const uint32_t ry = (yi + U) % n2u, rx = (xi + U) % n2u;
    unsigned long long ones = 0;
    for (uint32_t z = 0; z < n2v; ++z) {
        const bool hit = fabs(c - linspace_at(0.0, 2.0, n2v, z)) == best;
        psf[(size_t(z) * n2u + ry) * n2u + rx] = hit ? 1.0 : 0.0;
        ones += hit;
    }
    atomicAdd(count, ones);
}

__global__ void scale_d_kernel(double* __restrict__ p, size_t n,
                               const unsigned long long* __restrict__ count) {
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) p[i] *= 1.0 / sqrt(double(*count));
}

// flag != 0 when the (rolled) PSF is not mirror-symmetric in y bit for bit:
// psf[z][2U-1-y][x] != psf[z][y][x] for some element.
__global__ void psf_sym_check_kernel(const double* __restrict__ psf, uint32_t U, size_t n,
                                     int* __restrict__ flag) {
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;  // over [z][y < U][x]
    if (i >= n) return;
    const uint32_t n2u = 2 * U;
    const size_t x = i % n2u, zy = i / n2u;
    const size_t y = zy % U, z = zy / U;
    const size_t a = (z * n2u + y) * n2u + x, b = (z * n2u + (n2u - 1 - y)) * n2u + x;
    if (psf[a] != psf[b]) atomicOr(flag, 1);
}

// invpsf = conj(F) / (|F|^2 + 1/snr), or conj(F) (backprojection), FP64 -> c64.
__global__ void invpsf_kernel(const double2* __restrict__ F, float2* __restrict__ inv, size_t n,
                              double inv_snr, int backprop) {
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double2 f = F[i];
    if (backprop) {
        inv[i] = make_float2(float(f.x), float(-f.y));
    } else {
        const double d = f.x * f.x + f.y * f.y + inv_snr;
        inv[i] = make_float2(float(f.x / d), float(-f.y / d));
    }
}

// The inverse filter (layout [kz][2N][N + 1]) rearranged so that the
// 2048-point z-pass reads it contiguously: entry
// (((ky * nbx + bx) * 2 + h) * 8 + k2) * 1024 + tid holds the value thread
// tid of workgroup (bx, ky) applies as fil[h][k2] (column kx = 8 bx + tid % 8,
// kz = kl + 256 k2, kl = r / 16 + 16 (r % 16), r = tid / 8 + 128 h); 0 past
// the last column.
__global__ void invt_tile_kernel(const float2* __restrict__ inv, float2* __restrict__ invt,
                                 uint32_t N, uint32_t nbx, size_t n) {
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t tid = uint32_t(i % kZ2048Threads);
    size_t rest = i / kZ2048Threads;
    const uint32_t k2 = uint32_t(rest % 8);
    rest /= 8;
    const uint32_t h = uint32_t(rest % 2);
    rest /= 2;
    const uint32_t bx = uint32_t(rest % nbx), ky = uint32_t(rest / nbx);
    const uint32_t H = N + 1, Y2 = 2 * N;
    const uint32_t kx = bx * kLctCols + tid % kLctCols;
    const uint32_t r = tid / kLctCols + uint32_t(kZU) * h, kl = r / 16u + 16u * (r % 16u);
    invt[i] = kx < H ? inv[(size_t(kl + 256u * k2) * Y2 + ky) * H + kx] : make_float2(0.f, 0.f);
}

// Closed form of python/lct.py:resampling_operator (:350-419):
// mtx[r][j] = (1/M) sum_{x in [rM+1, (r+1)M], ceil(sqrt x) = j+1} 1/sqrt(x);
// mtxi = mtx^T.  Built in FP64 on the host, stored as float CSR.
void build_resampling(uint32_t M, std::vector<int>& rp, std::vector<int>& cl,
                      std::vector<float>& vl, std::vector<int>& rpT, std::vector<int>& clT,
                      std::vector<float>& vlT) {
    std::vector<std::map<int, double>> rows(M);
    const uint64_t MM = uint64_t(M) * M;
    for (uint64_t x = 1; x <= MM; ++x) {
        const uint32_t r = uint32_t((x - 1) / M);
        int j = int(std::ceil(std::sqrt(double(x)))) - 1;
        if (j < 0) j = 0;
        if (j > int(M) - 1) j = int(M) - 1;
        rows[r][j] += 1.0 / std::sqrt(double(x)) / double(M);
    }
    std::vector<std::map<int, double>> rowsT(M);
    rp.assign(1, 0);
    for (uint32_t r = 0; r < M; ++r) {
        for (auto& kv : rows[r]) {
            cl.push_back(kv.first);
            vl.push_back(float(kv.second));
            rowsT[kv.first][int(r)] = kv.second;
        }
        rp.push_back(int(cl.size()));
    }
    rpT.assign(1, 0);
    for (uint32_t r = 0; r < M; ++r) {
        for (auto& kv : rowsT[r]) {
            clT.push_back(kv.first);
            vlT.push_back(float(kv.second));
        }
        rpT.push_back(int(clT.size()));
    }
}

template <typename T>
fbp_status upload(fbp_context ctx, const std::vector<T>& v, T** dst, hipStream_t s) {
    FBP_CHECK_HIP(ctx, hipMalloc((void**)dst, std::max<size_t>(1, v.size()) * sizeof(T)));
    FBP_CHECK_HIP(ctx, hipMemcpyAsync(*dst, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, s));
    FBP_CHECK_HIP(ctx, hipStreamSynchronize(s));
    return FBP_OK;
}

void free_csr(Csr& c) {
    // Teardown: nothing to report a failed free to.
    if (c.rowptr) (void)hipFree(c.rowptr);
    if (c.cols) (void)hipFree(c.cols);
    if (c.vals) (void)hipFree(c.vals);
    c = Csr{};
}

}  // namespace

void fft_cache_destroy(fbp_context ctx) {
    if (!ctx || !ctx->fft_cache) return;
    FftCache* c = static_cast<FftCache*>(ctx->fft_cache);
    for (auto& kv : c->plans) rocfft_plan_destroy(kv.second.plan);
    if (c->info) rocfft_execution_info_destroy(c->info);
    if (c->work) (void)hipFree(c->work);
    if (c->khat) (void)hipFree(c->khat);
    if (c->invpsf) (void)hipFree(c->invpsf);
    if (c->invt) (void)hipFree(c->invt);
    if (c->tw) (void)hipFree(c->tw);
    if (c->tw256) (void)hipFree(c->tw256);
    free_csr(c->mtx);
    free_csr(c->mtxi);
    delete c;
    ctx->fft_cache = nullptr;
}

}  // namespace fbp

using namespace fbp;

extern "C" {

uint32_t fbp_filter_h_padded_length(uint32_t nt, float delta_t, float wl_mean, float wl_sigma) {
    uint32_t N, pad;
    if (!(delta_t > 0.f) || !(wl_mean > 0.f) ||
        !filter_h_lengths(nt, delta_t, wl_mean, wl_sigma, N, pad))
        return 0;
    return N;
}

fbp_status fbp_filter_h(fbp_context ctx, float* data, uint64_t batch, uint32_t nt, float delta_t,
                        float wl_mean, float wl_sigma, int32_t border, void* stream) {
    if (!ctx || !data || batch == 0 || nt == 0 || !(delta_t > 0.f) || !(wl_mean > 0.f) ||
        (border != 0 && border != 1))
        return fail(ctx, FBP_ERR_INVALID_ARGUMENT, "filter_h: bad argument");
    uint32_t N, pad;
    if (!filter_h_lengths(nt, delta_t, wl_mean, wl_sigma, N, pad))
        return fail(ctx, FBP_ERR_INVALID_ARGUMENT,
                    "filter_h: 6 sigma = %g rounds to no bins of %g (no padding) or the "
                    "padded length exceeds 2^24", double(6 * wl_sigma), double(delta_t));
    FBP_CHECK(bind(ctx));
    hipStream_t s = as_stream(stream);
    // Tap radius kept when the transform is longer than N: 2 t6 (pad = t6 - 1).
    const uint32_t R = std::min<uint32_t>(2 * (pad + 1), N / 2 - 1);
    const uint32_t M = filter_h_transform_length(N, R);
    const uint32_t nh = M / 2 + 1;
    FftCache* c = cache_of(ctx);

    // Half spectrum of Re(K) (the real input makes Re(IFFT(FFT(H) FFT(K)))
    // equal to IFFT(FFT(H) FFT(Re K))).  K built as FB/Laser.cu:138-172 does.
    char keybuf[160];
    snprintf(keybuf, sizeof(keybuf), "%u/%u/%a/%a/%a", N, M, double(delta_t), double(wl_mean),
             double(wl_sigma));
    if (c->khat_key != keybuf) {
        const float t_max = delta_t * float(N - 1);
        std::vector<float> tv(N), env(N), kre(N), shifted(N);
        float sum = 0.0f;
        for (uint32_t i = 0; i < N; ++i) tv[i] = float(i) * delta_t;
        for (uint32_t i = 0; i < N; ++i) {
            const float val = (tv[i] - t_max * 0.5f) / wl_sigma;
            env[i] = std::exp(-(val * val) * 0.5f);
            sum += env[i];
        }
        const float TWO_PI = 6.28318530717958648f;
        for (uint32_t i = 0; i < N; ++i) {
            const float phase = TWO_PI * tv[i] / wl_mean;
            kre[i] = env[i] * std::cos(phase) / sum;
        }
        for (uint32_t i = 0; i < N; ++i) shifted[i] = kre[(i + N / 2) % N];
        if (M != N) {
            // Lags 0..R and -R..-1 of the N-periodic kernel on the M circle.
            std::vector<float> km(M, 0.0f);
            for (uint32_t j = 0; j <= R; ++j) km[j] = shifted[j];
            for (uint32_t j = 1; j <= R; ++j) km[M - j] = shifted[N - j];
            shifted.swap(km);
        }
        float* dk = nullptr;
        FBP_CHECK(upload(ctx, shifted, &dk, s));
        if (c->khat) FBP_CHECK_HIP(ctx, hipFree(c->khat));
        c->khat = nullptr;
        c->khat_key.clear();
        FBP_CHECK_HIP(ctx, hipMalloc((void**)&c->khat, size_t(nh) * sizeof(float2)));
        FftPlan* pk = nullptr;
        FBP_CHECK(get_plan(ctx, rocfft_transform_type_real_forward, rocfft_precision_single, {M}, 1, &pk));
        FBP_CHECK(run_plan(ctx, pk, dk, c->khat, s));
        FBP_CHECK_HIP(ctx, hipStreamSynchronize(s));
        FBP_CHECK_HIP(ctx, hipFree(dk));
        c->khat_key = keybuf;
    }

    // Process time series in chunks so the work buffers stay bounded.
    const uint64_t chunk = std::min<uint64_t>(batch, std::max<uint64_t>(1, (uint64_t(1) << 28) / M));
    const size_t real_n = size_t(chunk) * M, cplx_n = size_t(chunk) * nh;
    FBP_CHECK(ensure_scratch2(ctx, real_n * sizeof(float) + cplx_n * sizeof(float2) + 256));
    float* real = static_cast<float*>(ctx->scratch2);
    float2* spec = reinterpret_cast<float2*>(real + ((real_n + 63) & ~size_t(63)));
    // Every chunk, the ragged last one included, runs the chunk-sized plans:
    // rows nb..chunk-1 of the work buffer then hold stale rows whose
    // (independent) transforms nobody reads.
    FftPlan *pf = nullptr, *pi = nullptr;
    FBP_CHECK(get_plan(ctx, rocfft_transform_type_real_forward, rocfft_precision_single, {M}, chunk, &pf));
    FBP_CHECK(get_plan(ctx, rocfft_transform_type_real_inverse, rocfft_precision_single, {M}, chunk, &pi));
    for (uint64_t b0 = 0; b0 < batch; b0 += chunk) {
        const uint64_t nb = std::min<uint64_t>(chunk, batch - b0);
        float* d = data + size_t(b0) * nt;
        hipLaunchKernelGGL(fh_pad_kernel, grid1(size_t(nb) * M), dim3(256), 0, s, d, real, nb, nt, N,
                           M, pad, border);
        FBP_CHECK(run_plan(ctx, pf, real, spec, s));
        hipLaunchKernelGGL(fh_mul_kernel, grid1(size_t(nb) * nh), dim3(256), 0, s, spec, c->khat, nb, nh);
        FBP_CHECK(run_plan(ctx, pi, spec, real, s));
        hipLaunchKernelGGL(fh_unpad_kernel, grid1(size_t(nb) * nt), dim3(256), 0, s, real, d, nb, nt, M,
                           pad, M == N ? 1.0f : float(double(N) / double(M)));
        FBP_CHECK_HIP(ctx, hipGetLastError());
    }
    return FBP_OK;
}

fbp_status fbp_filter_log_fft(fbp_context ctx, float* volume, fbp_uvec3 size, float sigma,
                              void* stream) {
    if (!ctx || !volume || size.x == 0 || size.y == 0 || size.z == 0)
        return fail(ctx, FBP_ERR_INVALID_ARGUMENT, "log_fft: bad argument");
    FBP_CHECK(bind(ctx));
    hipStream_t s = as_stream(stream);
    const size_t nc = size_t(size.x / 2 + 1) * size.y * size.z;
    FBP_CHECK(ensure_scratch2(ctx, nc * sizeof(float2)));
    float2* spec = static_cast<float2*>(ctx->scratch2);
    FftPlan *pf = nullptr, *pi = nullptr;
    FBP_CHECK(get_plan(ctx, rocfft_transform_type_real_forward, rocfft_precision_single,
                       {size.x, size.y, size.z}, 1, &pf));
    FBP_CHECK(get_plan(ctx, rocfft_transform_type_real_inverse, rocfft_precision_single,
                       {size.x, size.y, size.z}, 1, &pi));
    FBP_CHECK(run_plan(ctx, pf, volume, spec, s));
    hipLaunchKernelGGL(logfft_mul_kernel, grid1(nc), dim3(256), 0, s, spec, size.x, size.y, size.z,
                       sigma);
    FBP_CHECK_HIP(ctx, hipGetLastError());
    return run_plan(ctx, pi, spec, volume, s);
}

fbp_status fbp_lct(fbp_context ctx, const float* data, const fbp_lct_params* p, float* out,
                   void* stream) {
    if (!ctx || !data || !p || !out) return fail(ctx, FBP_ERR_INVALID_ARGUMENT, "lct: null argument");
    const uint32_t N = p->n, K = p->downsample;
    if (K > 16 || N == 0 || (N & (N - 1)) || p->m == 0 || (p->m & (p->m - 1)) || (p->m >> K) < 2 ||
        !(p->range_max > 0.f) || !(p->snr > 0.f))
        return fail(ctx, FBP_ERR_INVALID_ARGUMENT,
                    "lct: N, m must be powers of two with m >> downsample >= 2, downsample <= 16, snr > 0");
    const uint32_t M = p->m >> K;
    // z_trim = round(z_trim / 2) per downsampling round (python/lct.py:104),
    // Python's round: halves go to the even neighbour.
    uint32_t z_trim = p->z_trim;
    for (uint32_t k = 0; k < K; ++k) {
        const uint32_t q = z_trim >> 1;
        z_trim = (z_trim & 1u) ? q + (q & 1u) : q;
    }
    FBP_CHECK(bind(ctx));
    hipStream_t s = as_stream(stream);
    FftCache* c = cache_of(ctx);
    const size_t NN = size_t(N) * N;
    const size_t pad_n = size_t(2 * M) * (2 * N) * (2 * N);
    const size_t half_n = size_t(2 * M) * (2 * N) * (N + 1);

    // Resampling operators (closed form, FP64 -> float CSR), cached per M.
    if (c->op_m != M) {
        free_csr(c->mtx);
        free_csr(c->mtxi);
        c->op_m = 0;
        std::vector<int> rp, cl, rpT, clT;
        std::vector<float> vl, vlT;
        build_resampling(M, rp, cl, vl, rpT, clT, vlT);
        FBP_CHECK(upload(ctx, rp, &c->mtx.rowptr, s));
        FBP_CHECK(upload(ctx, cl, &c->mtx.cols, s));
        FBP_CHECK(upload(ctx, vl, &c->mtx.vals, s));
        FBP_CHECK(upload(ctx, rpT, &c->mtxi.rowptr, s));
        FBP_CHECK(upload(ctx, clT, &c->mtxi.cols, s));
        FBP_CHECK(upload(ctx, vlT, &c->mtxi.vals, s));
        c->op_m = M;
    }

    // Inverse filter of the NLOS blur kernel, built in FP64 and cached.
    const double slope = double(p->width) / double(p->range_max);
    char keybuf[200];
    snprintf(keybuf, sizeof(keybuf), "%u/%u/%a/%a/%d", N, M, slope, double(p->snr), p->is_backprop);
    if (c->inv_key != keybuf) {
        if (c->invpsf) FBP_CHECK_HIP(ctx, hipFree(c->invpsf));
        c->invpsf = nullptr;
        if (c->invt) FBP_CHECK_HIP(ctx, hipFree(c->invt));
        c->invt = nullptr;
        c->inv_key.clear();
        double* psf = nullptr;
        double2* F = nullptr;
        unsigned long long* cnt = nullptr;
        FBP_CHECK_HIP(ctx, hipMalloc((void**)&psf, pad_n * sizeof(double)));
        FBP_CHECK_HIP(ctx, hipMalloc((void**)&F, half_n * sizeof(double2)));
        FBP_CHECK_HIP(ctx, hipMalloc((void**)&cnt, sizeof(unsigned long long)));
        int* symflag = nullptr;
        FBP_CHECK_HIP(ctx, hipMalloc((void**)&symflag, sizeof(int)));
        FBP_CHECK_HIP(ctx, hipMemsetAsync(cnt, 0, sizeof(unsigned long long), s));
        FBP_CHECK_HIP(ctx, hipMemsetAsync(symflag, 0, sizeof(int), s));
        hipLaunchKernelGGL(psf_kernel, grid1(size_t(2 * N) * 2 * N), dim3(256), 0, s, psf, N, M, slope, cnt);
        hipLaunchKernelGGL(scale_d_kernel, grid1(pad_n), dim3(256), 0, s, psf, pad_n, cnt);
        hipLaunchKernelGGL(psf_sym_check_kernel, grid1(pad_n / 2), dim3(256), 0, s, psf, N,
                           pad_n / 2, symflag);
        int symhost = 1;
        FftPlan* pd = nullptr;
        FBP_CHECK(get_plan(ctx, rocfft_transform_type_real_forward, rocfft_precision_double,
                           {2 * N, 2 * N, 2 * M}, 1, &pd));
        FBP_CHECK(run_plan(ctx, pd, psf, F, s));
        FBP_CHECK_HIP(ctx, hipMalloc((void**)&c->invpsf, half_n * sizeof(float2)));
        hipLaunchKernelGGL(invpsf_kernel, grid1(half_n), dim3(256), 0, s, F, c->invpsf, half_n,
                           1.0 / double(p->snr), p->is_backprop);
        FBP_CHECK_HIP(ctx, hipGetLastError());
        FBP_CHECK_HIP(ctx, hipMemcpyAsync(&symhost, symflag, sizeof(int), hipMemcpyDeviceToHost, s));
        FBP_CHECK_HIP(ctx, hipStreamSynchronize(s));
        FBP_CHECK_HIP(ctx, hipFree(psf));
        FBP_CHECK_HIP(ctx, hipFree(F));
        FBP_CHECK_HIP(ctx, hipFree(cnt));
        FBP_CHECK_HIP(ctx, hipFree(symflag));
        c->inv_sym = symhost == 0 ? 1 : 0;
        c->inv_key = keybuf;
    }

    const dim3 tgrid((M + 31) / 32, (N + 31) / 32, N);
    const double inv_m1 = M > 1 ? 1.0 / double(M - 1) : 0.0;
    const float norm = float(1.0 / double(pad_n));
    // Crop + inverse resampling + max(., 0) into out: 16-byte stores when
    // the rows allow them and out is 16-byte aligned.
    // pitch: row and plane edge of the source planes (2N padded, N cropped).
    auto launch_final = [&](const float* src, uint32_t pitch) {
        if (N % 4 == 0 && M <= 65535u && reinterpret_cast<uintptr_t>(out) % 16u == 0)
            hipLaunchKernelGGL(lct_final4_kernel, dim3((N * N / 4 + 255) / 256, M), dim3(256), 0, s,
                               src, out, c->mtxi.rowptr, c->mtxi.cols, c->mtxi.vals, N, pitch,
                               norm);
        else
            hipLaunchKernelGGL(lct_final_kernel, grid1(NN * M), dim3(256), 0, s, src, out,
                               c->mtxi.rowptr, c->mtxi.cols, c->mtxi.vals, N, M, pitch, norm);
    };
    // Pruned path: the padded volume's upper M planes are zero and only its
    // first M output planes are kept, so the 2D transforms run on M planes
    // and one kernel does the z transforms and the filter in LDS.
    const uint32_t lg = 1u + uint32_t(std::lround(std::log2(double(M))));
    const size_t zpass_lds = (size_t(2 * M) * kLctCols + M) * sizeof(float2);
    if (zpass_lds <= size_t(ctx->max_lds_per_block)) {
        if (c->tw_m != M) {
            std::vector<float2> tw(M);
            for (uint32_t k = 0; k < M; ++k) {
                const double a = -2.0 * M_PI * double(k) / double(2 * M);
                tw[k] = make_float2(float(std::cos(a)), float(std::sin(a)));
            }
            if (c->tw) FBP_CHECK_HIP(ctx, hipFree(c->tw));
            c->tw = nullptr;
            c->tw_m = 0;
            FBP_CHECK(upload(ctx, tw, &c->tw, s));
            c->tw_m = M;
        }
        // Spectrum rows of N + 1 complex; for 2N = 256 (the y-pass form)
        // padded to a multiple of 16 (128-byte aligned rows).
        const bool ysplit = 2 * N == 256;
        const uint32_t sp = ysplit ? ((N + 1 + 15) / 16) * 16 : N + 1;
        const size_t plane = size_t(2 * N) * (2 * N), hplane = size_t(2 * N) * sp;
        const size_t d_bytes = (NN * M * sizeof(float) + 255) & ~size_t(255);
        const size_t r_bytes = (plane * M * sizeof(float) + 255) & ~size_t(255);
        FBP_CHECK(ensure_scratch2(ctx, d_bytes + r_bytes + hplane * M * sizeof(float2)));
        char* base = static_cast<char*>(ctx->scratch2);
        float* d = reinterpret_cast<float*>(base);
        float* planes = reinterpret_cast<float*>(base + d_bytes);
        float2* spec = reinterpret_cast<float2*>(base + d_bytes + r_bytes);
        hipLaunchKernelGGL(lct_permute_kernel, tgrid, dim3(32, 8), 0, s, data, d, N, M, K, z_trim,
                           p->is_diffuse, inv_m1);
        // 2D transforms of the M planes.  For 2N = 256: the resampling fused
        // with a 256-point x transform of the 128 non-zero rows of each plane
        // (the padded planes are never written), and the y-pass kernel along
        // ky on 128-byte row pieces, pruned to the non-zero input rows.
        // Otherwise the resampled padded planes and rocFFT's 2D plans.
        const dim3 ygrid((N + 1 + kYCols - 1) / kYCols, M);
        FftPlan *pf = nullptr, *pi = nullptr;
        if (ysplit) {
            if (!c->tw256) {
                std::vector<float2> w(256);
                for (int k = 0; k < 256; ++k) {
                    const double a = -2.0 * 3.14159265358979323846 * double(k) / 256.0;
                    w[k] = make_float2(float(std::cos(a)), float(std::sin(a)));
                }
                FBP_CHECK(upload(ctx, w, &c->tw256, s));
            }
            hipLaunchKernelGGL(lct_xr2c256_kernel, dim3(N / 16, M), dim3(256), 0, s, d, spec,
                               c->mtx.rowptr, c->mtx.cols, c->mtx.vals, c->tw256, sp);
            FBP_CHECK_HIP(ctx, hipGetLastError());
            hipLaunchKernelGGL((lct_ypass256_kernel<false, true>), ygrid, dim3(256), 0, s, spec,
                               c->tw256, N + 1, sp);
            FBP_CHECK_HIP(ctx, hipGetLastError());
        } else {
            if (N % 4 == 0)
                hipLaunchKernelGGL(lct_resample_planes4_kernel, dim3((N * N + 255) / 256, M),
                                   dim3(256), 0, s, d, planes, c->mtx.rowptr, c->mtx.cols,
                                   c->mtx.vals, N);
            else
                hipLaunchKernelGGL(lct_resample_planes_kernel, grid1(plane * M), dim3(256), 0, s,
                                   d, planes, c->mtx.rowptr, c->mtx.cols, c->mtx.vals, N, M);
            FBP_CHECK_HIP(ctx, hipGetLastError());
            FBP_CHECK(get_plan(ctx, rocfft_transform_type_real_forward, rocfft_precision_single,
                               {2 * N, 2 * N}, M, &pf));
            FBP_CHECK(get_plan(ctx, rocfft_transform_type_real_inverse, rocfft_precision_single,
                               {2 * N, 2 * N}, M, &pi));
            FBP_CHECK(run_plan(ctx, pf, planes, spec, s));
        }
        FBP_CHECK_HIP(ctx, hipFuncSetAttribute(reinterpret_cast<const void*>(lct_zpass_kernel),
                                               hipFuncAttributeMaxDynamicSharedMemorySize,
                                               int(zpass_lds)));
        const dim3 zgrid((N + 1 + kLctCols - 1) / kLctCols, 2 * N);
        const size_t lds2048 = (size_t(2048) * kLctCols + 4 * 256 + 1024) * sizeof(float2);
        if (M == 1024 && lds2048 <= size_t(ctx->max_lds_per_block)) {
            FBP_CHECK_HIP(ctx, hipFuncSetAttribute(
                                   reinterpret_cast<const void*>(lct_zpass2048_kernel),
                                   hipFuncAttributeMaxDynamicSharedMemorySize, int(lds2048)));
            // Mirror-symmetric PSF at 2N = 256: tiles of rows ky <= N only.
            const bool zsym = ysplit && c->inv_sym;
            if (!c->invt) {
                const size_t nt =
                    size_t(zgrid.x) * (zsym ? N + 1 : zgrid.y) * (256 / kZU) * 8 * kZ2048Threads;
                FBP_CHECK_HIP(ctx, hipMalloc((void**)&c->invt, nt * sizeof(float2)));
                hipLaunchKernelGGL(invt_tile_kernel, grid1(nt), dim3(256), 0, s, c->invpsf,
                                   c->invt, N, zgrid.x, nt);
                FBP_CHECK_HIP(ctx, hipGetLastError());
            }
            hipLaunchKernelGGL(lct_zpass2048_kernel, dim3(zgrid.x * zgrid.y), dim3(kZ2048Threads),
                               lds2048, s, spec, c->invt, c->tw, N, zgrid.x, sp,
                               zsym ? c->tw256 : nullptr);
        } else {
            hipLaunchKernelGGL(lct_zpass_kernel, zgrid, dim3(kLctThreads), zpass_lds, s, spec,
                               c->invpsf, c->tw, N, M, lg, sp);
        }
        FBP_CHECK_HIP(ctx, hipGetLastError());
        if (ysplit) {
            // Inverse y-pass keeping the rows y < N, the inverse x transform
            // keeping x < N: cropped 128 x 128 planes.
            hipLaunchKernelGGL((lct_ypass256_kernel<true, true>), ygrid, dim3(256), 0, s, spec,
                               c->tw256, N + 1, sp);
            FBP_CHECK_HIP(ctx, hipGetLastError());
            hipLaunchKernelGGL(lct_xc2r256_kernel, dim3(N / 16, M), dim3(256), 0, s, spec, planes,
                               c->tw256, sp);
            FBP_CHECK_HIP(ctx, hipGetLastError());
            launch_final(planes, N);
        } else {
            FBP_CHECK(run_plan(ctx, pi, spec, planes, s));
            launch_final(planes, 2 * N);
        }
        FBP_CHECK_HIP(ctx, hipGetLastError());
        return FBP_OK;
    }

    // Work buffers: d [M][N][N] | padded real [2M][2N][2N] | half spectrum.
    const size_t d_bytes = (NN * M * sizeof(float) + 255) & ~size_t(255);
    const size_t r_bytes = (pad_n * sizeof(float) + 255) & ~size_t(255);
    FBP_CHECK(ensure_scratch2(ctx, d_bytes + r_bytes + half_n * sizeof(float2)));
    char* base = static_cast<char*>(ctx->scratch2);
    float* d = reinterpret_cast<float*>(base);
    float* padv = reinterpret_cast<float*>(base + d_bytes);
    float2* spec = reinterpret_cast<float2*>(base + d_bytes + r_bytes);

    hipLaunchKernelGGL(lct_permute_kernel, tgrid, dim3(32, 8), 0, s, data, d, N, M, K, z_trim,
                       p->is_diffuse, inv_m1);
    hipLaunchKernelGGL(fill_kernel, grid1(pad_n), dim3(256), 0, s, padv, pad_n);
    hipLaunchKernelGGL(lct_resample_kernel, grid1(NN * M), dim3(256), 0, s, d, padv, c->mtx.rowptr,
                       c->mtx.cols, c->mtx.vals, N, M, 1);
    FBP_CHECK_HIP(ctx, hipGetLastError());
    FftPlan *pf = nullptr, *pi = nullptr;
    FBP_CHECK(get_plan(ctx, rocfft_transform_type_real_forward, rocfft_precision_single,
                       {2 * N, 2 * N, 2 * M}, 1, &pf));
    FBP_CHECK(get_plan(ctx, rocfft_transform_type_real_inverse, rocfft_precision_single,
                       {2 * N, 2 * N, 2 * M}, 1, &pi));
    FBP_CHECK(run_plan(ctx, pf, padv, spec, s));
    hipLaunchKernelGGL(lct_filter_kernel, grid1(half_n), dim3(256), 0, s, spec, c->invpsf, half_n);
    FBP_CHECK(run_plan(ctx, pi, spec, padv, s));
    launch_final(padv, 2 * N);
    FBP_CHECK_HIP(ctx, hipGetLastError());
    return FBP_OK;
}

}  // extern "C"
