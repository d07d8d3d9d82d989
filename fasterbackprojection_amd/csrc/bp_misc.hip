// Depth-plane backprojection, MIS-sampled confocal backprojection, transient
// quantisation and the confocal forward model.
#include <algorithm>
#include <cmath>
#include <vector>

#include "bp_common.h"
#include "bp_confocal_aligned.h"

namespace fbp {

// ------------------------------------------------------------ depth planes
// Reference: backprojectConfocal (FB/transient_reconstruction.cuh:41-64).
// One thread per output (c, d) gathers over all scan points in a fixed
// order (no atomics).  No 1/r^2 weight, as in the reference.  The exhaustive
// form is in bp_depth_exhaustive.hip.  Confocal captures on an aligned 64^2
// or 128^2 scan grid with the wall normal along y take the ring-convolution
// path instead (bp_confocal_aligned.hip, depth-plane mode): there the bin of
// a (pixel, scan point) pair depends only on their grid offset and the
// depth.

__global__ void depth_kernel(const float* __restrict__ I, const fbp_vec3* __restrict__ laser,
                             const double* __restrict__ depths, uint32_t D, uint32_t L, uint32_t T,
                             float t0, float dt, uint32_t discard, fbp_vec3 n, fbp_vec3 lpos,
                             fbp_vec3 spos, float* __restrict__ act, const int* __restrict__ skip) {
    // The aligned-grid path wrote the activation (its check passed).
    if (skip && *skip != 0) return;
    const uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= L * D) return;
    const uint32_t c = idx / D, d = idx % D;
    const float depth = static_cast<float>(depths[d]);
    // checkPos = laserTargets[c] + relayWallNormal * depth (componentwise).
    const float cx = rn_add(laser[c].x, rn_mul(n.x, depth));
    const float cy = rn_add(laser[c].y, rn_mul(n.y, depth));
    const float cz = rn_add(laser[c].z, rn_mul(n.z, depth));
    float sum = 0.f;
    for (uint32_t l = 0; l < L; ++l) {
        const float lx = laser[l].x, ly = laser[l].y, lz = laser[l].z;
        float dist = rn_sub(rn_mul(distance3(lx, ly, lz, cx, cy, cz), 2.0f), t0);
        if (discard == 0)
            dist = rn_add(dist, rn_add(distance3(lx, ly, lz, lpos.x, lpos.y, lpos.z),
                                       distance3(lx, ly, lz, spos.x, spos.y, spos.z)));
        const float q = rn_div(dist, dt);
        if (!(q > -1.0f) || !(q < float(T))) continue;
        const int bin = int(q);
        const float v = I[size_t(l) * T + bin];
        if (v > kEps) sum += v;
    }
    act[idx] = sum;
}

// --------------------------------------------------------------------- MIS
// Reference: backprojectConfocalVoxelMIS (FB/transient_reconstruction.cuh:194-260),
// getUniformRandom / getMISIndexAlias (FB/transient_utils.cuh:18-53).
//
// Sample s of voxel v (linear index z * X * Y + vy * X + vx) draws from the
// noise state v + s, so its scan point and weight depend on that sum alone:
// every state is shared by num_samples consecutive voxels.  mis_state_kernel
// draws each state once (scan point, weight factor, path offset);
// mis_walk_kernel gives each wave 64 consecutive voxels and walks the states
// v0 .. v0 + 63 + N - 1 in order, so at every step the wave reads one
// transient row (nearby bins for nearby voxels) instead of 64 random rows.
// Each voxel still adds its samples s = 0..N-1 in order (the reference runs
// one thread per (voxel, sample) with a racy +=).

__device__ __forceinline__ float mis_uniform(const float* __restrict__ noise, uint32_t n,
                                             uint32_t& state) {
    state = state * 747796405u + 2891336453u;
    return noise[state % n];
}

// One noise state's draw.  coef: REFERENCE mode pdf * safeRCP(globalPDF);
// UNBIASED mode the estimator's denominator (p_alias + 1/L) / 2 * N.  off:
// |Lpos - l| + |Spos - l| (0 when the first/last bounces are discarded).
struct MisState {
    uint32_t l;
    float coef;
    float off;
    float unused;
};

__global__ void mis_state_kernel(const float* __restrict__ cdf, const uint32_t* __restrict__ alias,
                                 const float* __restrict__ prob, const float* __restrict__ noise,
                                 uint32_t noise_size, uint32_t L, uint32_t num_samples,
                                 uint32_t mode, uint32_t discard,
                                 const fbp_vec3* __restrict__ laser, fbp_vec3 lpos, fbp_vec3 spos,
                                 size_t count, MisState* __restrict__ out) {
    const size_t sigma = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (sigma >= count) return;
    const float invL = safe_rcp(static_cast<float>(L));
    // Draws r0 (strategy), r1 (uniform index k), r2 (alias acceptance, used
    // by the alias strategy only); l = alias-table pick or k.
    uint32_t state = uint32_t(sigma);
    const float r0 = mis_uniform(noise, noise_size, state);
    const float r1 = mis_uniform(noise, noise_size, state);
    const float r2 = mis_uniform(noise, noise_size, state);
    uint32_t k = static_cast<uint32_t>(r1 * static_cast<float>(L));
    if (k >= L) k = L - 1;
    const bool by_alias = r0 < 0.5f;
    const uint32_t l = by_alias ? ((r2 < prob[k]) ? k : alias[k]) : k;
    const float cl = cdf[l], cm = cdf[l > 0 ? l - 1 : 0];
    const float pl = l > 0 ? cl - cm : cl;
    const float pdf = by_alias ? pl : 1.0f * invL;
    const float gpdf = by_alias ? pdf + 1.0f * invL : pdf + pl;
    MisState st;
    st.l = l;
    if (mode == FBP_MIS_REFERENCE) {
        st.coef = rn_mul(pdf, safe_rcp(gpdf));
    } else {
        // One-sample estimator over the mixture p = (p_alias + 1/L) / 2
        // (unbiased: E[f(l) / p(l)] / N = sum_l f(l) / N).
        const float pmix = 0.5f * (pl + invL);
        st.coef = pmix * float(num_samples);
    }
    st.off = 0.f;
    if (discard == 0) {
        const float lx = laser[l].x, ly = laser[l].y, lz = laser[l].z;
        st.off = rn_add(distance3(lx, ly, lz, lpos.x, lpos.y, lpos.z),
                        distance3(lx, ly, lz, spos.x, spos.y, spos.z));
    }
    st.unused = 0.f;
    out[sigma] = st;
}

// PURE (t0 = 0, bounces discarded, no voxel scattering): the bin comes from
// the exact squared distance s and the threshold row S_k = min{s : bin(s) >=
// k} staged in LDS (an estimate biased low, one compare: the reference's bin
// without its IEEE sqrt and division), the weight from a hardware
// reciprocal of 4 s (FP32 tolerance), the reference's IEEE weight for the
// rare pairs within 1.7 mm of their scan point (safeRCP's clamp).
// Otherwise every sample takes the reference's IEEE chain.
template <bool REF, bool SCATTER, bool PURE>
__global__ __launch_bounds__(256) void mis_walk_kernel(const float* __restrict__ I,
                                                       const fbp_vec3* __restrict__ laser,
                                                       const MisState* __restrict__ states,
                                                       uint32_t num_samples, VolumeGeom g,
                                                       uint32_t T, float t0, float dt,
                                                       uint32_t discard, float ext, uint32_t rows,
                                                       const float* __restrict__ thr,
                                                       float* __restrict__ volume) {
    extern __shared__ float srow[];  // PURE: S_0 .. S_T
    if constexpr (PURE) {
        for (uint32_t k = threadIdx.x; k <= T; k += blockDim.x) srow[k] = thr[k];
        __syncthreads();
    }
    // A wave owns a patch of `rows` voxel rows x (64 / rows) voxels along x
    // of one depth slice; its voxels' states lie in [v0, send).
    const uint32_t X = g.res[0], Y = g.res[1];
    const uint32_t wx = 64u / rows;
    const uint32_t pbx = (X + wx - 1) / wx, pby = (Y + rows - 1) / rows;
    const size_t wave = size_t(blockIdx.x) * (blockDim.x >> 6) + uint32_t(rfl(int(threadIdx.x >> 6)));
    if (wave >= size_t(pbx) * pby * g.res[2]) return;
    const uint32_t z = uint32_t(wave / (size_t(pbx) * pby));
    const uint32_t pr = uint32_t(wave % (size_t(pbx) * pby));
    const uint32_t x0 = (pr % pbx) * wx, y0 = (pr / pbx) * rows;
    const uint32_t j = threadIdx.x & 63;
    const uint32_t vx = min(x0 + j % wx, X - 1), vy = min(y0 + j / wx, Y - 1);
    const bool valid = x0 + j % wx < X && y0 + j / wx < Y;
    const size_t v0 = (size_t(z) * Y + y0) * X + x0;
    const size_t vid = (size_t(z) * Y + vy) * X + vx;
    const float px = voxel_center(g.hmin[0], g.vs[0], vx);
    const float py = voxel_center(g.hmin[1], g.vs[1], z);
    const float pz = voxel_center(g.hmin[2], g.vs[2], vy);
    auto path = [&](const MisState& st) {
        const fbp_vec3 lp = laser[st.l];
        float dist = rn_sub(rn_mul(distance3(lp.x, lp.y, lp.z, px, py, pz), 2.0f), t0);
        if (discard == 0) dist = rn_add(dist, st.off);
        return dist;
    };
    auto add = [&](float& sum, float light, float wd, float coef) {
        if constexpr (REF)
            sum += rn_mul(rn_mul(coef, light), wd);
        else
            sum += (light * wd) / coef;
    };
    float sum = 0.f;
    const size_t send = v0 + size_t(rows - 1) * X + wx - 1 + num_samples;
    if constexpr (SCATTER) {
        // Voxel-scattering walk (FB/transient_reconstruction.cuh:235-250).
        for (size_t sg = v0; sg < send; ++sg) {
            const bool act = valid && sg >= vid && sg < vid + num_samples;
            if (!act) continue;
            const MisState st = states[sg];
            const float dist = path(st);
            const float wd = safe_rcp(rn_mul(dist, dist));
            const float* row = I + size_t(st.l) * T;
            scatter_walk(dist, ext, dt, T, [&](int bin) {
                const float light = row[uint32_t(bin)];
                if (light > kEps) add(sum, light, wd, st.coef);
            });
        }
    } else {
        // Eight states per pass: their loads are independent, so the chains
        // overlap; every load is issued (indices clamped in range) and the
        // contributions are added in state order.
        constexpr int U = 8;
        // Pure-chain estimate (s * rsq(s)) * cb ~ 2 r / dt, biased low by
        // 2^-18 relative so the bin is ka or ka + 1.
        const float cb = (2.0f / dt) * (1.0f - 3.814697265625e-06f);
        for (size_t s0 = v0; s0 < send; s0 += U) {
            float light[U], wd[U], coef[U], s2[U];
            bool ok[U];
            // State sg serves this voxel iff 0 <= sg - vid < N (32-bit wrap).
            const uint32_t rel0 = uint32_t(s0 - vid);
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const size_t sg = s0 + u;
                const MisState st = states[sg < send ? sg : send - 1];
                const bool act = valid && rel0 + uint32_t(u) < num_samples;
                int bin;
                if constexpr (PURE) {
                    const fbp_vec3 lp = laser[st.l];
                    s2[u] = dot3(sq(rn_sub(px, lp.x)), sq(rn_sub(py, lp.y)), sq(rn_sub(pz, lp.z)));
                    const float est = fmaxf((s2[u] * __builtin_amdgcn_rsqf(s2[u])) * cb, 0.f);
                    const int ka = min(int(est), int(T));
                    const int k = ka + int(srow[min(ka + 1, int(T))] <= s2[u] && ka < int(T));
                    ok[u] = act && k < int(T);
                    bin = ok[u] ? k : 0;
                    wd[u] = __builtin_amdgcn_rcpf(s2[u] * 4.0f);
                } else {
                    const float dist = path(st);
                    wd[u] = safe_rcp(rn_mul(dist, dist));
                    const float q = rn_div(dist, dt);
                    ok[u] = act && (q > -1.0f) && (q < float(T));
                    bin = ok[u] ? int(q) : 0;
                }
                light[u] = I[size_t(ok[u] ? st.l : 0) * T + uint32_t(bin)];
                coef[u] = st.coef;
            }
#pragma unroll
            for (int u = 0; u < U; ++u) {
                if (ok[u] && light[u] > kEps) {
                    float w = wd[u];
                    if constexpr (PURE) {
                        if (!(s2[u] > 3.0e-6f)) {  // dist^2 = 4 s near safeRCP's 1e-5
                            const float d = rn_mul(rn_sqrt(s2[u]), 2.0f);
                            w = safe_rcp(rn_mul(d, d));
                        }
                    }
                    add(sum, light[u], w, coef[u]);
                }
            }
        }
    }
    if (valid) volume[vid] = sum;
}

// Inclusive FP64 prefix of the row sums (one workgroup of 1024 threads).
__global__ __launch_bounds__(1024) void row_prefix_kernel(const double* __restrict__ sum, uint32_t L,
                                                          double* __restrict__ cum) {
    __shared__ double part[1024];
    const uint32_t t = threadIdx.x;
    const uint32_t per = (L + 1023) / 1024;
    const uint32_t b = t * per, e = min(L, b + per);
    double local = 0.0;
    for (uint32_t l = b; l < e; ++l) local += sum[l];
    part[t] = local;
    __syncthreads();
    for (uint32_t off = 1; off < 1024; off <<= 1) {
        const double v = t >= off ? part[t - off] : 0.0;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    double run = part[t] - local;
    for (uint32_t l = b; l < e; ++l) {
        run += sum[l];
        cum[l] = run;
    }
}

// Entry p of the virtual noise array of a seeded draw: a uniform in [0, 1)
// with 24 random bits from a counter hash of (seed, p) (the "lowbias32"
// integer mix, three rounds).
__device__ __forceinline__ uint32_t mix32(uint32_t x) {
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}
__device__ __forceinline__ float seeded_noise(uint64_t seed, uint32_t p) {
    const uint32_t key = mix32(uint32_t(seed) ^ mix32(uint32_t(seed >> 32) + 0x9e3779b9u));
    return float(mix32(p ^ key) >> 8) * (1.0f / 16777216.0f);
}

// mis_uniform over noise, or over the seeded virtual array when noise is null.
__device__ __forceinline__ float mis_uniform_or_seeded(const float* __restrict__ noise, uint32_t n,
                                                       uint64_t seed, uint32_t& state) {
    state = state * 747796405u + 2891336453u;
    return noise ? noise[state % n] : seeded_noise(seed, state % n);
}

// Scan-point samples of every voxel tile for FBP_MIS_TILED.
// cum == nullptr: every block builds the inclusive FP64 prefix of sum in
// LDS itself (thread t adds its contiguous range of rows in order, then a
// scan over the 256 range totals), so no separate prefix launch is needed
// (L <= kMisLdsRows).
constexpr uint32_t kMisLdsRows = 8192;

__global__ __launch_bounds__(256) void mis_tile_samples_kernel(
    const double* __restrict__ sum, const double* __restrict__ cum_in, uint32_t L,
    const float* __restrict__ noise, uint32_t noise_size, uint64_t seed, uint32_t samples,
    uint32_t tiles, uint32_t* __restrict__ sample_l, float* __restrict__ sample_w) {
    extern __shared__ double scum[];  // L (prefix built here)
    __shared__ double part[256];
    const double* cum = cum_in;
    if (!cum_in) {
        const uint32_t t = threadIdx.x;
        const uint32_t per = (L + 255) / 256;
        const uint32_t b = min(L, t * per), e = min(L, b + per);
        double local = 0.0;
        for (uint32_t r = b; r < e; ++r) local += sum[r];
        part[t] = local;
        __syncthreads();
        for (uint32_t off = 1; off < 256; off <<= 1) {
            const double v = t >= off ? part[t - off] : 0.0;
            __syncthreads();
            part[t] += v;
            __syncthreads();
        }
        double run = part[t] - local;
        for (uint32_t r = b; r < e; ++r) {
            run += sum[r];
            scum[r] = run;
        }
        __syncthreads();
        cum = scum;
    }
    const size_t id = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (id >= size_t(tiles) * samples) return;
    const uint32_t tile = uint32_t(id / samples), i = uint32_t(id % samples);
    const double total = cum[L - 1];
    uint32_t state = tile * samples + i;
    const float u0 = mis_uniform_or_seeded(noise, noise_size, seed, state);
    const float u1 = mis_uniform_or_seeded(noise, noise_size, seed, state);
    uint32_t l;
    if (total > 0.0 && u0 < 0.5f) {
        const double target = double(u1) * total;
        uint32_t lo = 0, hi = L - 1;  // first l with cum[l] > target
        while (lo < hi) {
            const uint32_t mid = (lo + hi) >> 1;
            if (cum[mid] > target)
                hi = mid;
            else
                lo = mid + 1;
        }
        l = lo;
    } else {
        l = uint32_t(u1 * float(L));
        if (l >= L) l = L - 1;
    }
    const double pmix = total > 0.0 ? 0.5 * (sum[l] / total + 1.0 / double(L)) : 1.0 / double(L);
    sample_l[id] = l;
    sample_w[id] = float(1.0 / (pmix * double(samples)));
}

// Per-scan-point total intensity (I > 1e-5 samples), in FP64.
__global__ void row_sum_kernel(const float* __restrict__ I, uint32_t L, uint32_t T,
                               double* __restrict__ out) {
    const uint32_t l = blockIdx.x;
    double acc = 0.0;
    for (uint32_t t = threadIdx.x; t < T; t += blockDim.x) {
        float v = I[size_t(l) * T + t];
        if (v > kEps) acc += v;
    }
    __shared__ double red[256];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if (int(threadIdx.x) < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[l] = red[0];
}

// ------------------------------------------------------------ quantisation
// scale = max_t I (after zeroing I <= 1e-5); q = I / scale in F16 / BF16.
template <int STORAGE>
__global__ void quantize_kernel(const float* __restrict__ src, uint32_t cols, void* __restrict__ dst,
                                float* __restrict__ scales) {
    const size_t row = blockIdx.x;
    const float* r = src + row * cols;
    float mx = 0.f;
    for (uint32_t t = threadIdx.x; t < cols; t += blockDim.x) {
        float v = r[t];
        v = (v > kEps) ? v : 0.f;
        mx = fmaxf(mx, v);
    }
    __shared__ float red[256];
    red[threadIdx.x] = mx;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if (int(threadIdx.x) < w) red[threadIdx.x] = fmaxf(red[threadIdx.x], red[threadIdx.x + w]);
        __syncthreads();
    }
    const float scale = red[0] > 0.f ? red[0] : 1.0f;
    const float inv = 1.0f / scale;
    if (threadIdx.x == 0) scales[row] = scale;
    uint16_t* out = static_cast<uint16_t*>(dst) + row * cols;
    for (uint32_t t = threadIdx.x; t < cols; t += blockDim.x) {
        float v = r[t];
        v = (v > kEps) ? v * inv : 0.f;
        if constexpr (STORAGE == FBP_STORAGE_F16) {
            out[t] = __half_as_ushort(__float2half_rn(v));
        } else {
            // round-to-nearest-even to bfloat16
            uint32_t b = __float_as_uint(v);
            b += 0x7fffu + ((b >> 16) & 1u);
            out[t] = uint16_t(b >> 16);
        }
    }
}

// ----------------------------------------------------------- forward model
// One workgroup per scan point (transient row): for every point p,
// I[l][round(2 r / delta)] += albedo / r^4.  Contributions are accumulated in
// LDS as 2^-32 fixed point in 64-bit integers, so the order in which threads
// add them does not change the result (deterministic).
constexpr double kRenderFix = 4294967296.0;

__global__ void render_confocal_kernel(const fbp_vec3* __restrict__ scan, uint32_t num_scan,
                                       const fbp_vec3* __restrict__ pts, const float* __restrict__ alb,
                                       uint32_t num_points, uint32_t T, float delta,
                                       float* __restrict__ transient) {
    extern __shared__ unsigned long long hist[];
    const uint32_t l = blockIdx.x;
    if (l >= num_scan) return;
    for (uint32_t t = threadIdx.x; t < T; t += blockDim.x) hist[t] = 0ull;
    __syncthreads();
    const float sx = scan[l].x, sy = scan[l].y, sz = scan[l].z;
    for (uint32_t p = threadIdx.x; p < num_points; p += blockDim.x) {
        const float dx = pts[p].x - sx, dy = pts[p].y - sy, dz = pts[p].z - sz;
        const float r2 = dx * dx + dy * dy + dz * dz;
        const float r = sqrtf(r2);
        const int bin = int(rintf(2.0f * r / delta));
        if (bin < 0 || bin >= int(T) || !(r2 > 0.f)) continue;
        const double v = double(alb[p]) / (double(r2) * double(r2));
        atomicAdd(&hist[bin], (unsigned long long)(v * kRenderFix + 0.5));
    }
    __syncthreads();
    float* row = transient + size_t(l) * T;
    for (uint32_t t = threadIdx.x; t < T; t += blockDim.x) row[t] = float(double(hist[t]) / kRenderFix);
}

// One workgroup per (laser, sensor) pair: I[l][s][round((|l-p| + |p-s|) / delta)]
// += albedo / (|l-p|^2 |p-s|^2), fixed-point LDS accumulation as above.
__global__ void render_exhaustive_kernel(const fbp_vec3* __restrict__ lasers,
                                         const fbp_vec3* __restrict__ sensors, uint32_t S,
                                         const fbp_vec3* __restrict__ pts,
                                         const float* __restrict__ alb, uint32_t num_points,
                                         uint32_t T, float delta, float* __restrict__ transient) {
    extern __shared__ unsigned long long hist[];
    const uint64_t pair = blockIdx.x;
    const uint32_t l = uint32_t(pair / S), s = uint32_t(pair % S);
    for (uint32_t t = threadIdx.x; t < T; t += blockDim.x) hist[t] = 0ull;
    __syncthreads();
    const float lx = lasers[l].x, ly = lasers[l].y, lz = lasers[l].z;
    const float sx = sensors[s].x, sy = sensors[s].y, sz = sensors[s].z;
    for (uint32_t p = threadIdx.x; p < num_points; p += blockDim.x) {
        const float ax = pts[p].x - lx, ay = pts[p].y - ly, az = pts[p].z - lz;
        const float bx = pts[p].x - sx, by = pts[p].y - sy, bz = pts[p].z - sz;
        const float ra2 = ax * ax + ay * ay + az * az, rb2 = bx * bx + by * by + bz * bz;
        const int bin = int(rintf((sqrtf(ra2) + sqrtf(rb2)) / delta));
        if (bin < 0 || bin >= int(T) || !(ra2 > 0.f) || !(rb2 > 0.f)) continue;
        const double v = double(alb[p]) / (double(ra2) * double(rb2));
        atomicAdd(&hist[bin], (unsigned long long)(v * kRenderFix + 0.5));
    }
    __syncthreads();
    float* row = transient + pair * T;
    for (uint32_t t = threadIdx.x; t < T; t += blockDim.x) row[t] = float(double(hist[t]) / kRenderFix);
}

// ----------------------------------------------------- forward-model residual
// For scan point rows[r]: render the volume's voxels above `thr` (albedo
// v - thr, at their centres; with polarity -1 the voxels below it, albedo
// thr - v) with the same 1/r^4 model, box-blur both that
// and the measured row over +-blur bins, and accumulate sum p, sum m,
// sum pm, sum p^2, sum m^2 over the bins in FP64 (one workgroup per row
// writes its five sums).
__global__ void residual_row_kernel(const float* __restrict__ volume, VolumeGeom g,
                                    const float* __restrict__ mm, float thr_frac, float polarity,
                                    const fbp_vec3* __restrict__ scan, const uint32_t* __restrict__ rows,
                                    const float* __restrict__ I, uint32_t T, float delta, int blur,
                                    double* __restrict__ sums) {
    extern __shared__ unsigned long long hist[];
    float* pm = reinterpret_cast<float*>(hist + T);  // [2][T] blurred rows
    const uint32_t r = blockIdx.x;
    const uint32_t l = rows[r];
    for (uint32_t t = threadIdx.x; t < T; t += blockDim.x) hist[t] = 0ull;
    __syncthreads();
    const float mn = mm[0], mx = mm[1];
    const float thr = mn + thr_frac * (mx - mn);
    const float sx = scan[l].x, sy = scan[l].y, sz = scan[l].z;
    const size_t n = size_t(g.res[0]) * g.res[1] * g.res[2];
    for (size_t v = threadIdx.x; v < n; v += blockDim.x) {
        const float a = polarity * (volume[v] - thr);
        if (!(a > 0.f)) continue;
        const uint32_t vx = uint32_t(v % g.res[0]), vy = uint32_t((v / g.res[0]) % g.res[1]),
                       z = uint32_t(v / (size_t(g.res[0]) * g.res[1]));
        const float dx = voxel_center(g.hmin[0], g.vs[0], vx) - sx;
        const float dy = voxel_center(g.hmin[1], g.vs[1], z) - sy;
        const float dz = voxel_center(g.hmin[2], g.vs[2], vy) - sz;
        const float r2 = dx * dx + dy * dy + dz * dz;
        const int bin = int(rintf(2.0f * sqrtf(r2) / delta));
        if (bin < 0 || bin >= int(T) || !(r2 > 0.f)) continue;
        atomicAdd(&hist[bin], (unsigned long long)(double(a) / (double(r2) * r2) * kRenderFix + 0.5));
    }
    __syncthreads();
    for (uint32_t t = threadIdx.x; t < T; t += blockDim.x) {
        double sp = 0.0, sm = 0.0;
        for (int k = -blur; k <= blur; ++k) {
            const int u = int(t) + k;
            if (u < 0 || u >= int(T)) continue;
            sp += double(hist[u]) / kRenderFix;
            sm += double(I[size_t(l) * T + u]);
        }
        pm[t] = float(sp);
        pm[T + t] = float(sm);
    }
    __syncthreads();
    __shared__ double red[5][256];
    double q5[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    for (uint32_t t = threadIdx.x; t < T; t += blockDim.x) {
        const double p = pm[t], m = pm[T + t];
        q5[0] += p;
        q5[1] += m;
        q5[2] += p * m;
        q5[3] += p * p;
        q5[4] += m * m;
    }
    for (int q = 0; q < 5; ++q) red[q][threadIdx.x] = q5[q];
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if (int(threadIdx.x) < w)
            for (int q = 0; q < 5; ++q) red[q][threadIdx.x] += red[q][threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int q = 0; q < 5; ++q) sums[size_t(r) * 5 + q] = red[q][0];
}

}  // namespace fbp

using namespace fbp;

extern "C" {

fbp_status fbp_backproject_depth_confocal(fbp_context ctx, const fbp_reconstruction_info* info,
                                          const fbp_transient* tr, const fbp_vec3* laser_targets,
                                          const double* depths, uint32_t num_depths,
                                          float* activation, void* stream) {
    if (!ctx || !info || !tr || !tr->data || !laser_targets || !depths || !activation)
        return fail(ctx, FBP_ERR_INVALID_ARGUMENT, "depth confocal: null argument");
    if (tr->storage != FBP_STORAGE_F32 || tr->layout != FBP_LAYOUT_TIME_FASTEST)
        return fail(ctx, FBP_ERR_UNSUPPORTED, "depth confocal: F32 [l][t] transient only");
    if (!num_depths || !info->num_laser_targets || !info->num_time_bins || !(info->time_step > 0.f))
        return fail(ctx, FBP_ERR_INVALID_ARGUMENT, "depth confocal: empty grid");
    if (uint64_t(info->num_laser_targets) * num_depths > 0xffffffffull)
        return fail(ctx, FBP_ERR_UNSUPPORTED, "depth confocal: more than 2^32 - 1 outputs");
    FBP_CHECK(bind(ctx));
    const uint32_t L = info->num_laser_targets;
    const uint32_t n = L * num_depths;
    hipStream_t s = as_stream(stream);
    // Aligned scan grid: ring convolution per depth plane (decided on the
    // device by the grid check; the gather below returns at once then).
    const AlignedPlan ap = plan_aligned_depth(ctx, *info, *tr, num_depths);
    const int* skip = nullptr;
    ctx->path_flag_valid = ap.usable;
    if (ap.usable) {
        FBP_CHECK(ensure_scratch(ctx, aligned_scratch_bytes(ap)));
        if (!ctx->path_flag)
            FBP_CHECK_HIP(ctx, hipMalloc(reinterpret_cast<void**>(&ctx->path_flag), 256));
        FBP_CHECK(launch_aligned_depth(ctx, ap, *info, *tr, laser_targets, depths, num_depths,
                                       activation, ctx->path_flag, ctx->scratch, s));
        skip = ctx->path_flag;
    }
    hipLaunchKernelGGL(depth_kernel, dim3(unsigned((uint64_t(n) + 255) / 256)), dim3(256), 0, s,
                       static_cast<const float*>(tr->data), laser_targets, depths, num_depths, L,
                       info->num_time_bins, info->time_offset, info->time_step,
                       info->discard_first_last_bounces, info->relay_wall_normal,
                       info->laser_position, info->sensor_position, activation, skip);
    FBP_CHECK_HIP(ctx, hipGetLastError());
    return FBP_OK;
}

fbp_status fbp_build_alias_tables_host(const float* cdf, uint32_t n, uint32_t* alias_table,
                                       float* prob_table) {
    // Laser::buildAliasTables (FB/Laser.cu:253-315): same arithmetic, same
    // stack (pop from the back) order.
    if (!cdf || !alias_table || !prob_table || n == 0) return FBP_ERR_INVALID_ARGUMENT;
    std::vector<float> p(n);
    p[0] = cdf[0];
    for (uint32_t i = 1; i < n; ++i) p[i] = cdf[i] - cdf[i - 1];
    std::vector<uint32_t> small_, large_;
    float sum = 0.f;
    for (uint32_t i = 0; i < n; ++i) {
        alias_table[i] = 0;
        if (p[i] < 0.f) return FBP_ERR_PROBABILITIES;
        prob_table[i] = p[i] * float(n);
        sum += p[i];
        if (prob_table[i] < 1.0f)
            small_.push_back(i);
        else
            large_.push_back(i);
    }
    if (std::fabs(sum - 1.0f) > 1e-5f) return FBP_ERR_PROBABILITIES;
    while (!small_.empty() && !large_.empty()) {
        const uint32_t s = small_.back();
        small_.pop_back();
        const uint32_t l = large_.back();
        large_.pop_back();
        alias_table[s] = l;
        prob_table[l] += prob_table[s] - 1.0f;
        if (prob_table[l] < 1.0f)
            small_.push_back(l);
        else
            large_.push_back(l);
    }
    while (!small_.empty()) {
        prob_table[small_.back()] = 1.0f;
        small_.pop_back();
    }
    while (!large_.empty()) {
        prob_table[large_.back()] = 1.0f;
        large_.pop_back();
    }
    return FBP_OK;
}

fbp_status fbp_backproject_confocal_mis(fbp_context ctx, const fbp_reconstruction_info* info,
                                        const fbp_transient* tr, const fbp_vec3* laser_targets,
                                        const fbp_mis_params* params, float* volume, void* stream) {
    if (!ctx || !info || !tr || !tr->data || !laser_targets || !params || !volume ||
        (!params->noise && params->mode != FBP_MIS_TILED))
        return fail(ctx, FBP_ERR_INVALID_ARGUMENT, "mis: null argument");
    if (tr->storage != FBP_STORAGE_F32 || tr->layout != FBP_LAYOUT_TIME_FASTEST)
        return fail(ctx, FBP_ERR_UNSUPPORTED, "mis: F32 [l][t] transient only");
    if (!params->noise_size || !params->num_samples || !info->num_laser_targets ||
        !info->num_time_bins || !(info->time_step > 0.f) || params->mode > FBP_MIS_TILED)
        return fail(ctx, FBP_ERR_INVALID_ARGUMENT, "mis: bad parameters");
    FBP_CHECK(bind(ctx));
    hipStream_t s = as_stream(stream);
    const uint32_t L = info->num_laser_targets, T = info->num_time_bins;
    if (params->mode == FBP_MIS_TILED) {
        // scratch2: row sums (double L) | prefix (double L) | samples (u32, f32 per tile)
        const uint32_t tiles = make_geom(*info, 4, 4, 4).num_tiles;
        const size_t ns = size_t(tiles) * params->num_samples;
        FBP_CHECK(ensure_scratch2(ctx, size_t(L) * 16 + ns * 8 + 512));
        char* base = static_cast<char*>(ctx->scratch2);
        double* d_sum = reinterpret_cast<double*>(base);
        double* d_cum = d_sum + L;
        uint32_t* d_l = reinterpret_cast<uint32_t*>(d_cum + L);
        float* d_w = reinterpret_cast<float*>(d_l + ns);
        hipLaunchKernelGGL(row_sum_kernel, dim3(L), dim3(256), 0, s,
                           static_cast<const float*>(tr->data), L, T, d_sum);
        // Short row lists: every samples block builds the prefix in LDS.
        const bool lds_prefix = L <= kMisLdsRows;
        if (!lds_prefix)
            hipLaunchKernelGGL(row_prefix_kernel, dim3(1), dim3(1024), 0, s, d_sum, L, d_cum);
        hipLaunchKernelGGL(mis_tile_samples_kernel, dim3(unsigned((ns + 255) / 256)), dim3(256),
                           lds_prefix ? size_t(L) * sizeof(double) : 0, s, d_sum,
                           lds_prefix ? nullptr : d_cum, L, params->noise, params->noise_size,
                           params->seed, params->num_samples, tiles, d_l, d_w);
        FBP_CHECK_HIP(ctx, hipGetLastError());
        return confocal_voxel_impl(ctx, info, tr, laser_targets, volume, stream, d_l, d_w,
                                   params->num_samples);
    }
    // Scratch: row sums (double L) | cdf (float L) | alias (u32 L) | prob (float L)
    const size_t bytes = size_t(L) * (8 + 4 + 4 + 4) + 256;
    FBP_CHECK(ensure_scratch(ctx, bytes));
    char* base = static_cast<char*>(ctx->scratch);
    double* d_sum = reinterpret_cast<double*>(base);
    float* d_cdf = reinterpret_cast<float*>(base + size_t(L) * 8);
    uint32_t* d_alias = reinterpret_cast<uint32_t*>(base + size_t(L) * 12);
    float* d_prob = reinterpret_cast<float*>(base + size_t(L) * 16);
    hipLaunchKernelGGL(row_sum_kernel, dim3(L), dim3(256), 0, s, static_cast<const float*>(tr->data),
                       L, T, d_sum);
    FBP_CHECK_HIP(ctx, hipGetLastError());
    // Spatial CDF: inclusive prefix over scan points, then min-max normalised
    // (FB/Laser.cu:547-574, scanning the intensity rather than the volume).
    std::vector<double> h_sum(L);
    FBP_CHECK_HIP(ctx, hipMemcpyAsync(h_sum.data(), d_sum, size_t(L) * 8, hipMemcpyDeviceToHost, s));
    FBP_CHECK_HIP(ctx, hipStreamSynchronize(s));
    std::vector<double> pre(L);
    double run = 0.0;
    for (uint32_t l = 0; l < L; ++l) {
        run += h_sum[l];
        pre[l] = run;
    }
    const double mn = *std::min_element(pre.begin(), pre.end());
    const double mx = *std::max_element(pre.begin(), pre.end());
    std::vector<float> cdf(L);
    for (uint32_t l = 0; l < L; ++l)
        cdf[l] = (mx > mn) ? float((pre[l] - mn) / (mx - mn)) : 0.f;
    std::vector<uint32_t> alias(L);
    std::vector<float> prob(L);
    fbp_status st = fbp_build_alias_tables_host(cdf.data(), L, alias.data(), prob.data());
    if (st != FBP_OK) return fail(ctx, st, "mis: alias tables rejected the scan-point CDF");
    FBP_CHECK_HIP(ctx, hipMemcpyAsync(d_cdf, cdf.data(), size_t(L) * 4, hipMemcpyHostToDevice, s));
    FBP_CHECK_HIP(ctx, hipMemcpyAsync(d_alias, alias.data(), size_t(L) * 4, hipMemcpyHostToDevice, s));
    FBP_CHECK_HIP(ctx, hipMemcpyAsync(d_prob, prob.data(), size_t(L) * 4, hipMemcpyHostToDevice, s));
    VolumeGeom g = make_geom(*info, 1, 1, 1);
    const size_t nv = size_t(g.res[0]) * g.res[1] * g.res[2];
    // Wave patches: rows x (64 / rows) voxels, as square as the state window
    // allows (each lane is active for N of the wave's N + (rows - 1) X +
    // 64 / rows - 1 states; squarer patches read fewer transient lines).
    uint32_t rows = 1;
    while (rows < 8 && 4 * (2 * rows - 1) * size_t(g.res[0]) <= params->num_samples) rows *= 2;
    // States 0 .. nv + N + rows X + 64 (a patch's walk may run past nv - 1).
    const size_t nstate = nv + params->num_samples + size_t(rows) * g.res[0] + 64;
    FBP_CHECK(ensure_scratch2(ctx, nstate * sizeof(MisState) + 256));
    MisState* states = static_cast<MisState*>(ctx->scratch2);
    hipLaunchKernelGGL(mis_state_kernel, dim3(unsigned((nstate + 255) / 256)), dim3(256), 0, s,
                       d_cdf, d_alias, d_prob, params->noise, params->noise_size, L,
                       params->num_samples, params->mode, info->discard_first_last_bounces,
                       laser_targets, info->laser_position, info->sensor_position, nstate, states);
    FBP_CHECK_HIP(ctx, hipGetLastError());
    const size_t waves = size_t((g.res[0] + 64 / rows - 1) / (64 / rows)) *
                         ((g.res[1] + rows - 1) / rows) * g.res[2];
    const dim3 grid(unsigned((waves + 3) / 4));
    const float ext = voxel_extent(info->hidden_volume_voxel_size.x,
                                   info->hidden_volume_voxel_size.y,
                                   info->hidden_volume_voxel_size.z);
    const bool ref = params->mode == FBP_MIS_REFERENCE;
    const float* I = static_cast<const float*>(tr->data);
    const bool pure = info->time_offset == 0.f && info->discard_first_last_bounces != 0 &&
                      !ctx->voxel_scattering && (size_t(T) + 1) * sizeof(float) <= 65536;
    const float* thr = nullptr;
    if (pure) FBP_CHECK(pure_threshold_row(ctx, T, info->time_step, s, &thr));
    const size_t shmem = pure ? (size_t(T) + 1) * sizeof(float) : 0;
#define FBP_MIS_WALK(REF_, SCAT_, PURE_)                                                          \
    hipLaunchKernelGGL((mis_walk_kernel<REF_, SCAT_, PURE_>), grid, dim3(256), shmem, s, I,      \
                       laser_targets, states, params->num_samples, g, T, info->time_offset,      \
                       info->time_step, info->discard_first_last_bounces, ext, rows, thr, volume)
    if (ctx->voxel_scattering) {
        if (ref)
            FBP_MIS_WALK(true, true, false);
        else
            FBP_MIS_WALK(false, true, false);
    } else if (pure) {
        if (ref)
            FBP_MIS_WALK(true, false, true);
        else
            FBP_MIS_WALK(false, false, true);
    } else {
        if (ref)
            FBP_MIS_WALK(true, false, false);
        else
            FBP_MIS_WALK(false, false, false);
    }
#undef FBP_MIS_WALK
    FBP_CHECK_HIP(ctx, hipGetLastError());
    // The host vectors above are consumed by the async copies: wait for them.
    FBP_CHECK_HIP(ctx, hipStreamSynchronize(s));
    return FBP_OK;
}

fbp_status fbp_quantize_transient(fbp_context ctx, const float* src, uint64_t rows, uint32_t cols,
                                  int32_t storage, void* dst, float* scales, void* stream) {
    if (!ctx || !src || !dst || !scales || rows == 0 || cols == 0)
        return fail(ctx, FBP_ERR_INVALID_ARGUMENT, "quantize: bad argument");
    if (storage != FBP_STORAGE_F16 && storage != FBP_STORAGE_BF16)
        return fail(ctx, FBP_ERR_INVALID_ARGUMENT, "quantize: storage must be F16 or BF16");
    if (rows > 0x7fffffffull) return fail(ctx, FBP_ERR_UNSUPPORTED, "quantize: too many rows");
    FBP_CHECK(bind(ctx));
    if (storage == FBP_STORAGE_F16)
        hipLaunchKernelGGL(quantize_kernel<FBP_STORAGE_F16>, dim3(unsigned(rows)), dim3(256), 0,
                           as_stream(stream), src, cols, dst, scales);
    else
        hipLaunchKernelGGL(quantize_kernel<FBP_STORAGE_BF16>, dim3(unsigned(rows)), dim3(256), 0,
                           as_stream(stream), src, cols, dst, scales);
    FBP_CHECK_HIP(ctx, hipGetLastError());
    return FBP_OK;
}

fbp_status fbp_render_exhaustive(fbp_context ctx, const fbp_vec3* laser_points, uint32_t num_lasers,
                                 const fbp_vec3* sensor_points, uint32_t num_sensors,
                                 const fbp_vec3* points, const float* albedo, uint32_t num_points,
                                 uint32_t num_time_bins, float delta, float* transient,
                                 void* stream) {
    if (!ctx || !laser_points || !sensor_points || !points || !albedo || !transient ||
        !num_lasers || !num_sensors || !num_time_bins || !(delta > 0.f))
        return fail(ctx, FBP_ERR_INVALID_ARGUMENT, "render_exhaustive: bad argument");
    if (size_t(num_time_bins) * 8 > 65536)
        return fail(ctx, FBP_ERR_UNSUPPORTED, "render_exhaustive: at most 8192 time bins");
    FBP_CHECK(bind(ctx));
    const uint64_t pairs = uint64_t(num_lasers) * num_sensors;
    if (pairs >= (1ull << 31))
        return fail(ctx, FBP_ERR_UNSUPPORTED, "render_exhaustive: at most 2^31 - 1 histograms");
    hipLaunchKernelGGL(render_exhaustive_kernel, dim3(unsigned(pairs)), dim3(256),
                       size_t(num_time_bins) * 8, as_stream(stream), laser_points, sensor_points,
                       num_sensors, points, albedo, num_points, num_time_bins, delta, transient);
    FBP_CHECK_HIP(ctx, hipGetLastError());
    return FBP_OK;
}

fbp_status fbp_volume_residual(fbp_context ctx, const fbp_reconstruction_info* info,
                               const float* volume, const float* transient,
                               const fbp_vec3* laser_targets, const uint32_t* rows,
                               uint32_t num_rows, float threshold_fraction, int32_t polarity,
                               int32_t blur, double* host_residual, void* stream) {
    if (!ctx || !info || !volume || !transient || !laser_targets || !rows || !num_rows ||
        !host_residual || blur < 0 || (polarity != 1 && polarity != -1))
        return fail(ctx, FBP_ERR_INVALID_ARGUMENT, "residual: bad argument");
    const uint32_t T = info->num_time_bins;
    if (size_t(T) * 16 > 65536) return fail(ctx, FBP_ERR_UNSUPPORTED, "residual: at most 4096 bins");
    FBP_CHECK(bind(ctx));
    hipStream_t s = as_stream(stream);
    VolumeGeom g = make_geom(*info, 1, 1, 1);
    const size_t nv = size_t(g.res[0]) * g.res[1] * g.res[2];
    float mnv, mxv;
    FBP_CHECK(fbp_minmax(ctx, volume, nv, &mnv, &mxv, stream));
    // minmax leaves {min, max} on the device in its scratch block; keep our
    // own copy after it for the kernel.
    FBP_CHECK(ensure_scratch2(ctx, size_t(num_rows) * 5 * sizeof(double) + 64));
    double* sums = static_cast<double*>(ctx->scratch2);
    float* mm = reinterpret_cast<float*>(sums + size_t(num_rows) * 5);
    const float hmm[2] = {mnv, mxv};
    FBP_CHECK_HIP(ctx, hipMemcpyAsync(mm, hmm, sizeof(hmm), hipMemcpyHostToDevice, s));
    const size_t shmem = size_t(T) * 8 + size_t(T) * 2 * 4;
    hipLaunchKernelGGL(residual_row_kernel, dim3(num_rows), dim3(256), shmem, s, volume, g, mm,
                       threshold_fraction, float(polarity), laser_targets, rows, transient, T, info->time_step, blur,
                       sums);
    FBP_CHECK_HIP(ctx, hipGetLastError());
    std::vector<double> h(size_t(num_rows) * 5);
    FBP_CHECK_HIP(ctx, hipMemcpyAsync(h.data(), sums, h.size() * 8, hipMemcpyDeviceToHost, s));
    FBP_CHECK_HIP(ctx, hipStreamSynchronize(s));
    // Per-row centred second moments, pooled over the rows: 1 - Pearson
    // correlation of rendered and measured rows.
    double cov = 0.0, vp = 0.0, vm = 0.0;
    for (uint32_t r = 0; r < num_rows; ++r) {
        const double* q = &h[size_t(r) * 5];
        cov += q[2] - q[0] * q[1] / T;
        vp += q[3] - q[0] * q[0] / T;
        vm += q[4] - q[1] * q[1] / T;
    }
    *host_residual = (vp > 0.0 && vm > 0.0) ? 1.0 - cov / std::sqrt(vp * vm) : 1.0;
    return FBP_OK;
}

fbp_status fbp_render_confocal(fbp_context ctx, const fbp_vec3* scan_points, uint32_t num_scan,
                               const fbp_vec3* points, const float* albedo, uint32_t num_points,
                               uint32_t num_time_bins, float delta, float* transient,
                               void* stream) {
    if (!ctx || !scan_points || !points || !albedo || !transient || !num_scan || !num_time_bins ||
        !(delta > 0.f))
        return fail(ctx, FBP_ERR_INVALID_ARGUMENT, "render: bad argument");
    FBP_CHECK(bind(ctx));
    if (size_t(num_time_bins) * 8 > 65536)
        return fail(ctx, FBP_ERR_UNSUPPORTED, "render: at most 8192 time bins");
    hipLaunchKernelGGL(render_confocal_kernel, dim3(num_scan), dim3(256),
                       size_t(num_time_bins) * 8, as_stream(stream),
                       scan_points, num_scan, points, albedo, num_points, num_time_bins, delta,
                       transient);
    FBP_CHECK_HIP(ctx, hipGetLastError());
    return FBP_OK;
}

}  // extern "C"
