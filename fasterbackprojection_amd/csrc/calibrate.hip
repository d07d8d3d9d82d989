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
// No host round trip; four launches: one pass over the volume (the fit's
// sums and each block's extreme candidates), one block selecting the k of
// each side, the exact neighbourhoods, and the affine map (each block
// solving the small fit itself).
#include <algorithm>
#include <vector>

#include "bp_common.h"
#include "topk.h"

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

// ---- extremes: (value, index) pairs as 64-bit keys whose ascending order
// is value ascending (lows) or descending (highs), ties by smaller index.
constexpr int kScanChunk = 512;     // voxels per selection round (2 per thread)
constexpr uint32_t kMaxScanBlocks = 1024;
constexpr uint64_t kPadKey = kTopkPad;

__device__ __forceinline__ uint32_t ordered_bits(float v) {
    const uint32_t u = __float_as_uint(v);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ uint64_t key_low(float v, uint32_t i) {
    return (uint64_t(ordered_bits(v)) << 32) | i;
}
__device__ __forceinline__ uint64_t key_high(float v, uint32_t i) {
    return (uint64_t(~ordered_bits(v)) << 32) | i;
}

// One pass over the volume, blockIdx.y = side (0: lows, 1: highs).  Block
// b owns voxels [b seg, (b + 1) seg): its k keys of the side,
// cand[side][b][0..k), over 512-voxel chunks merged with the best k so far;
// the lows' blocks also form the least-squares sums of reference r on
// volume v in FP64, sums[b][0..4] = (n, sum v, sum v^2, sum r, sum v r).
__global__ __launch_bounds__(256) void calib_scan_kernel(
    const float* __restrict__ v, const float* __restrict__ r, uint32_t n, uint32_t seg,
    uint32_t k, uint64_t* __restrict__ cand, double* __restrict__ sums) {
    __shared__ uint64_t best[kMaxExtremes];
    __shared__ double red[5][256];
    const uint32_t t = threadIdx.x, b = blockIdx.x, nb = gridDim.x;
    const bool high = blockIdx.y == 1;
    const uint32_t begin = b * seg, end = min(n, begin + seg);
    double acc[5] = {0, 0, 0, 0, 0};
    bool first = true;
    for (uint32_t c0 = begin; c0 < end; c0 += kScanChunk) {
        uint64_t kk[2];
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            const uint32_t i = c0 + t + 256u * uint32_t(j);
            kk[j] = kPadKey;
            if (i < end) {
                const float x = v[i];
                if (!high) {
                    const double xd = x, yd = r[i];
                    acc[0] += 1.0;
                    acc[1] += xd;
                    acc[2] += xd * xd;
                    acc[3] += yd;
                    acc[4] += xd * yd;
                }
                kk[j] = high ? key_high(x, i) : key_low(x, i);
            }
        }
        // This chunk's keys and the best k so far (one more key per thread).
        uint64_t kc[3] = {kk[0], kk[1], (!first && t < k) ? best[t] : kPadKey};
        __syncthreads();  // best is read above and rewritten below
        block_select_k<3>(kc, k, best);
        first = false;
    }
    if (t < k) cand[(size_t(high ? nb : 0) + b) * k + t] = best[t];
    if (high) return;
    for (int q = 0; q < 5; ++q) red[q][t] = acc[q];
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (int(t) < w)
            for (int q = 0; q < 5; ++q) red[q][t] += red[q][t + w];
        __syncthreads();
    }
    if (t == 0)
        for (int q = 0; q < 5; ++q) sums[size_t(b) * 5 + q] = red[q][0];
}

// Two blocks after the scan over nb blocks, one per side (blockIdx.x = 0:
// lows, 1: highs): the global k of the side: the k blocks with the smallest
// first keys hold all of them (a block outside those has k smaller keys
// before its first), so the k smallest of their k x k keys are the answer;
// sel[side * k + i], i < k.  Block 0 also reduces the fit sums in a fixed
// order into fit[0..4].
__global__ __launch_bounds__(256) void calib_select_kernel(const uint64_t* __restrict__ cand,
                                                           const double* __restrict__ sums,
                                                           uint32_t nb, uint32_t seg, uint32_t k,
                                                           uint32_t* __restrict__ sel,
                                                           double* __restrict__ fit) {
    __shared__ double red[5][256];
    __shared__ uint64_t sk[kMaxExtremes];
    __shared__ uint32_t blk[kMaxExtremes];
    const uint32_t t = threadIdx.x, side = blockIdx.x;
    const uint64_t* c = cand + size_t(side) * nb * k;
    // The first key of every block (nb <= 1024: 4 per thread).
    uint64_t a[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const uint32_t i = t + 256u * uint32_t(j);
        a[j] = i < nb ? c[size_t(i) * k] : kPadKey;
    }
    if (side == 0) {
        // Fit sums: thread t adds blocks t, t + 256, ... in order, then the tree.
        double s5[5] = {0, 0, 0, 0, 0};
        for (uint32_t i = t; i < nb; i += 256)
            for (int q = 0; q < 5; ++q) s5[q] += sums[size_t(i) * 5 + q];
        for (int q = 0; q < 5; ++q) red[q][t] = s5[q];
        __syncthreads();
        for (int w = 128; w > 0; w >>= 1) {
            if (int(t) < w)
                for (int q = 0; q < 5; ++q) red[q][t] += red[q][t + w];
            __syncthreads();
        }
        if (t < 5) fit[t] = red[t][0];
    }
    block_select_k<4>(a, k, sk);
    // (fewer than k blocks: the pads select no block)
    constexpr uint32_t kNoBlock = 0xffffffffu;
    if (t < k) blk[t] = sk[t] == kPadKey ? kNoBlock : uint32_t(sk[t] & 0xffffffffu) / seg;
    __syncthreads();
    // Their k x k keys (k <= 32: 4 per thread).
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const uint32_t i = t + 256u * uint32_t(j);
        a[j] = kPadKey;
        if (i < k * k && blk[i / k] != kNoBlock) a[j] = c[size_t(blk[i / k]) * k + i % k];
    }
    block_select_k<4>(a, k, sk);
    if (t < k) sel[side * k + t] = uint32_t(sk[t] & 0xffffffffu);
}

// Exact LoG contribution of one chunk of scan points at candidate ci
// (ci < k: lows, else highs): block (ci, chunk); thread t < 125
// backprojects neighbour t = (dz, dy, dx) of the candidate (clamped to the
// volume) over the chunk, eight scan points per round with every load of
// the round issued first.  All threads read the same scan point's row at
// nearby bins (coalesced).  partial[ci * chunks + chunk] = sum_t w[t] bp_t
// (FP64, fixed order).
// pthr (pure chain, t0 = 0 and bounces discarded; else null): the
// threshold row S_k, staged in LDS; each bin is then found as in the tile
// kernels (a v_rsq estimate biased low and one threshold compare, exact),
// with weight 1 / (2 r)^2 from the same estimate; samples within 4.5 mm of
// dist 0 take point_sample's IEEE chain.  256 threads: tap t % 128 (< 125)
// over half t / 128 of the chunk, kNeighU samples per round with every load
// of the round issued first; the chunk's scan points staged in LDS.
constexpr int kNeighU = 16;
constexpr uint32_t kNeighMaxChunk = 128;

// Backprojection of voxel (px, py, pz) over scan points [h0, h1) (positions
// in slp, indexed from l0), FP64 sum.  No branch between the loads of a
// round: a near sample (PURE) is masked there and redone afterwards.
template <bool PURE>
__device__ __forceinline__ double neigh_samples(const PointArgs& a, const float (*slp)[3],
                                                const float* sthr, uint32_t l0, uint32_t h0,
                                                uint32_t h1, float px, float py, float pz) {
    const int T = int(a.T);
    // Bias of the estimate: 2^-18 relative, far above its error.
    const float cb = (1.0f / a.dt) * (1.0f - 3.814697265625e-06f);
    double acc = 0.0;
    bool any_near = false;
    for (uint32_t lb = h0; lb < h1; lb += kNeighU) {
        float light[kNeighU], wt[kNeighU];
#pragma unroll
        for (int u = 0; u < kNeighU; ++u) {
            const uint32_t l = min(lb + uint32_t(u), h1 - 1);
            const float lx = slp[l - l0][0], ly = slp[l - l0][1], lz = slp[l - l0][2];
            const bool live = lb + uint32_t(u) < h1;
            if constexpr (PURE) {
                // 4 x the exact squared distance, glm::dot order.
                const float A = rn_add(sq(rn_sub(px, lx)) * 4.0f, sq(rn_sub(py, ly)) * 4.0f);
                const float s4 = A + sq(rn_sub(pz, lz)) * 4.0f;
                const float rs = __builtin_amdgcn_rsqf(s4);
                const int ka = int((s4 * rs) * cb);
                const int kb = min(max(ka + 1, 0), T);
                const float th = sthr[kb];
                const int kk = ka + int((ka + 1 <= T) & (s4 >= th));
                const bool near = s4 < 2.1e-5f;
                any_near |= live && near;
                const bool ok = live & !near & (kk >= 0) & (kk < T);
                const float x = a.I[size_t(l) * a.T + uint32_t(ok ? kk : 0)];
                light[u] = (ok & (x > kEps)) ? x : 0.f;
                wt[u] = rs * rs;
            } else {
                // point_sample's arithmetic.
                float dist = rn_sub(rn_mul(distance3(lx, ly, lz, px, py, pz), 2.0f), a.t0);
                if (a.discard == 0)
                    dist = rn_add(dist, rn_add(distance3(lx, ly, lz, a.lpos.x, a.lpos.y, a.lpos.z),
                                               distance3(lx, ly, lz, a.spos.x, a.spos.y, a.spos.z)));
                const float q = rn_div(dist, a.dt);
                const bool ok = live & (q > -1.0f) & (q < float(a.T));
                const float x = a.I[size_t(l) * a.T + uint32_t(ok ? int(q) : 0)];
                light[u] = (ok & (x > kEps)) ? x : 0.f;
                wt[u] = safe_rcp(rn_mul(dist, dist));
            }
        }
#pragma unroll
        for (int u = 0; u < kNeighU; ++u)
            acc += PURE ? double(light[u] * wt[u]) : double(rn_mul(light[u], wt[u]));
    }
    if (PURE && any_near) {
        // dist within 4.5 mm of 0: the IEEE chain (rare).
        for (uint32_t l = h0; l < h1; ++l) {
            const float lx = slp[l - l0][0], ly = slp[l - l0][1], lz = slp[l - l0][2];
            const float A = rn_add(sq(rn_sub(px, lx)) * 4.0f, sq(rn_sub(py, ly)) * 4.0f);
            if (A + sq(rn_sub(pz, lz)) * 4.0f < 2.1e-5f) acc += point_sample(a, l, px, py, pz);
        }
    }
    return acc;
}

__global__ __launch_bounds__(256) void calib_neigh_kernel(PointArgs a, Weights w,
                                                          const uint32_t* __restrict__ sel,
                                                          uint32_t k, uint32_t chunk_len,
                                                          const float* __restrict__ pthr,
                                                          double* __restrict__ partial) {
    __shared__ double red[256];
    __shared__ float slp[kNeighMaxChunk][3];
    extern __shared__ float sthr[];  // T + 1 thresholds x 4 (pure chain)
    const uint32_t ci = blockIdx.x, chunk = blockIdx.y, t = threadIdx.x;
    const uint32_t tap = t & 127, half = t >> 7;
    const uint32_t l0 = chunk * chunk_len, l1 = min(a.L, l0 + chunk_len);
    for (uint32_t i = t; i < (l1 - l0) * 3; i += 256) {
        const fbp_vec3 lp = a.laser[l0 + i / 3];
        slp[i / 3][i % 3] = (i % 3 == 0) ? lp.x : (i % 3 == 1) ? lp.y : lp.z;
    }
    if (pthr)
        for (uint32_t i = t; i <= a.T; i += 256) sthr[i] = pthr[i] * 4.0f;
    __syncthreads();
    double acc = 0.0;
    if (tap < uint32_t(kTaps)) {
        const uint32_t c = sel[ci];
        const int x = int(c % a.X), y = int((c / a.X) % a.Y), z = int(c / (a.X * a.Y));
        const int nx = min(max(x + int(tap % 5) - 2, 0), int(a.X) - 1);
        const int ny = min(max(y + int((tap / 5) % 5) - 2, 0), int(a.Y) - 1);
        const int nz = min(max(z + int(tap / 25) - 2, 0), int(a.Z) - 1);
        const float px = voxel_center(a.hmin[0], a.vs[0], uint32_t(nx));
        const float py = voxel_center(a.hmin[1], a.vs[1], uint32_t(nz));
        const float pz = voxel_center(a.hmin[2], a.vs[2], uint32_t(ny));
        const uint32_t hl = (l1 - l0 + 1) / 2;
        const uint32_t h0 = l0 + half * hl, h1 = min(l1, h0 + hl);
        if (pthr)
            acc = neigh_samples<true>(a, slp, sthr, l0, h0, h1, px, py, pz);
        else
            acc = neigh_samples<false>(a, slp, sthr, l0, h0, h1, px, py, pz);
        acc *= double(w.w[tap]);
    }
    red[t] = acc;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (int(t) < s) red[t] += red[t + s];
        __syncthreads();
    }
    if (t == 0) partial[size_t(ci) * gridDim.y + chunk] = red[0];
}

// Every block solves the fit (the same values in every block): L ~ a v + b
// from fit[0..4], the exact LoG of each candidate from its chunk partials,
// the true range over the candidates; then out = coef0 v + coef1 with
// coef = (a / range, (b - min) / range).
__global__ __launch_bounds__(256) void calib_apply_kernel(float* __restrict__ v, size_t n,
                                                          const double* __restrict__ fit,
                                                          const double* __restrict__ partial,
                                                          uint32_t k, uint32_t chunks) {
    __shared__ double lpart[2 * kMaxExtremes][4];
    __shared__ double lval[2 * kMaxExtremes];
    __shared__ float coef[2];
    const uint32_t t = threadIdx.x;
    // Candidate c's chunk partials: four threads, each a fixed quarter in
    // order, then the quarters in order (the same sum in every block).
    for (uint32_t i = t; i < 2 * k * 4; i += 256) {
        const uint32_t c = i / 4, q = i % 4;
        const uint32_t per = (chunks + 3) / 4, c0 = q * per, c1 = min(chunks, c0 + per);
        double l = 0.0;
        for (uint32_t j = c0; j < c1; ++j) l += partial[size_t(c) * chunks + j];
        lpart[c][q] = l;
    }
    __syncthreads();
    if (t < 2 * k) lval[t] = ((lpart[t][0] + lpart[t][1]) + lpart[t][2]) + lpart[t][3];
    __syncthreads();
    if (t == 0) {
        double S[5];
        for (int q = 0; q < 5; ++q) S[q] = fit[q];
        const double det = S[0] * S[2] - S[1] * S[1];
        const double av = det > 0.0 ? (S[0] * S[4] - S[1] * S[3]) / det : 0.0;
        const double bv = S[0] > 0.0 ? (S[3] - av * S[1]) / S[0] : 0.0;
        double mn = lval[0], mx = lval[k];
        for (uint32_t i = 1; i < k; ++i) mn = fmin(mn, lval[i]);
        for (uint32_t i = k + 1; i < 2 * k; ++i) mx = fmax(mx, lval[i]);
        const double range = mx - mn;
        // safeRCP of the range, as the min-max normalisation (FB/Laser.cu:317-339)
        const double inv = (range > 1e-5 || range < -1e-5) ? 1.0 / range : 3.402823466e+38;
        coef[0] = float(av * inv);
        coef[1] = float((bv - mn) * inv);
    }
    __syncthreads();
    const size_t i = size_t(blockIdx.x) * blockDim.x + t;
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

namespace fbp {

fbp_status learned_calibrate_impl(fbp_context ctx, const fbp_reconstruction_info* info,
                                  const fbp_transient* tr, const fbp_vec3* laser_targets,
                                  float* volume, const float* reference, int32_t kernel_size,
                                  float sigma, uint32_t extremes, hipStream_t s) {
    if (!ctx || !info || !tr || !tr->data || !laser_targets || !volume || !reference)
        return fail(ctx, FBP_ERR_INVALID_ARGUMENT, "calibrate: null argument");
    FBP_CHECK(check_confocal_f32(ctx, info, tr, "calibrate"));
    if (kernel_size != 5)
        return fail(ctx, FBP_ERR_UNSUPPORTED, "calibrate: LoG kernel size 5 only");
    if (extremes == 0 || extremes > uint32_t(kMaxExtremes))
        return fail(ctx, FBP_ERR_INVALID_ARGUMENT, "calibrate: 1 <= extremes <= %d", kMaxExtremes);
    FBP_CHECK(bind(ctx));

    Weights w{};
    FBP_CHECK(fbp_log_kernel_host(kernel_size, sigma, w.w));
    const uint32_t X = info->voxel_resolution.x, Y = info->voxel_resolution.y,
                   Z = info->voxel_resolution.z;
    const size_t n = size_t(X) * Y * Z;
    if (n < extremes || n >= (size_t(1) << 31))
        return fail(ctx, FBP_ERR_UNSUPPORTED, "calibrate: volume size out of range");
    const uint32_t k = extremes;

    // Scan-point chunks per candidate: enough blocks to fill the device, at
    // most kNeighMaxChunk scan points each.
    const uint32_t L = info->num_laser_targets;
    const uint32_t target = 8u * uint32_t(ctx->num_cus > 0 ? ctx->num_cus : 256);
    uint32_t chunks = std::max(1u, std::min(L, (target + 2 * k - 1) / (2 * k)));
    chunks = std::max(chunks, (L + kNeighMaxChunk - 1) / kNeighMaxChunk);
    const uint32_t chunk_len = (L + chunks - 1) / chunks;
    chunks = (L + chunk_len - 1) / chunk_len;

    // Volume segments of the scan pass: at most kMaxScanBlocks blocks.
    uint32_t nb = uint32_t(std::min<size_t>(kMaxScanBlocks, (n + kScanChunk - 1) / kScanChunk));
    const uint32_t seg = uint32_t((n + nb - 1) / nb);
    nb = uint32_t((n + seg - 1) / seg);

    const size_t bytes = align256(size_t(2) * nb * k * 8) + align256(size_t(nb) * 5 * 8) +
                         align256(size_t(2) * k * 4) + align256(5 * 8) +
                         align256(size_t(2) * k * chunks * 8) + 256;
    FBP_CHECK(ensure_scratch(ctx, bytes));
    char* cur = static_cast<char*>(ctx->scratch);
    auto take = [&](size_t b) {
        char* p = cur;
        cur += align256(b);
        return p;
    };
    uint64_t* cand = reinterpret_cast<uint64_t*>(take(size_t(2) * nb * k * 8));
    double* sums = reinterpret_cast<double*>(take(size_t(nb) * 5 * 8));
    uint32_t* sel = reinterpret_cast<uint32_t*>(take(size_t(2) * k * 4));
    double* fit = reinterpret_cast<double*>(take(5 * 8));
    double* partial = reinterpret_cast<double*>(take(size_t(2) * k * chunks * 8));

    hipLaunchKernelGGL(calib_scan_kernel, dim3(nb, 2), dim3(256), 0, s, volume, reference,
                       uint32_t(n), seg, k, cand, sums);
    hipLaunchKernelGGL(calib_select_kernel, dim3(2), dim3(256), 0, s, cand, sums, nb, seg, k, sel,
                       fit);
    const PointArgs a = point_args(*info, static_cast<const float*>(tr->data), laser_targets);
    // Pure chain: thresholds from the context's cached row, staged in LDS.
    const float* pthr = nullptr;
    const size_t thr_lds = (size_t(a.T) + 1) * sizeof(float);
    if (info->time_offset == 0.f && info->discard_first_last_bounces != 0 && thr_lds <= 49152)
        FBP_CHECK(pure_threshold_row(ctx, a.T, a.dt, s, &pthr));
    hipLaunchKernelGGL(calib_neigh_kernel, dim3(2 * k, chunks), dim3(256), pthr ? thr_lds : 0, s,
                       a, w, sel, k, chunk_len, pthr, partial);
    hipLaunchKernelGGL(calib_apply_kernel, dim3(unsigned((n + 255) / 256)), dim3(256), 0, s, volume,
                       n, fit, partial, k, chunks);
    FBP_CHECK_HIP(ctx, hipGetLastError());
    return FBP_OK;
}

}  // namespace fbp

extern "C" fbp_status fbp_learned_calibrate(fbp_context ctx, const fbp_reconstruction_info* info,
                                            const fbp_transient* tr,
                                            const fbp_vec3* laser_targets, float* volume,
                                            const float* reference, int32_t kernel_size,
                                            float sigma, uint32_t extremes, void* stream) {
    return fbp::learned_calibrate_impl(ctx, info, tr, laser_targets, volume, reference,
                                       kernel_size, sigma, extremes, as_stream(stream));
}
