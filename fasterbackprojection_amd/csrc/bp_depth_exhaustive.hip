// Exhaustive depth planes: the activation of pixel c (laser target c) at
// depth d sums, over every (laser l, sensor s) pair, the transient sample
// I[l][s][bin] with bin = (int)((|p - l| + |p - s| - t0 [+ path offsets]) / dt)
// at p = laser[c] + n * depth[d], for 0 <= bin < T and I > 1e-5 (no 1/r^2
// weight).
//
// Reference: backprojectExhaustive (FB/transient_reconstruction.cuh:66-91)
// launched by Laser::reconstructDepthExhaustive (FB/Laser.cu:431-480), one
// thread per (pixel, depth) summing the pairs ls = s * L + l in order with a
// gather per pair.
//
// Here a workgroup of 256 threads owns 2048 consecutive outputs o = c * D + d
// (eight per thread, 256 apart: eight independent chains per pair) and a
// contiguous chunk of pairs in the reference's order;
// the chunks are added in FP64 in a fixed order (deterministic).  For each
// pair, the bins that any point of the workgroup can reach lie in a window
// [kmin, kmin + W) found from the points' bounding box; the windows of a
// batch of pairs are staged through LDS (the next batch fetched into
// registers while the current one runs), so the per-pair gather of the
// reference becomes an LDS read.  Bin estimate in bin units as in the
// exhaustive voxel kernel: q = sqrt(|p - l|^2) / dt (one v_rsq) + |p - s| / dt
// (cached while s is unchanged, L consecutive pairs) + (offsets - t0) / dt;
// a point whose q lies within the estimate's error bound of a bin boundary
// adds nothing on the fast path and takes the reference's IEEE chain in a
// branch that the wave skips when no lane needs it.  Pairs whose points may
// come within millimetres of the laser or sensor point, or of zero distance,
// take the IEEE chain for every point.
#include "bp_common.h"

namespace fbp {

namespace {

constexpr int kDxThreads = 256;
constexpr int kDxOut = 8;                         // outputs per thread (independent chains)
constexpr int kDxBatch = 16;                      // pairs per staged batch (at most)
constexpr int kDxCap = 4096;                      // window floats per staging buffer
constexpr int kDxPrefetch = kDxCap / kDxThreads;  // registers per thread per batch

struct DepthExArgs {
    const float* data;  // I[l][s][t], F32
    const fbp_vec3* laser;
    const fbp_vec3* sensor;
    const double* depths;
    float* partial;  // [chunk][L * D]
    uint32_t D, L, S, T;
    float t0, dt, inv_dt;
    uint32_t discard;
    float n[3], lpos[3], spos[3];
    uint64_t pairs, chunk_len;
    uint32_t nblk;  // workgroups per chunk
};

// Parameters of one staged pair (computed by one thread, read by all).
struct DepthPair {
    float lx, ly, lz;
    int base;      // LDS index of bin k: base + k (buffer, batch slot and kmin folded in)
    float hme;     // 0.5 - error bound of the fast quotient (bins)
    float offq;    // (off - t0) / dt
    float off;     // |l - laser position| + |s - sensor position| (IEEE chain), or 0
    int near;      // 1: every point takes the IEEE chain
    uint32_t row;  // l * S + s
    int kmin;
};

// Positions of one pair, loaded a batch ahead of their use.
struct PairRaw {
    float l[3], s[3];
    uint32_t row;
};

__device__ __forceinline__ PairRaw load_pair(const DepthExArgs& a, uint32_t pair) {
    const uint32_t s = pair / a.L, l = pair % a.L;
    PairRaw r;
    r.l[0] = a.laser[l].x;
    r.l[1] = a.laser[l].y;
    r.l[2] = a.laser[l].z;
    r.s[0] = a.sensor[s].x;
    r.s[1] = a.sensor[s].y;
    r.s[2] = a.sensor[s].z;
    r.row = l * a.S + s;
    return r;
}

__device__ DepthPair depth_pair(const DepthExArgs& a, const PairRaw& r, int slot_base,
                                const float (&lo)[3], const float (&hi)[3]) {
    const float* lp = r.l;
    const float* sp = r.s;
    float off = 0.f;
    if (a.discard == 0)
        off = rn_add(distance3(lp[0], lp[1], lp[2], a.lpos[0], a.lpos[1], a.lpos[2]),
                     distance3(sp[0], sp[1], sp[2], a.spos[0], a.spos[1], a.spos[2]));
    const float rlmin = box_min_dist(lp, lo, hi), rlmax = box_max_dist(lp, lo, hi);
    const float rsmin = box_min_dist(sp, lo, hi), rsmax = box_max_dist(sp, lo, hi);
    const float dmin = rlmin + rsmin - a.t0 + off;
    const float dmax = rlmax + rsmax - a.t0 + off;
    DepthPair P;
    P.lx = lp[0];
    P.ly = lp[1];
    P.lz = lp[2];
    P.kmin = int(floorf(dmin * a.inv_dt)) - 3;
    P.base = slot_base - P.kmin;
    // Error bound of the fast quotient against fl(dist / dt), in bins: the
    // squared laser distance by FMAs (<= 5 ulp from the chain's), v_rsq
    // (1 ulp), the product and the cached sensor term stay within
    // 7.5 * 2^-24 * dist, the offset terms within 3 * 2^-24 * mag, and the
    // chain's own roundings within 4 * 2^-24 * mag: 10 * 2^-23 * mag / dt
    // covers them.
    const float mag = rlmax + rsmax + fabsf(a.t0) + fabsf(off);
    const float eps = 10.f * 1.1920929e-7f * mag * a.inv_dt + 1.0e-6f;
    P.hme = 0.5f - eps;
    P.offq = (off - a.t0) * a.inv_dt;
    P.off = off;
    P.near = ((dmin < 0.0045f && dmax > -0.0045f) || rlmin < 1.0e-3f || rsmin < 1.0e-3f) ? 1 : 0;
    P.row = r.row;
    return P;
}

// The reference's chain for point p and pair (l, s): ds = |p - s| (IEEE),
// off as in DepthPair.
__device__ __forceinline__ float exact_quotient(const DepthExArgs& a, float lx, float ly, float lz,
                                                float px, float py, float pz, float ds,
                                                float off) {
    float dist = rn_sub(rn_add(distance3(lx, ly, lz, px, py, pz), ds), a.t0);
    if (a.discard == 0) dist = rn_add(dist, off);
    return rn_div(dist, a.dt);
}

// Grid: chunks * nblk workgroups, chunk-major (the workgroups of one chunk
// read the same transient rows at about the same time).  256 threads, each
// with the kDxOut outputs o = blk * 256 * kDxOut + j * 256 + tid.
// COFF: t0 != 0 or path offsets (q needs the pair's (off - t0) / dt).
// Pipeline per batch b of pairs: fetch the windows of b + 1 (parameters
// already in LDS), run b, store b + 1's windows, compute the parameters of
// b + 2 from positions loaded during b - 1 and load those of b + 3; one
// barrier.
template <bool COFF>
__global__ __launch_bounds__(kDxThreads) void depth_exhaustive_window_kernel(DepthExArgs a) {
    __shared__ float win[2 * kDxCap];
    __shared__ DepthPair prm[3][kDxBatch];
    __shared__ float boxs[kDxThreads / 64][6];

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const uint32_t blk = blockIdx.x % a.nblk, chunk = blockIdx.x / a.nblk;
    const uint32_t p_begin = uint32_t(uint64_t(chunk) * a.chunk_len);
    const uint32_t p_end = uint32_t(min(a.pairs, uint64_t(p_begin) + a.chunk_len));
    const uint32_t n_out = a.L * a.D;

    float px[kDxOut], py[kDxOut], pz[kDxOut];
    float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
#pragma unroll
    for (int j = 0; j < kDxOut; ++j) {
        const uint32_t o = blk * (kDxThreads * kDxOut) + uint32_t(j * kDxThreads + tid);
        // Outputs past the end carry the last output's point (inside the box).
        const uint32_t oc = o < n_out ? o : n_out - 1;
        const uint32_t c = oc / a.D, d = oc % a.D;
        const float depth = static_cast<float>(a.depths[d]);
        // checkPos = laserTargets[c] + relayWallNormal * depth (componentwise).
        px[j] = rn_add(a.laser[c].x, rn_mul(a.n[0], depth));
        py[j] = rn_add(a.laser[c].y, rn_mul(a.n[1], depth));
        pz[j] = rn_add(a.laser[c].z, rn_mul(a.n[2], depth));
        lo[0] = fminf(lo[0], px[j]);
        lo[1] = fminf(lo[1], py[j]);
        lo[2] = fminf(lo[2], pz[j]);
        hi[0] = fmaxf(hi[0], px[j]);
        hi[1] = fmaxf(hi[1], py[j]);
        hi[2] = fmaxf(hi[2], pz[j]);
    }

    // Bounding box of the workgroup's points.
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1)
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            lo[k] = fminf(lo[k], __shfl_xor(lo[k], m));
            hi[k] = fmaxf(hi[k], __shfl_xor(hi[k], m));
        }
    if (lane == 0)
        for (int k = 0; k < 3; ++k) {
            boxs[wave][k] = lo[k];
            boxs[wave][3 + k] = hi[k];
        }
    __syncthreads();
#pragma unroll
    for (int w = 0; w < kDxThreads / 64; ++w)
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            lo[k] = fminf(lo[k], boxs[w][k]);
            hi[k] = fmaxf(hi[k], boxs[w][3 + k]);
        }
    // Over the box, each leg's distance varies by at most its diagonal.
    const float diag = sqrtf(sq(hi[0] - lo[0]) + sq(hi[1] - lo[1]) + sq(hi[2] - lo[2]));
    const float wf = 2.f * diag * a.inv_dt;

    uint32_t s = p_begin / a.L, l = p_begin % a.L;
    float ds[kDxOut];
#pragma unroll
    for (int j = 0; j < kDxOut; ++j)
        ds[j] = distance3(px[j], py[j], pz[j], a.sensor[s].x, a.sensor[s].y, a.sensor[s].z);

    if (!(wf < float(kDxCap / 2))) {
        // Points spread over metres (windows beyond the staging buffers):
        // the reference's per-pair gather.  Uniform over the workgroup; no
        // barrier follows.
        float sum[kDxOut] = {};
        for (uint32_t ls = p_begin; ls < p_end; ++ls) {
            const float sx = a.sensor[s].x, sy = a.sensor[s].y, sz = a.sensor[s].z;
            const float lx = a.laser[l].x, ly = a.laser[l].y, lz = a.laser[l].z;
            float off = 0.f;
            if (a.discard == 0)
                off = rn_add(distance3(lx, ly, lz, a.lpos[0], a.lpos[1], a.lpos[2]),
                             distance3(sx, sy, sz, a.spos[0], a.spos[1], a.spos[2]));
            const float* row = a.data + (size_t(l) * a.S + s) * a.T;
#pragma unroll
            for (int j = 0; j < kDxOut; ++j) {
                const float q = exact_quotient(a, lx, ly, lz, px[j], py[j], pz[j], ds[j], off);
                if (q > -1.0f && q < float(a.T)) {
                    const float v = row[int(q)];
                    if (v > kEps) sum[j] += v;
                }
            }
            if (++l == a.L && ls + 1 < p_end) {
                l = 0;
                ++s;
#pragma unroll
                for (int j = 0; j < kDxOut; ++j)
                    ds[j] = distance3(px[j], py[j], pz[j], a.sensor[s].x, a.sensor[s].y,
                                      a.sensor[s].z);
            }
        }
#pragma unroll
        for (int j = 0; j < kDxOut; ++j) {
            const uint32_t o = blk * (kDxThreads * kDxOut) + uint32_t(j * kDxThreads + tid);
            if (o < n_out) a.partial[size_t(chunk) * n_out + o] = sum[j];
        }
        return;
    }
    // Window: kmin = floor(dmin / dt) - 3 and (dmax - dmin) / dt <= wf.
    const int W = int(wf) + 8;
    const uint32_t B = uint32_t(min(kDxBatch, kDxCap / W));
    const float inv_W = 1.0f / float(W);
    const uint32_t nbatch = (p_end - p_begin + B - 1) / B;
    auto batch_len = [&](uint32_t b) { return min(B, p_end - p_begin - b * B); };

    float reg[kDxPrefetch];
    // Every register's load is issued (unused ones read element 0), so the
    // loads of a batch are in flight together and waited for once, at the
    // store.
    auto fetch = [&](uint32_t nb, const DepthPair* pr) {
#pragma unroll
        for (int i = 0; i < kDxPrefetch; ++i) {
            const int f = tid + i * kDxThreads;
            const bool in = f < int(nb) * W;
            // f / W through the reciprocal: exact for f < 2^22.
            const int jl = in ? int((float(f) + 0.5f) * inv_W) : 0;
            const int w = f - jl * W;
            const int k = pr[jl].kmin + w;
            const bool ok = in && k >= 0 && k < int(a.T);
            const float v = a.data[ok ? size_t(pr[jl].row) * a.T + uint32_t(k) : 0];
            reg[i] = (ok && v > kEps) ? v : 0.f;
        }
    };
    auto store = [&](int buf, uint32_t nb) {
#pragma unroll
        for (int i = 0; i < kDxPrefetch; ++i) {
            const int f = tid + i * kDxThreads;
            if (f < int(nb) * W) win[buf * kDxCap + f] = reg[i];
        }
    };

    // Prologue: parameters of batches 0 and 1, positions of batch 2,
    // windows of batch 0.
    PairRaw raw{};
    for (uint32_t b = 0; b < 2 && b < nbatch; ++b)
        if (uint32_t(tid) < batch_len(b))
            prm[b][tid] = depth_pair(a, load_pair(a, p_begin + b * B + tid),
                                     int(b & 1) * kDxCap + tid * W, lo, hi);
    if (2 < nbatch && uint32_t(tid) < batch_len(2)) raw = load_pair(a, p_begin + 2 * B + tid);
    __syncthreads();
    fetch(batch_len(0), prm[0]);
    store(0, batch_len(0));
    __syncthreads();

    float sum[kDxOut] = {}, corr[kDxOut] = {}, dsq[kDxOut];
#pragma unroll
    for (int j = 0; j < kDxOut; ++j) dsq[j] = ds[j] * a.inv_dt;
    for (uint32_t b = 0; b < nbatch; ++b) {
        const int cur = int(b & 1);
        const uint32_t nb = batch_len(b);
        const bool has_next = b + 1 < nbatch;
        if (has_next) fetch(batch_len(b + 1), prm[(b + 1) % 3]);
        const DepthPair* pr = prm[b % 3];
        for (uint32_t jl = 0; jl < nb; ++jl) {
            const DepthPair& P = pr[jl];
            const float lx = P.lx, ly = P.ly, lz = P.lz;
            const int base = P.base;
            if (P.near) {
#pragma unroll
                for (int j = 0; j < kDxOut; ++j) {
                    const float q = exact_quotient(a, lx, ly, lz, px[j], py[j], pz[j], ds[j], P.off);
                    if (q > -1.0f && q < float(a.T)) sum[j] += win[base + int(q)];
                }
            } else {
                const float hme = P.hme;
                float offq = 0.f;
                if constexpr (COFF) offq = P.offq;
#pragma unroll
                for (int j = 0; j < kDxOut; ++j) {
                    const float dx = px[j] - lx, dy = py[j] - ly, dz = pz[j] - lz;
                    const float sl = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
                    float q = fmaf(sl * __builtin_amdgcn_rsqf(sl), a.inv_dt, dsq[j]);
                    if constexpr (COFF) q = q + offq;
                    const int k = int(q);
                    const float f = __builtin_amdgcn_fractf(q);
                    const bool band = fabsf(f - 0.5f) > hme;
                    const float v = win[base + k];
                    sum[j] += band ? 0.f : v;
                    if (band) {
                        // q within the error bound of a bin boundary.
                        const float qe =
                            exact_quotient(a, lx, ly, lz, px[j], py[j], pz[j], ds[j], P.off);
                        if (qe > -1.0f && qe < float(a.T)) corr[j] += win[base + int(qe)];
                    }
                }
            }
            // Next pair: l fastest (uniform).
            if (++l == a.L) {
                l = 0;
                if (p_begin + b * B + jl + 1 < p_end) {
                    ++s;
#pragma unroll
                    for (int j = 0; j < kDxOut; ++j) {
                        ds[j] = distance3(px[j], py[j], pz[j], a.sensor[s].x, a.sensor[s].y,
                                          a.sensor[s].z);
                        dsq[j] = ds[j] * a.inv_dt;
                    }
                }
            }
        }
        if (has_next) store(cur ^ 1, batch_len(b + 1));
        if (b + 2 < nbatch) {
            if (uint32_t(tid) < batch_len(b + 2))
                prm[(b + 2) % 3][tid] = depth_pair(a, raw, cur * kDxCap + tid * W, lo, hi);
            if (b + 3 < nbatch && uint32_t(tid) < batch_len(b + 3))
                raw = load_pair(a, p_begin + (b + 3) * B + tid);
        }
        __syncthreads();
    }
#pragma unroll
    for (int j = 0; j < kDxOut; ++j) {
        const uint32_t o = blk * (kDxThreads * kDxOut) + uint32_t(j * kDxThreads + tid);
        if (o < n_out) a.partial[size_t(chunk) * n_out + o] = sum[j] + corr[j];
    }
}

__global__ void depth_reduce_kernel(const float* __restrict__ partial, uint32_t chunks, size_t n,
                                    float* __restrict__ act) {
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double s = 0.0;
    for (uint32_t c = 0; c < chunks; ++c) s += double(partial[size_t(c) * n + i]);
    act[i] = float(s);
}

}  // namespace

}  // namespace fbp

using namespace fbp;

extern "C" fbp_status fbp_backproject_depth_exhaustive(
    fbp_context ctx, const fbp_reconstruction_info* info, const fbp_transient* tr,
    const fbp_vec3* laser_targets, const fbp_vec3* sensor_targets, const double* depths,
    uint32_t num_depths, float* activation, void* stream) {
    if (!ctx || !info || !tr || !tr->data || !laser_targets || !sensor_targets || !depths ||
        !activation)
        return fail(ctx, FBP_ERR_INVALID_ARGUMENT, "depth exhaustive: null argument");
    if (tr->storage != FBP_STORAGE_F32)
        return fail(ctx, FBP_ERR_UNSUPPORTED, "depth exhaustive: F32 transient only");
    if (!num_depths || !info->num_laser_targets || !info->num_sensor_targets ||
        !info->num_time_bins || !(info->time_step > 0.f))
        return fail(ctx, FBP_ERR_INVALID_ARGUMENT, "depth exhaustive: empty grid");
    const uint32_t L = info->num_laser_targets, S = info->num_sensor_targets;
    const size_t n = size_t(L) * num_depths;
    if (n > 0xffffffffu || uint64_t(L) * S > 0xffffffffu)
        return fail(ctx, FBP_ERR_UNSUPPORTED, "depth exhaustive: more than 2^32 outputs or pairs");
    FBP_CHECK(bind(ctx));
    hipStream_t s = as_stream(stream);

    DepthExArgs a{};
    a.data = static_cast<const float*>(tr->data);
    a.laser = laser_targets;
    a.sensor = sensor_targets;
    a.depths = depths;
    a.D = num_depths;
    a.L = L;
    a.S = S;
    a.T = info->num_time_bins;
    a.t0 = info->time_offset;
    a.dt = info->time_step;
    a.inv_dt = 1.0f / info->time_step;
    a.discard = info->discard_first_last_bounces;
    a.n[0] = info->relay_wall_normal.x;
    a.n[1] = info->relay_wall_normal.y;
    a.n[2] = info->relay_wall_normal.z;
    a.lpos[0] = info->laser_position.x;
    a.lpos[1] = info->laser_position.y;
    a.lpos[2] = info->laser_position.z;
    a.spos[0] = info->sensor_position.x;
    a.spos[1] = info->sensor_position.y;
    a.spos[2] = info->sensor_position.z;
    a.pairs = uint64_t(L) * S;
    a.nblk = uint32_t((n + kDxThreads * kDxOut - 1) / (kDxThreads * kDxOut));
    // About 12 workgroups per CU over the launch (3 resident at 158 VGPRs),
    // chunks of at least 256 pairs.
    const uint64_t want = 12ull * uint64_t(ctx->num_cus > 0 ? ctx->num_cus : 256);
    uint64_t chunks = std::max<uint64_t>(1, std::min<uint64_t>((want + a.nblk - 1) / a.nblk,
                                                               (a.pairs + 255) / 256));
    a.chunk_len = (a.pairs + chunks - 1) / chunks;
    chunks = (a.pairs + a.chunk_len - 1) / a.chunk_len;
    if (chunks * a.nblk > 0x7fffffffull)
        return fail(ctx, FBP_ERR_UNSUPPORTED, "depth exhaustive: grid too large");
    FBP_CHECK(ensure_scratch(ctx, size_t(chunks) * n * sizeof(float)));
    a.partial = static_cast<float*>(ctx->scratch);
    const bool coff = !(a.t0 == 0.f && a.discard != 0);
    const dim3 grid(unsigned(chunks * a.nblk));
    if (coff)
        hipLaunchKernelGGL(depth_exhaustive_window_kernel<true>, grid, dim3(kDxThreads), 0, s, a);
    else
        hipLaunchKernelGGL(depth_exhaustive_window_kernel<false>, grid, dim3(kDxThreads), 0, s, a);
    FBP_CHECK_HIP(ctx, hipGetLastError());
    hipLaunchKernelGGL(depth_reduce_kernel, dim3(unsigned((n + 255) / 256)), dim3(256), 0, s,
                       a.partial, uint32_t(chunks), n, activation);
    FBP_CHECK_HIP(ctx, hipGetLastError());
    return FBP_OK;
}
