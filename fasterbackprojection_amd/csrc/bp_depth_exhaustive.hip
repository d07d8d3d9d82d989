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
#include "group_common.h"

namespace fbp {

namespace {

constexpr int kGVecD = 2;   // 4-bin loads per lane per batch in the group form
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

// ---------------------------------------------------------------------------
// Laser-group form of the exhaustive depth planes (the laser-group voxel
// kernel's scheme on the depth-plane points): for a wall normal along y the
// point p = laser[c] + n * depth[d] keeps the pixel's x and z, so a lane holds
// 4 pixels x 4 depths (4 x and z values, 16 y values; the 64 lanes of a wave
// take 64 consecutive depths of one pixel per slot).  A workgroup of eight
// waves owns a tile of 32 pixels (consecutive in a Morton order of the
// pixels on the wall, so a tile is a compact block) x 256 depth slots and a
// range of (group of 8 lasers, sensor) batches.  Each lane keeps the eight exact laser
// distances of its 16 points while the group is unchanged and computes its
// 16 exact sensor distances once per sensor; the update is the reference's
// dist = fl(rl + rs) (- t0, + off), the two-term quotient (exact at every
// bin boundary by the host's check) and one LDS read of the staged sample:
// 6 VALU.  Pairs whose points may come within a millimetre of the laser or
// sensor point, or of zero distance, take the reference's chain per point;
// a tile whose bin window would exceed the staging capacity reads its samples
// from the transient directly.
constexpr int kDgPix = 32;                 // pixels per tile
constexpr int kDgDep = 256;                // depth slots per tile (16 per lane)
constexpr int kDgWin = 4 * 64 * kGVecD;    // staged window capacity (bins) per pair
// Pair parameters are computed kDgPrmBatches batches at a time, by the first
// kDgPrmBatches * kGroup threads, into one half of a two-block LDS ring.
constexpr int kDgPrmBatches = 16;

struct DepthGroupArgs {
    const float* data;  // I[l][s][t], F32
    const fbp_vec3* laser;
    const fbp_vec3* sensor;
    const double* depths;
    const uint32_t* perm;  // pixels in Morton order on the wall
    float* partial;        // [chunk][L * D]
    uint32_t L, S, T, D;
    float t0, dt, inv_dt, inv_dt_lo, ny;  // inv_dt + inv_dt_lo = 1 / dt
    uint32_t discard;
    float lpos[3], spos[3];
    uint32_t groups, groups_per_chunk, sensor_parts, sensors_per_part;
    uint32_t ptiles, num_tiles, num_blocks;
};

struct __align__(16) DgPair {
    int base;      // g * kDgWin - kmin4
    int kmin;
    int kmin4;
    int fast;      // 1: fast body, 0: the reference's chain per point, -1: no pair
    float lx, ly, lz, off;
    float sx, sy, sz, pad;
};

struct __align__(16) DgLaser {
    float x, y, z, offl;
    float rmin, rmax, pad[2];
};

// Pixels ordered along a Morton curve of their (x, z) position on the wall,
// quantised to 16 bits over the pixels' bounding box; ties by index.  One
// workgroup, L <= 4096.
__global__ __launch_bounds__(1024) void depth_pixel_order_kernel(const fbp_vec3* laser, uint32_t L,
                                                                 uint32_t* perm) {
    __shared__ uint64_t key[4096];
    __shared__ float red[4][16];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    float b[4] = {INFINITY, -INFINITY, INFINITY, -INFINITY};
    for (uint32_t i = tid; i < L; i += 1024) {
        b[0] = fminf(b[0], laser[i].x);
        b[1] = fmaxf(b[1], laser[i].x);
        b[2] = fminf(b[2], laser[i].z);
        b[3] = fmaxf(b[3], laser[i].z);
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        b[0] = fminf(b[0], __shfl_xor(b[0], o));
        b[1] = fmaxf(b[1], __shfl_xor(b[1], o));
        b[2] = fminf(b[2], __shfl_xor(b[2], o));
        b[3] = fmaxf(b[3], __shfl_xor(b[3], o));
    }
    if (lane == 0)
        for (int k = 0; k < 4; ++k) red[k][wave] = b[k];
    __syncthreads();
    for (int w = 0; w < 16; ++w) {
        b[0] = fminf(b[0], red[0][w]);
        b[1] = fmaxf(b[1], red[1][w]);
        b[2] = fminf(b[2], red[2][w]);
        b[3] = fmaxf(b[3], red[3][w]);
    }
    const float sx = b[1] > b[0] ? 65535.f / (b[1] - b[0]) : 0.f;
    const float sz = b[3] > b[2] ? 65535.f / (b[3] - b[2]) : 0.f;
    uint32_t N = 1;
    while (N < L) N <<= 1;
    for (uint32_t i = tid; i < N; i += 1024) {
        uint64_t k = ~0ull;
        if (i < L) {
            const uint32_t qx = uint32_t(fminf(fmaxf((laser[i].x - b[0]) * sx, 0.f), 65535.f));
            const uint32_t qz = uint32_t(fminf(fmaxf((laser[i].z - b[2]) * sz, 0.f), 65535.f));
            uint32_t code = 0;
            for (int bit = 0; bit < 16; ++bit)
                code |= (((qx >> bit) & 1u) << (2 * bit)) | (((qz >> bit) & 1u) << (2 * bit + 1));
            k = (uint64_t(code) << 32) | i;
        }
        key[i] = k;
    }
    __syncthreads();
    // Bitonic sort of key[0 .. N).
    for (uint32_t size = 2; size <= N; size <<= 1) {
        for (uint32_t stride = size >> 1; stride > 0; stride >>= 1) {
            for (uint32_t i = tid; i < N; i += 1024) {
                const uint32_t j = i ^ stride;
                if (j > i) {
                    const bool up = (i & size) == 0;
                    const uint64_t x = key[i], y = key[j];
                    if ((x > y) == up) {
                        key[i] = y;
                        key[j] = x;
                    }
                }
            }
            __syncthreads();
        }
    }
    for (uint32_t i = tid; i < L; i += 1024) perm[i] = uint32_t(key[i]);
}

template <bool COFF>
__global__ __launch_bounds__(512, 1) void depth_group_kernel(DepthGroupArgs a) {
    extern __shared__ float smem[];
    // [2][kGroup][kDgWin] staged samples, [2 * kDgPrmBatches][kGroup] pair
    // parameters and per-batch masks, the chunk's lasers and sensors, box
    // reduction scratch.
    float* winbuf = smem;
    DgPair* prmbuf = reinterpret_cast<DgPair*>(winbuf + 2 * kGroup * kDgWin);
    // Per batch slot: bits 0-7 pairs on the fast body, bits 8-15 pairs on
    // the reference's chain.
    uint32_t* bmask = reinterpret_cast<uint32_t*>(prmbuf + 2 * kDgPrmBatches * kGroup);
    DgLaser* lasl = reinterpret_cast<DgLaser*>(bmask + 2 * kDgPrmBatches);
    float4* senl = reinterpret_cast<float4*>(lasl + a.groups_per_chunk * kGroup);
    float* red = reinterpret_cast<float*>(senl + a.sensors_per_part);  // [6][8]

    const int tid = threadIdx.x;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int lane = tid & 63;
    const uint32_t nb8 = a.num_blocks / 8u;
    const uint32_t bid = blockIdx.x < nb8 * 8u ? (blockIdx.x % 8u) * nb8 + blockIdx.x / 8u
                                               : blockIdx.x;
    const uint32_t tile = bid % a.num_tiles, chunk = bid / a.num_tiles;
    const uint32_t tp = tile % a.ptiles, td = tile / a.ptiles;

    // Wave w holds pixels 4w .. 4w + 3 of the tile; the lane's slot v is
    // pixel j = v / 4 at depth dl + 64 (v % 4), so the 64 lanes of a wave
    // take 64 consecutive depths of one pixel at a time (neighbouring bins:
    // window reads with few bank conflicts).  Pixels past the last one
    // repeat the tile's first pixel (inside the tile's box), depths past the
    // last one repeat the last depth; neither is written out.
    const uint32_t dl = td * kDgDep + uint32_t(lane);
    uint32_t cj[4];
    float px[4], pz[4], py[16];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const uint32_t pi = tp * kDgPix + 4u * uint32_t(wave) + uint32_t(j);
        cj[j] = a.perm[pi < a.L ? pi : tp * kDgPix];
        px[j] = a.laser[cj[j]].x;
        pz[j] = a.laser[cj[j]].z;
        const float lcy = a.laser[cj[j]].y;
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const uint32_t d = min(dl + 64u * uint32_t(u), a.D - 1);
            py[4 * j + u] = lcy + a.ny * static_cast<float>(a.depths[d]);
        }
    }
    const uint32_t pix0 = tp * kDgPix + 4u * uint32_t(wave);
    // Box of the tile's points.
    float b[6] = {INFINITY, -INFINITY, INFINITY, -INFINITY, INFINITY, -INFINITY};
#pragma unroll
    for (int v = 0; v < 16; ++v)
        if (pix0 + uint32_t(v / 4) < a.L && dl + 64u * uint32_t(v % 4) < a.D) {
            b[0] = fminf(b[0], px[v / 4]);
            b[1] = fmaxf(b[1], px[v / 4]);
            b[2] = fminf(b[2], py[v]);
            b[3] = fmaxf(b[3], py[v]);
            b[4] = fminf(b[4], pz[v / 4]);
            b[5] = fmaxf(b[5], pz[v / 4]);
        }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1)
#pragma unroll
        for (int k = 0; k < 6; k += 2) {
            b[k] = fminf(b[k], __shfl_xor(b[k], o));
            b[k + 1] = fmaxf(b[k + 1], __shfl_xor(b[k + 1], o));
        }
    if (lane == 0)
        for (int k = 0; k < 6; ++k) red[k * 8 + wave] = b[k];
    __syncthreads();
    float lo[3], hi[3];
#pragma unroll
    for (int ax = 0; ax < 3; ++ax) {
        lo[ax] = INFINITY;
        hi[ax] = -INFINITY;
        for (int w = 0; w < 8; ++w) {
            lo[ax] = fminf(lo[ax], red[(2 * ax) * 8 + w]);
            hi[ax] = fmaxf(hi[ax], red[(2 * ax + 1) * 8 + w]);
        }
    }
    const float ex = hi[0] - lo[0], ey = hi[1] - lo[1], ez = hi[2] - lo[2];
    // Bins the tile can reach for one pair: both legs vary by at most the box
    // diagonal; + 3 below kmin, aligned start and margin.
    const float wt = ceilf(2.f * sqrtf(ex * ex + ey * ey + ez * ez) * a.inv_dt) + 16.f;
    const bool fits = wt <= float(kDgWin);

    const uint32_t lgc = chunk / a.sensor_parts, sp = chunk % a.sensor_parts;
    const uint32_t g_begin = lgc * a.groups_per_chunk;
    const uint32_t g_end = min(a.groups, g_begin + a.groups_per_chunk);
    const uint32_t s_begin = sp * a.sensors_per_part;
    const uint32_t s_end = min(a.S, s_begin + a.sensors_per_part);
    const uint32_t ns = s_end > s_begin ? s_end - s_begin : 0u;
    const uint32_t nbatch = (g_end > g_begin ? g_end - g_begin : 0u) * ns;
    const float cq = a.inv_dt, cql = a.inv_dt_lo;

    const uint32_t l_begin = g_begin * kGroup;
    const uint32_t nl = min(a.L, g_end * kGroup) - min(a.L, l_begin);
    for (uint32_t i = tid; i < nl; i += 512) {
        const uint32_t l = l_begin + i;
        DgLaser G;
        const float lp3[3] = {a.laser[l].x, a.laser[l].y, a.laser[l].z};
        G.x = lp3[0];
        G.y = lp3[1];
        G.z = lp3[2];
        G.offl = a.discard == 0
                     ? distance3(lp3[0], lp3[1], lp3[2], a.lpos[0], a.lpos[1], a.lpos[2])
                     : 0.f;
        G.rmin = box_dist_fast(lp3, lo, hi, false);
        G.rmax = box_dist_fast(lp3, lo, hi, true);
        G.pad[0] = G.pad[1] = 0.f;
        lasl[i] = G;
    }
    for (uint32_t i = tid; i < ns; i += 512) {
        const uint32_t s = s_begin + i;
        const float x = a.sensor[s].x, y = a.sensor[s].y, z = a.sensor[s].z;
        const float offs =
            a.discard == 0 ? distance3(x, y, z, a.spos[0], a.spos[1], a.spos[2]) : 0.f;
        senl[i] = make_float4(x, y, z, offs);
    }
    __syncthreads();

    auto params = [&](uint32_t b, int g, DgPair* out) {
        DgPair P;
        const uint32_t li = (b / ns) * kGroup + uint32_t(g);
        const float4 S4 = senl[b % ns];
        const float sp3[3] = {S4.x, S4.y, S4.z};
        P.sx = sp3[0];
        P.sy = sp3[1];
        P.sz = sp3[2];
        P.pad = 0.f;
        const float offs = S4.w;
        const float rsmin = box_dist_fast(sp3, lo, hi, false);
        const float rsmax = box_dist_fast(sp3, lo, hi, true);
        if (l_begin + li >= a.L) {
            P.fast = -1;
            P.base = P.kmin = P.kmin4 = 0;
            P.lx = P.ly = P.lz = P.off = 0.f;
            *out = P;
            return;
        }
        const DgLaser G = lasl[li];
        P.lx = G.x;
        P.ly = G.y;
        P.lz = G.z;
        P.off = a.discard == 0 ? rn_add(G.offl, offs) : 0.f;
        const float dmin = G.rmin + rsmin - a.t0 + P.off;
        const float dmax = G.rmax + rsmax - a.t0 + P.off;
        P.kmin = int(floorf(dmin * a.inv_dt)) - 3;
        P.kmin4 = P.kmin & ~3;
        P.base = g * kDgWin - P.kmin4;
        const bool near = dmin < 0.0045f || G.rmin < 1.0e-3f || rsmin < 1.0e-3f;
        (void)dmax;
        P.fast = (fits && !near) ? 1 : 0;
        *out = P;
    };
    // Parameters of batch bn's pair g for the threads of a block (lanes
    // 8j .. 8j + 7 of a wave hold one batch; every lane of the wave takes
    // part, with valid = false past the last batch) and the batch's masks.
    auto params_block = [&](uint32_t bn, int g, bool valid) {
        int fast = -1;
        if (valid) {
            DgPair* o = prmbuf + (bn % (2u * kDgPrmBatches)) * kGroup + g;
            params(bn, g, o);
            fast = o->fast;
        }
        const uint64_t fm = __ballot(fast == 1), em = __ballot(fast == 0);
        const int sh = lane & ~(kGroup - 1);
        if (valid && g == 0)
            bmask[bn % (2u * kDgPrmBatches)] =
                uint32_t((fm >> sh) & 0xffu) | (uint32_t((em >> sh) & 0xffu) << 8);
    };
    // Raw samples of batch (gi, si): wave g fetches pair g's bins kmin4 + 4 j
    // .. + 3 from the pair's row (a scalar address); lanes whose bins lie
    // outside [0, T) read bins inside the row (the store zeroes them), and a
    // slot without a staged pair reads row 0.
    auto fetch = [&](uint32_t gi, uint32_t si, const DgPair* prm, float4 (&reg)[kGVecD]) {
        const int fast = rfl(prm[wave].fast);
        const int kmin4 = rfl(prm[wave].kmin4);
        const uint32_t grp = g_begin + gi;
        const uint32_t s = s_begin + si;
        const uint64_t pair =
            (fits && fast >= 0) ? uint64_t(grp * kGroup + uint32_t(wave)) * a.S + s : 0u;
        const float* row = a.data + pair * a.T;
#pragma unroll
        for (int i = 0; i < kGVecD; ++i) {
            const int k0 = kmin4 + 4 * (lane + 64 * i);
            const uint32_t kc = uint32_t(min(max(k0, 0), int(a.T) - 4));
            reg[i] = *reinterpret_cast<const float4*>(row + kc);
        }
    };
    auto store = [&](float* win, const DgPair* prm, const float4 (&reg)[kGVecD]) {
        const int fast = rfl(prm[wave].fast);
        if (!fits || fast < 0) return;
        const int kmin4 = rfl(prm[wave].kmin4);
#pragma unroll
        for (int i = 0; i < kGVecD; ++i) {
            const int j = lane + 64 * i;
            const int k0 = kmin4 + 4 * j;
            const bool ok = k0 >= 0 && k0 < int(a.T);
            const float4 r = reg[i];
            const float4 e = make_float4(ok && r.x > kEps ? r.x : 0.f, ok && r.y > kEps ? r.y : 0.f,
                                         ok && r.z > kEps ? r.z : 0.f, ok && r.w > kEps ? r.w : 0.f);
            *reinterpret_cast<float4*>(win + wave * kDgWin + 4 * j) = e;
        }
    };

    float acc[16];
#pragma unroll
    for (int v = 0; v < 16; ++v) acc[v] = 0.f;
    float rl[kGroup][16];
#pragma unroll
    for (int g = 0; g < kGroup; ++g)
#pragma unroll
        for (int v = 0; v < 16; ++v) rl[g][v] = 0.f;

    const int per = kGroup * kDgWin;
    float4 reg[kGVecD];
    if (nbatch > 0) {
        // Parameters of the first two blocks of batches.
        if (tid < 2 * kDgPrmBatches * kGroup)
            params_block(uint32_t(tid / kGroup), tid % kGroup, uint32_t(tid / kGroup) < nbatch);
        __syncthreads();
        fetch(0, 0, prmbuf, reg);
        store(winbuf, prmbuf, reg);
        __syncthreads();
    }
    // Batch b = gi * ns + si (laser group gi of the chunk, sensor si of the
    // part), kept as running counters.
    uint32_t gi = 0, si = 0;
    for (uint32_t b = 0; b < nbatch; ++b) {
        const int cur = int(b & 1u);
        const DgPair* prm = prmbuf + (b % (2u * kDgPrmBatches)) * kGroup;
        const bool has_next = b + 1 < nbatch;
        const DgPair* prm_next = prmbuf + ((b + 1) % (2u * kDgPrmBatches)) * kGroup;
        uint32_t gn = gi, sn = si + 1;
        if (sn == ns) {
            sn = 0;
            ++gn;
        }
        if (has_next) fetch(gn, sn, prm_next, reg);
        if (si == 0) {
#pragma unroll
            for (int g = 0; g < kGroup; ++g) {
                const float lxg = rfl(prm[g].lx), lyg = rfl(prm[g].ly), lzg = rfl(prm[g].lz);
                float XL2[4], ZL2[4];
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    XL2[j] = sq(rn_sub(px[j], lxg));
                    ZL2[j] = sq(rn_sub(pz[j], lzg));
                }
#pragma unroll
                for (int v = 0; v < 16; ++v)
                    rl[g][v] = rn_sqrt(dot3(XL2[v / 4], sq(rn_sub(py[v], lyg)), ZL2[v / 4]));
            }
        }
        float rs[16];
        const float sxp = rfl(prm[0].sx), syp = rfl(prm[0].sy), szp = rfl(prm[0].sz);
        {
            float XS2[4], ZS2[4];
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                XS2[j] = sq(rn_sub(px[j], sxp));
                ZS2[j] = sq(rn_sub(pz[j], szp));
            }
#pragma unroll
            for (int v = 0; v < 16; ++v)
                rs[v] = sqrt_rn_normal(dot3(XS2[v / 4], sq(rn_sub(py[v], syp)), ZS2[v / 4]));
        }
        const uint32_t masks = uint32_t(rfl(int(bmask[b % (2u * kDgPrmBatches)])));
        const int fast_pairs = int(masks & 0xffu);
        int exact_pairs = int(masks >> 8);
        auto fast_body = [&](int g, int base, float off) {
#pragma unroll
            for (int h = 0; h < 16; h += 8) {
                float e[8];
#pragma unroll
                for (int v = 0; v < 8; ++v) {
                    float d = rl[g][h + v] + rs[h + v];
                    if constexpr (COFF) d = (d - a.t0) + off;
                    e[v] = winbuf[base + static_cast<int>(group_quotient(d, cq, cql))];
                }
                __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                for (int v = 0; v < 8; ++v) acc[h + v] += e[v];
            }
        };
        if (fast_pairs == (1 << kGroup) - 1) {
            int bases[kGroup];
            float offs[kGroup];
#pragma unroll
            for (int g = 0; g < kGroup; ++g) {
                bases[g] = rfl(prm[g].base) + cur * per;
                offs[g] = rfl(prm[g].off);
            }
#pragma unroll
            for (int g = 0; g < kGroup; ++g) fast_body(g, bases[g], offs[g]);
        } else {
#pragma unroll
            for (int g = 0; g < kGroup; ++g)
                if (fast_pairs & (1 << g)) fast_body(g, rfl(prm[g].base) + cur * per, rfl(prm[g].off));
        }
        // The reference's chain per point (samples from the window when the
        // tile's window fits, else from the transient).
        while (exact_pairs) {
            const int g = __builtin_ctz(exact_pairs);
            exact_pairs &= exact_pairs - 1;
            const DgPair& P = prm[g];
            const float lx = rfl(P.lx), ly = rfl(P.ly), lz = rfl(P.lz), off = rfl(P.off);
            const int base = rfl(P.base) + cur * per;
            const uint64_t pair = uint64_t((g_begin + gi) * kGroup + uint32_t(g)) * a.S +
                                  (s_begin + si);
#pragma unroll
            for (int v = 0; v < 16; ++v) {
                const float qx = px[v / 4], qz = pz[v / 4];
                const float sl = dot3(sq(rn_sub(qx, lx)), sq(rn_sub(py[v], ly)), sq(rn_sub(qz, lz)));
                const float ss =
                    dot3(sq(rn_sub(qx, sxp)), sq(rn_sub(py[v], syp)), sq(rn_sub(qz, szp)));
                const float de = exhaustive_dist(sl, ss, a.t0, a.discard, off);
                const int k = bin_of(de, a.dt);
                if (k >= 0 && k < int(a.T)) {
                    float I;
                    if (fits) {
                        I = winbuf[base + k];
                    } else {
                        I = a.data[pair * a.T + uint32_t(k)];
                        I = I > kEps ? I : 0.f;
                    }
                    acc[v] += I;
                }
            }
        }
        if (has_next) store(winbuf + (cur ^ 1) * per, prm_next, reg);
        // First batch of a block: the block after next goes to the slots of
        // the block before, whose last batch ended with the previous barrier
        // (and is first needed by the fetch at the end of the next block).
        if (b > 0 && b % kDgPrmBatches == 0 && tid < kDgPrmBatches * kGroup) {
            const uint32_t bn = b + kDgPrmBatches + uint32_t(tid / kGroup);
            params_block(bn, tid % kGroup, bn < nbatch);
        }
        __syncthreads();
        gi = gn;
        si = sn;
    }
    float* out = a.partial + size_t(chunk) * a.L * a.D;
#pragma unroll
    for (int v = 0; v < 16; ++v) {
        const uint32_t d = dl + 64u * uint32_t(v % 4);
        if (pix0 + uint32_t(v / 4) < a.L && d < a.D) out[size_t(cj[v / 4]) * a.D + d] = acc[v];
    }
}

// Laser-group form when the capture qualifies (wall normal along y, at
// least 8 lasers and 64 sensors, at most 4096 pixels, T a multiple of 4,
// 16-byte aligned samples, the bin quotient check passed); *used = false
// (nothing launched) otherwise.
fbp_status depth_group_launch(fbp_context ctx, const fbp_reconstruction_info* info,
                              const fbp_transient* tr, const fbp_vec3* laser_targets,
                              const fbp_vec3* sensor_targets, const double* depths,
                              uint32_t num_depths, float* activation, hipStream_t s, bool* used) {
    *used = false;
    DepthGroupArgs a{};
    a.L = info->num_laser_targets;
    a.S = info->num_sensor_targets;
    a.T = info->num_time_bins;
    a.D = num_depths;
    if (info->relay_wall_normal.x != 0.f || info->relay_wall_normal.z != 0.f || a.L < uint32_t(kGroup) ||
        a.L > 4096u || a.S < 64u || a.T < 4u || (a.T & 3u) != 0u ||
        (reinterpret_cast<uintptr_t>(tr->data) % 16u) != 0u)
        return FBP_OK;
    if (!group_quotient_cached(ctx, info->time_step, a.T)) return FBP_OK;
    a.data = static_cast<const float*>(tr->data);
    a.laser = laser_targets;
    a.sensor = sensor_targets;
    a.depths = depths;
    a.t0 = info->time_offset;
    a.dt = info->time_step;
    a.inv_dt = 1.0f / info->time_step;
    a.inv_dt_lo = group_inv_dt_lo(info->time_step);
    a.ny = info->relay_wall_normal.y;
    a.discard = info->discard_first_last_bounces;
    a.lpos[0] = info->laser_position.x;
    a.lpos[1] = info->laser_position.y;
    a.lpos[2] = info->laser_position.z;
    a.spos[0] = info->sensor_position.x;
    a.spos[1] = info->sensor_position.y;
    a.spos[2] = info->sensor_position.z;
    a.groups = (a.L + kGroup - 1) / kGroup;
    a.ptiles = (a.L + kDgPix - 1) / kDgPix;
    a.num_tiles = a.ptiles * ((a.D + kDgDep - 1) / kDgDep);
    const uint64_t cus = uint64_t(ctx->num_cus > 0 ? ctx->num_cus : 256);
    const uint64_t want = std::max<uint64_t>(1, (4 * cus + a.num_tiles - 1) / a.num_tiles);
    uint64_t lgc = std::min<uint64_t>(a.groups, want);
    a.groups_per_chunk = uint32_t((a.groups + lgc - 1) / lgc);
    lgc = (a.groups + a.groups_per_chunk - 1) / a.groups_per_chunk;
    uint64_t parts = std::max<uint64_t>(1, want / lgc);
    parts = std::min<uint64_t>(parts, std::max<uint64_t>(1, a.S / 64));
    a.sensors_per_part = uint32_t((a.S + parts - 1) / parts);
    auto lds_bytes = [&](uint32_t spp) {
        return size_t(2) * kGroup * kDgWin * sizeof(float) +
               size_t(2) * kDgPrmBatches * (kGroup * sizeof(DgPair) + sizeof(uint32_t)) +
               size_t(a.groups_per_chunk) * kGroup * sizeof(DgLaser) + size_t(spp) * sizeof(float4) +
               48 * sizeof(float);
    };
    while (a.sensors_per_part > 64 && lds_bytes(a.sensors_per_part) > size_t(ctx->max_lds_per_block))
        a.sensors_per_part = (a.sensors_per_part + 1) / 2;
    const size_t shmem = lds_bytes(a.sensors_per_part);
    if (shmem > size_t(ctx->max_lds_per_block)) return FBP_OK;
    a.sensor_parts = uint32_t((a.S + a.sensors_per_part - 1) / a.sensors_per_part);
    const uint64_t chunks = lgc * a.sensor_parts;
    const size_t n = size_t(a.L) * a.D;
    if (chunks * a.num_tiles > 0x7fffffffull) return FBP_OK;
    // Partial sums [chunk][L * D], then the pixel order.
    const size_t part_bytes = (size_t(chunks) * n * sizeof(float) + 255) & ~size_t(255);
    FBP_CHECK(ensure_scratch(ctx, part_bytes + size_t(a.L) * sizeof(uint32_t)));
    a.partial = static_cast<float*>(ctx->scratch);
    uint32_t* perm = reinterpret_cast<uint32_t*>(static_cast<char*>(ctx->scratch) + part_bytes);
    a.perm = perm;
    a.num_blocks = uint32_t(chunks * a.num_tiles);
    hipLaunchKernelGGL(depth_pixel_order_kernel, dim3(1), dim3(1024), 0, s, laser_targets, a.L, perm);
    FBP_CHECK_HIP(ctx, hipGetLastError());
    const bool coff = !(a.t0 == 0.f && a.discard != 0);
    void (*kern)(DepthGroupArgs) = coff ? depth_group_kernel<true> : depth_group_kernel<false>;
    FBP_CHECK_HIP(ctx, hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                                           hipFuncAttributeMaxDynamicSharedMemorySize, int(shmem)));
    hipLaunchKernelGGL(kern, dim3(a.num_blocks), dim3(512), shmem, s, a);
    FBP_CHECK_HIP(ctx, hipGetLastError());
    hipLaunchKernelGGL(depth_reduce_kernel, dim3(unsigned((n + 255) / 256)), dim3(256), 0, s,
                       a.partial, uint32_t(chunks), n, activation);
    FBP_CHECK_HIP(ctx, hipGetLastError());
    *used = true;
    return FBP_OK;
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

    ctx->last_exhaustive_group = 0;
    if (ctx->exhaustive_group_kernel) {
        bool used = false;
        FBP_CHECK(depth_group_launch(ctx, info, tr, laser_targets, sensor_targets, depths,
                                     num_depths, activation, s, &used));
        if (used) {
            ctx->last_exhaustive_group = 1;
            return FBP_OK;
        }
    }

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
