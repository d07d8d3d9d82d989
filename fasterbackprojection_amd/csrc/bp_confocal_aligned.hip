// Confocal voxel backprojection on aligned grids (ring convolution form).
//
// Same result as backprojectConfocalVoxel (FB/transient_reconstruction.cuh:93-136)
// for the pure chain (t0 = 0, first/last bounces discarded): vol[v] =
// sum over scan points l of I'[l][bin(v, l)] * safeRCP(dist^2), with
// I' = I thresholded at EPS and the bin/dist of the IEEE chain in chain.h.
//
// When the scan points form an nx x nx grid on a plane of constant world y
// and the voxel centres along world x and z sit on the same lattice, the
// rounded per-axis squares fl(fl(p - l)^2) depend only on |vx - ix| and
// |vy - iz|.  A device-side check proves that for every (voxel, scan point)
// pair before anything else runs; then bin and weight are functions of the
// offset (k, m) = (vx - ix, vy - iz) and of the depth index z only, and the
// depth plane z of the volume is
//   vol_z[vy][vx] = sum_b sum_{(k,m) in ring(z,b)} w(k,m,z) * P_b[vy-m][vx-k]
// where P_b is the plane of bin b of the transient (time-major) and
// ring(z, b) = {(k, m) : bin(k, m, z) = b}.  Each workgroup keeps whole
// depth planes of the volume in registers (8x8 voxel tiles, one voxel per
// lane per tile; a wave owns two tile rows), stages one transient plane at
// a time in LDS with zero borders (32 columns, 8 rows), and for every ring
// offset does one LDS read and one FMA per voxel: no per-pair distance,
// square root or division.  The offsets of a depth are enumerated once per
// call by a table kernel: quadrant k, m >= 0 (the four sign combinations
// share bin and weight), sorted by (bin, k / 32, index).  Within a k / 32
// bucket the column groups (4 tiles = 32 voxels) that can meet the plane
// are known at compile time, so the inner bodies are straight-line code;
// tile rows that miss the plane read the wave's own (in-bounds) position
// with weight 0.
//
// Partial sums per bin chunk are reduced in FP64 in a fixed order, so the
// result is deterministic.  Whether this path or the general kernel runs is
// decided on the device by the check kernel's flag: both paths are
// launched, each returns at once unless the flag selects it.
#include <cstdlib>

#include "bp_common.h"
#include "bp_confocal_aligned.h"

namespace fbp {

namespace {

constexpr int kPadRows = 9;    // 1 + 8 zero rows above the data, 8 below
constexpr int kPadCols = 32;   // zero columns on each side (groups of 4 tiles)
constexpr int kKeyBits = 14;   // quadrant index bits in the sort key
constexpr int kBucketBits = 2; // k / 32 bits in the sort key (above the index)
constexpr int kBinShift = kKeyBits + kBucketBits;

struct AlignedParams {
    uint32_t nx;            // scan grid edge, L = nx * nx (l = iz * nx + ix)
    uint32_t X, VY, ZD;     // voxels along vx (world x), vy (world z), z (world y)
    uint32_t kq, mq;        // quadrant extents max(X, nx), max(VY, nx)
    uint32_t T;
    float t0, dt;
    uint32_t discard;
    float hmin[3], vs[3];
};

__device__ __forceinline__ float px_of(const AlignedParams& a, uint32_t vx) {
    return voxel_center(a.hmin[0], a.vs[0], vx);
}
__device__ __forceinline__ float pz_of(const AlignedParams& a, uint32_t vy) {
    return voxel_center(a.hmin[2], a.vs[2], vy);
}

// fl(fl(px - lx)^2) for |vx - ix| = k, from a representative pair.
__device__ __forceinline__ float x2_ref(const AlignedParams& a, const fbp_vec3* laser, uint32_t k) {
    return k < a.X ? sq(rn_sub(px_of(a, k), laser[0].x)) : sq(rn_sub(px_of(a, 0), laser[k].x));
}
__device__ __forceinline__ float z2_ref(const AlignedParams& a, const fbp_vec3* laser, uint32_t m) {
    return m < a.VY ? sq(rn_sub(pz_of(a, m), laser[0].z))
                    : sq(rn_sub(pz_of(a, 0), laser[size_t(m) * a.nx].z));
}

// Device-resident cache header of the offset tables: the parameters and
// scan points they were built for, and whether the build completed.
struct TableCache {
    AlignedParams key;
    int valid;
};

__device__ __forceinline__ bool same_params(const AlignedParams& x, const AlignedParams& y) {
    bool eq = x.nx == y.nx && x.X == y.X && x.VY == y.VY && x.ZD == y.ZD && x.kq == y.kq &&
              x.mq == y.mq && x.T == y.T && x.discard == y.discard;
    // Bit patterns: -0 / +0 and NaNs must not compare equal by value.
    eq = eq && __float_as_uint(x.t0) == __float_as_uint(y.t0) &&
         __float_as_uint(x.dt) == __float_as_uint(y.dt);
    for (int i = 0; i < 3; ++i)
        eq = eq && __float_as_uint(x.hmin[i]) == __float_as_uint(y.hmin[i]) &&
             __float_as_uint(x.vs[i]) == __float_as_uint(y.vs[i]);
    return eq;
}

// flag[0] = 1 iff the grid conditions hold for every scan point and pair.
// flag[1] = 1 iff the cached tables were built for these parameters and
// exactly these scan points (then the table kernel is skipped); when
// flag[0] = 1 and flag[1] = 0 the cache is marked invalid until rebuilt.
__global__ __launch_bounds__(1024) void aligned_check_kernel(const fbp_vec3* __restrict__ laser,
                                                             AlignedParams a, int* __restrict__ flag,
                                                             TableCache* __restrict__ cache,
                                                             const fbp_vec3* __restrict__ cached_laser) {
    __shared__ int ok, same;
    if (threadIdx.x == 0) {
        ok = 1;
        same = cache->valid == 1 && same_params(cache->key, a);
    }
    __syncthreads();
    bool good = true, eq = true;
    const uint32_t L = a.nx * a.nx;
    const float y0 = laser[0].y;
    const bool cmp = same != 0;
    for (uint32_t l = threadIdx.x; l < L; l += blockDim.x) {
        const uint32_t ix = l % a.nx, iz = l / a.nx;
        const fbp_vec3 p = laser[l];
        good = good && p.y == y0 && p.x == laser[ix].x && p.z == laser[size_t(iz) * a.nx].z;
        if (cmp) {
            const fbp_vec3 c = cached_laser[l];
            eq = eq && __float_as_uint(p.x) == __float_as_uint(c.x) &&
                 __float_as_uint(p.y) == __float_as_uint(c.y) &&
                 __float_as_uint(p.z) == __float_as_uint(c.z);
        }
    }
    for (uint32_t q = threadIdx.x; q < a.X * a.nx; q += blockDim.x) {
        const uint32_t vx = q % a.X, ix = q / a.X;
        const float v = sq(rn_sub(px_of(a, vx), laser[ix].x));
        good = good && v == x2_ref(a, laser, vx > ix ? vx - ix : ix - vx);
    }
    for (uint32_t q = threadIdx.x; q < a.VY * a.nx; q += blockDim.x) {
        const uint32_t vy = q % a.VY, iz = q / a.VY;
        const float v = sq(rn_sub(pz_of(a, vy), laser[size_t(iz) * a.nx].z));
        good = good && v == z2_ref(a, laser, vy > iz ? vy - iz : iz - vy);
    }
    if (!good) atomicAnd(&ok, 0);
    if (!eq) atomicAnd(&same, 0);
    __syncthreads();
    if (threadIdx.x == 0) {
        flag[0] = ok;
        flag[1] = ok && same;
        if (ok && !same) cache->valid = 0;
    }
}

// One workgroup per depth index z: bins and weights of the quadrant
// offsets (k, m) in [0, kq) x [0, mq), sorted by (bin, k / 32, m * kq + k),
// plus the first entry of every (bin, bucket): starts[z][4 b + q], and
// starts[z][4 T] = number of offsets whose bin lies inside [0, T).
// work[z][b] = the convolution work (column groups x sign combinations) of
// the entries with bin < b, for balancing the bin chunks.
__global__ __launch_bounds__(1024) void aligned_table_kernel(const fbp_vec3* __restrict__ laser,
                                                             AlignedParams a, const int* __restrict__ flag,
                                                             uint32_t sort_n, uint2* __restrict__ entries,
                                                             uint32_t* __restrict__ starts,
                                                             uint32_t* __restrict__ work,
                                                             TableCache* __restrict__ cache,
                                                             fbp_vec3* __restrict__ cached_laser) {
    if (flag[0] == 0 || flag[1] == 1) return;
    if (blockIdx.x == 0) {
        for (uint32_t l = threadIdx.x; l < a.nx * a.nx; l += blockDim.x) cached_laser[l] = laser[l];
        if (threadIdx.x == 0) cache->key = a;
    }
    extern __shared__ uint32_t keys[];  // sort_n (power of two), then sort_n work sums
    uint32_t* wsum = keys + sort_n;
    __shared__ float X2[128], Z2[128];
    __shared__ uint32_t part[1024];
    const uint32_t z = blockIdx.x;
    const uint32_t tid = threadIdx.x;
    for (uint32_t k = tid; k < a.kq; k += blockDim.x) X2[k] = x2_ref(a, laser, k);
    for (uint32_t m = tid; m < a.mq; m += blockDim.x) Z2[m] = z2_ref(a, laser, m);
    const float Y2 = sq(rn_sub(voxel_center(a.hmin[1], a.vs[1], z), laser[0].y));
    __syncthreads();
    const uint32_t n = a.kq * a.mq;
    for (uint32_t e = tid; e < sort_n; e += blockDim.x) {
        uint32_t key = 0xffffffffu;
        if (e < n) {
            const float s = dot3(X2[e % a.kq], Y2, Z2[e / a.kq]);
            const int bin = bin_of(confocal_dist(s, a.t0, a.discard, 0.f), a.dt);
            if (bin >= 0 && bin < int(a.T))
                key = (uint32_t(bin) << kBinShift) | ((e % a.kq) / 32u) << kKeyBits | e;
        }
        keys[e] = key;
    }
    __syncthreads();
    // Bitonic sort (ascending) of the unique keys.
    for (uint32_t size = 2; size <= sort_n; size <<= 1) {
        for (uint32_t stride = size >> 1; stride > 0; stride >>= 1) {
            for (uint32_t i = tid; i < sort_n; i += blockDim.x) {
                const uint32_t j = i ^ stride;
                if (j > i) {
                    const uint32_t ki = keys[i], kj = keys[j];
                    const bool up = (i & size) == 0;
                    if ((ki > kj) == up) {
                        keys[i] = kj;
                        keys[j] = ki;
                    }
                }
            }
            __syncthreads();
        }
    }
    const uint32_t groups = a.kq / 32u;
    uint2* out = entries + size_t(z) * n;
    for (uint32_t i = tid; i < sort_n; i += blockDim.x) {
        const uint32_t key = keys[i];
        uint32_t wk = 0;
        if (i < n) {
            uint2 ent = make_uint2(0xffffffffu, 0u);
            if (key != 0xffffffffu) {
                const uint32_t e = key & ((1u << kKeyBits) - 1u);
                const uint32_t k = e % a.kq, m = e / a.kq;
                const float s = dot3(X2[k], Y2, Z2[m]);
                const float dist = confocal_dist(s, a.t0, a.discard, 0.f);
                ent = make_uint2(k | (m << 16), __float_as_uint(safe_rcp(rn_mul(dist, dist))));
                wk = (groups - k / 32u) * (k > 0 ? 2u : 1u) * (m > 0 ? 2u : 1u);
            }
            out[i] = ent;
        }
        wsum[i] = wk;
    }
    __syncthreads();
    // Inclusive scan of wsum: per-thread runs, then a scan of the run sums.
    const uint32_t run = sort_n / blockDim.x;  // sort_n is a multiple of 1024 or smaller
    uint32_t local = 0;
    if (run > 0)
        for (uint32_t i = 0; i < run; ++i) local += wsum[tid * run + i];
    part[tid] = local;
    __syncthreads();
    for (uint32_t off = 1; off < blockDim.x; off <<= 1) {
        const uint32_t v = tid >= off ? part[tid - off] : 0u;
        __syncthreads();
        part[tid] += v;
        __syncthreads();
    }
    if (run > 0) {
        uint32_t acc = part[tid] - local;
        for (uint32_t i = 0; i < run; ++i) {
            acc += wsum[tid * run + i];
            wsum[tid * run + i] = acc;
        }
    } else if (tid == 0) {
        for (uint32_t i = 1; i < sort_n; ++i) wsum[i] += wsum[i - 1];
    }
    __syncthreads();
    for (uint32_t i = tid; i <= 4 * a.T; i += blockDim.x) {
        const uint32_t target = ((i >> 2) << kBinShift) | ((i & 3u) << kKeyBits);
        uint32_t lo = 0, hi = n;
        while (lo < hi) {
            const uint32_t mid = (lo + hi) >> 1;
            if (keys[mid] < target)
                lo = mid + 1;
            else
                hi = mid;
        }
        starts[size_t(z) * (4 * a.T + 1) + i] = lo;
        if ((i & 3u) == 0) work[size_t(z) * (a.T + 1) + (i >> 2)] = lo == 0 ? 0u : wsum[lo - 1];
    }
}

// [L][T] -> [T][L] through 64x64 LDS tiles.
__global__ __launch_bounds__(256) void transpose_lt_kernel(const float* __restrict__ in,
                                                           float* __restrict__ out, uint32_t L,
                                                           uint32_t T, const int* __restrict__ flag) {
    if (*flag == 0) return;
    __shared__ float tile[64][65];
    const uint32_t t0 = blockIdx.x * 64, l0 = blockIdx.y * 64;
    const uint32_t tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
    for (uint32_t r = ty; r < 64; r += 4) {
        const uint32_t l = l0 + r, t = t0 + tx;
        tile[r][tx] = (l < L && t < T) ? in[size_t(l) * T + t] : 0.f;
    }
    __syncthreads();
    for (uint32_t r = ty; r < 64; r += 4) {
        const uint32_t t = t0 + r, l = l0 + tx;
        if (l < L && t < T) out[size_t(t) * L + l] = tile[tx][r];
    }
}

struct ConvArgs {
    const int* flag;
    const float* tm;          // transient, [T][L] (time-major)
    const uint2* entries;     // [ZD][n]
    const uint32_t* starts;   // [ZD][4 T + 1]
    const uint32_t* work;     // [ZD][T + 1]
    float* partial;           // [chunks][ZD][VY][X]
    uint32_t nx, X, VY, ZD, T, n, chunks, stride;
};

// Prefetch registers for one transient plane: 4 float4 per thread (planes
// of at most 16384 samples with 1024 threads).
constexpr int kStageVec = 4;

// One signed offset applied to the wave's tile row, tile columns
// 4*GLO .. 4*GHI+3, reading at LDS index a (+ 8 per tile column).  All reads
// are issued before the FMAs.
template <int TXC, int GLO, int GHI>
__device__ __forceinline__ void apply_offset(float (&acc)[TXC], const float* plane, int a, float w) {
    constexpr int C0 = 4 * GLO, C1 = 4 * GHI + 4;
    float v[TXC];
#pragma unroll
    for (int c = C0; c < C1; ++c) v[c] = plane[a + 8 * c];
#pragma unroll
    for (int c = C0; c < C1; ++c) acc[c] = fmaf(w, v[c], acc[c]);
}

// The entries of one k / 32 bucket Q for one depth: every sign combination
// of every entry.  For +k the column groups Q..G-1 can meet the plane, for
// -k the groups 0..G-1-Q (X = nx = 32 G); a group that meets the plane only
// in part reads zeros from the border.
template <int TXC, int Q>
__device__ __forceinline__ void apply_bucket(float (&acc)[TXC], const float* plane,
                                             const uint2* __restrict__ ent, uint32_t e,
                                             uint32_t e_end, int lane_row, int y0, int nx,
                                             int stride) {
    constexpr int G = TXC / 4;
    for (; e < e_end; ++e) {
        const uint2 q = ent[e];
        const uint32_t km = uint32_t(rfl(int(q.x)));
        const float w = rfl(__uint_as_float(q.y));
        const int k = int(km & 0xffffu), m = int(km >> 16);
        for (int ms = 0; ms < 2; ++ms) {
            if (ms == 1 && m == 0) break;
            const int sm = ms ? -m : m;
            // Does the wave's tile row (voxel rows y0 .. y0+7) meet the plane?
            if (y0 + 7 - sm < 0 || y0 - sm >= nx) continue;
            const int a = lane_row - sm * stride;
            apply_offset<TXC, Q, G - 1>(acc, plane, a - k, w);
            if (k > 0) apply_offset<TXC, 0, G - 1 - Q>(acc, plane, a + k, w);
        }
    }
}

template <int TXC>
__device__ __forceinline__ void apply_plane(float (&acc)[TXC], const float* plane,
                                            const uint2* __restrict__ ent,
                                            const uint32_t* __restrict__ sb, int lane_row, int y0,
                                            int nx, int stride) {
    constexpr int G = TXC / 4;
    uint32_t bq[5];
#pragma unroll
    for (int q = 0; q <= 4; ++q) bq[q] = uint32_t(rfl(int(sb[q])));
    apply_bucket<TXC, 0>(acc, plane, ent, bq[0], bq[1], lane_row, y0, nx, stride);
    apply_bucket<TXC, 1>(acc, plane, ent, bq[1], bq[2], lane_row, y0, nx, stride);
    if constexpr (G == 4) {
        apply_bucket<TXC, 2>(acc, plane, ent, bq[2], bq[3], lane_row, y0, nx, stride);
        apply_bucket<TXC, 3>(acc, plane, ent, bq[3], bq[4], lane_row, y0, nx, stride);
    }
}

// Workgroup: ZB consecutive depth planes x one chunk of bins.  Wave w owns
// voxel rows 8w .. 8w+7 of those planes (TXC tiles of 8x8, one voxel per
// lane per tile and depth).
template <int TXC, int ZB>
__global__ __launch_bounds__(1024) void aligned_conv_kernel(ConvArgs a,
                                                            const uint2* __restrict__ entries,
                                                            const uint32_t* __restrict__ starts,
                                                            const uint32_t* __restrict__ work) {
    static_assert(ZB == 1 || ZB == 2 || ZB == 4, "1, 2 or 4 depth planes per workgroup");
    if (*a.flag == 0) return;
    extern __shared__ float plane[];
    const uint32_t tid = threadIdx.x;
    const uint32_t nthreads = blockDim.x;
    const int wave = __builtin_amdgcn_readfirstlane(int(tid >> 6));
    const int lane = int(tid & 63);
    const int lx = lane & 7, ly = lane >> 3;
    // Chunk-major order: the workgroups resident at one time work on the
    // same share of every depth's bins, so the planes they stage overlap
    // (L2 reuse across neighbouring depths).
    const uint32_t groups = a.ZD / ZB;
    const uint32_t chunk = blockIdx.x / groups, group = blockIdx.x % groups;
    const uint32_t z0 = group * ZB;
    const int nx = int(a.nx);
    const int stride = int(a.stride);
    const uint32_t L = a.nx * a.nx;
    const uint32_t rows = a.nx + 2 * kPadRows - 1;
    const uint32_t srow = 4 * a.T + 1;  // starts per depth
    const uint32_t wrow = a.T + 1;      // work prefix per depth

    // Zero the whole plane buffer once: the borders stay zero.
    for (uint32_t i = tid; i < rows * a.stride; i += nthreads) plane[i] = 0.f;

    // Bin range of this chunk: equal shares of the group's convolution work.
    auto group_work = [&](uint32_t b) {
        uint32_t w = 0;
#pragma unroll
        for (int j = 0; j < ZB; ++j) w += work[size_t(z0 + j) * wrow + b];
        return w;
    };
    auto first_bin_at = [&](uint64_t target) {
        uint32_t lo = 0, hi = a.T;  // smallest b with group_work(b) >= target
        while (lo < hi) {
            const uint32_t mid = (lo + hi) >> 1;
            if (group_work(mid) < target)
                lo = mid + 1;
            else
                hi = mid;
        }
        return lo;
    };
    const uint64_t W = group_work(a.T);
    const uint32_t b_begin = chunk == 0 ? 0u : first_bin_at(W * chunk / a.chunks);
    const uint32_t b_end = chunk + 1 == a.chunks ? a.T : first_bin_at(W * (chunk + 1) / a.chunks);

    auto nonempty = [&](uint32_t b) {
        bool any = false;
#pragma unroll
        for (int j = 0; j < ZB; ++j) {
            const uint32_t* sj = starts + size_t(z0 + j) * srow;
            any = any || (sj[4 * b + 4] > sj[4 * b]);
        }
        return any;
    };
    auto next_bin = [&](uint32_t b) {
        while (b < b_end && !nonempty(b)) ++b;
        return b;
    };

    float acc0[TXC], acc1[TXC], acc2[TXC], acc3[TXC];
#pragma unroll
    for (int c = 0; c < TXC; ++c) acc0[c] = acc1[c] = acc2[c] = acc3[c] = 0.f;

    float4 stage[kStageVec];
    auto load_plane = [&](uint32_t b) {
        const float4* src = reinterpret_cast<const float4*>(a.tm + size_t(b) * L);
#pragma unroll
        for (int i = 0; i < kStageVec; ++i) {
            const uint32_t f = (tid + uint32_t(i) * nthreads) * 4;
            stage[i] = f < L ? src[f / 4] : make_float4(0.f, 0.f, 0.f, 0.f);
        }
    };
    auto thr = [](float v) { return v > kEps ? v : 0.f; };
    auto store4 = [&](uint32_t f, float4 v) {
        const uint32_t iz = f / a.nx, ix = f % a.nx;
        v.x = thr(v.x);
        v.y = thr(v.y);
        v.z = thr(v.z);
        v.w = thr(v.w);
        *reinterpret_cast<float4*>(plane + (kPadRows + iz) * a.stride + kPadCols + ix) = v;
    };
    // Planes that fit the prefetch registers are loaded while the previous
    // plane is processed; larger ones (few threads) are copied directly.
    const bool prefetch = L <= 4u * kStageVec * nthreads;

    const int y0 = 8 * wave;
    const int lane_row = (kPadRows + y0 + ly) * stride + kPadCols + lx;
    auto ent_of = [&](int j) { return entries + size_t(z0 + j) * a.n; };
    auto st_of = [&](int j) { return starts + size_t(z0 + j) * srow; };
    uint32_t b = next_bin(b_begin);
    if (prefetch && b < b_end) load_plane(b);
    while (b < b_end) {
        __syncthreads();
        if (prefetch) {
#pragma unroll
            for (int i = 0; i < kStageVec; ++i) {
                const uint32_t f = (tid + uint32_t(i) * nthreads) * 4;
                if (f < L) store4(f, stage[i]);
            }
        } else {
            const float4* src = reinterpret_cast<const float4*>(a.tm + size_t(b) * L);
            for (uint32_t f = tid * 4; f < L; f += 4 * nthreads) store4(f, src[f / 4]);
        }
        __syncthreads();
        const uint32_t bn = next_bin(b + 1);
        if (prefetch && bn < b_end) load_plane(bn);
        apply_plane<TXC>(acc0, plane, ent_of(0), st_of(0) + 4 * b, lane_row, y0, nx, stride);
        if constexpr (ZB >= 2)
            apply_plane<TXC>(acc1, plane, ent_of(1), st_of(1) + 4 * b, lane_row, y0, nx, stride);
        if constexpr (ZB == 4) {
            apply_plane<TXC>(acc2, plane, ent_of(2), st_of(2) + 4 * b, lane_row, y0, nx, stride);
            apply_plane<TXC>(acc3, plane, ent_of(3), st_of(3) + 4 * b, lane_row, y0, nx, stride);
        }
        b = bn;
    }

    // Partial sums of this chunk for the ZB depth planes.
    float* out0 = a.partial + ((size_t(chunk) * a.ZD + z0) * a.VY + y0 + ly) * a.X + lx;
    const size_t zs = size_t(a.VY) * a.X;
#pragma unroll
    for (int c = 0; c < TXC; ++c) {
        out0[8 * c] = acc0[c];
        if constexpr (ZB >= 2) out0[zs + 8 * c] = acc1[c];
        if constexpr (ZB == 4) {
            out0[2 * zs + 8 * c] = acc2[c];
            out0[3 * zs + 8 * c] = acc3[c];
        }
    }
}

__global__ void cache_reset_kernel(TableCache* cache) {
    if (threadIdx.x == 0) cache->valid = 0;
}

__global__ void aligned_reduce_kernel(const float* __restrict__ partial, float* __restrict__ volume,
                                      size_t nvox, uint32_t chunks, const int* __restrict__ flag,
                                      TableCache* __restrict__ cache) {
    if (*flag == 0) return;
    // The tables (cached or rebuilt by the table kernel, which has finished)
    // now match the key recorded in the cache header.
    if (blockIdx.x == 0 && threadIdx.x == 0) cache->valid = 1;
    const size_t v = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (v >= nvox) return;
    double s = 0.0;
    for (uint32_t c = 0; c < chunks; ++c) s += double(partial[size_t(c) * nvox + v]);
    volume[v] = float(s);
}

uint32_t lds_stride(uint32_t nx) {
    uint32_t s = nx + 2 * kPadCols;
    while (s % 64 != 8 && s % 64 != 24 && s % 64 != 40 && s % 64 != 56) ++s;
    return s;
}

uint32_t pow2_at_least(uint32_t n) {
    uint32_t p = 1;
    while (p < n) p <<= 1;
    return p;
}

}  // namespace

AlignedPlan plan_aligned_confocal(fbp_context ctx, const fbp_reconstruction_info& info,
                                  const fbp_transient& tr) {
    AlignedPlan p{};
    p.usable = false;
    if (ctx->disable_aligned) return p;
    const bool pure = info.time_offset == 0.f && info.discard_first_last_bounces != 0;
    if (!pure || tr.storage != FBP_STORAGE_F32) return p;
    const uint32_t L = info.num_laser_targets;
    uint32_t nx = uint32_t(std::lround(std::sqrt(double(L))));
    if (nx * nx != L || nx < 8 || nx > 128 || nx % 4 != 0) return p;
    const uint32_t X = info.voxel_resolution.x, VY = info.voxel_resolution.y,
                   ZD = info.voxel_resolution.z;
    // X = 8 * TXC with TXC in {8, 16}; VY a multiple of 16 up to 128.
    // X = nx = 32 * (TXC / 4) with TXC in {8, 16}; VY a multiple of 16, at
    // most nx.
    if (!(X == 128 || X == 64) || X != nx || VY % 16 != 0 || VY == 0 || VY > nx || ZD == 0)
        return p;
    if (info.num_time_bins == 0 || info.num_time_bins >= (1u << 16)) return p;
    p.nx = nx;
    p.stride = lds_stride(nx);
    p.lds_bytes = size_t(nx + 2 * kPadRows - 1) * p.stride * sizeof(float);
    if (p.lds_bytes > size_t(ctx->max_lds_per_block)) return p;
    p.kq = std::max(X, nx);
    p.mq = std::max(VY, nx);
    p.n = p.kq * p.mq;
    p.sort_n = pow2_at_least(p.n);
    if (size_t(p.sort_n) * 8 + 6144 > size_t(ctx->max_lds_per_block)) return p;
    p.zb = 1;
    const uint32_t groups = ZD / p.zb;
    // ~8 workgroups per CU over the whole launch (one resident per CU).
    const uint32_t target = 8u * uint32_t(ctx->num_cus > 0 ? ctx->num_cus : 256);
    p.chunks = std::max(1u, std::min(64u, (target + groups - 1) / groups));
#ifdef FBP_TUNING_ENV
    if (const char* e = std::getenv("FBP_ALIGNED_ZB")) p.zb = uint32_t(std::atoi(e));
    if (const char* e = std::getenv("FBP_ALIGNED_CHUNKS")) p.chunks = uint32_t(std::atoi(e));
    if (ZD % p.zb) p.zb = 1;
#endif
    const size_t nvox = size_t(X) * VY * ZD;
    p.bytes_entries = size_t(ZD) * p.n * sizeof(uint2);
    p.bytes_starts = size_t(ZD) * (4 * size_t(info.num_time_bins) + 1) * sizeof(uint32_t);
    p.bytes_work = size_t(ZD) * (size_t(info.num_time_bins) + 1) * sizeof(uint32_t);
    p.bytes_tm = tr.layout == FBP_LAYOUT_TIME_MAJOR ? 0 : size_t(L) * info.num_time_bins * 4;
    p.bytes_partial = size_t(p.chunks) * nvox * sizeof(float);
    p.usable = true;
    return p;
}

static size_t align256(size_t b) { return (b + 255) & ~size_t(255); }

size_t aligned_scratch_bytes(const AlignedPlan& p) {
    return align256(p.bytes_tm) + align256(p.bytes_partial);
}

template <int TXC, int ZB>
static fbp_status launch_conv(fbp_context ctx, const ConvArgs& c, unsigned blocks, unsigned threads,
                              size_t lds, hipStream_t s) {
    auto kern = aligned_conv_kernel<TXC, ZB>;
    FBP_CHECK_HIP(ctx, hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                                           hipFuncAttributeMaxDynamicSharedMemorySize, int(lds)));
    hipLaunchKernelGGL(kern, dim3(blocks), dim3(threads), lds, s, c, c.entries, c.starts, c.work);
    FBP_CHECK_HIP(ctx, hipGetLastError());
    return FBP_OK;
}

fbp_status launch_aligned_confocal(fbp_context ctx, const AlignedPlan& p,
                                   const fbp_reconstruction_info& info, const fbp_transient& tr,
                                   const fbp_vec3* laser, float* volume, int* flag, void* base,
                                   hipStream_t s) {
    AlignedParams a{};
    a.nx = p.nx;
    a.X = info.voxel_resolution.x;
    a.VY = info.voxel_resolution.y;
    a.ZD = info.voxel_resolution.z;
    a.kq = p.kq;
    a.mq = p.mq;
    a.T = info.num_time_bins;
    a.t0 = info.time_offset;
    a.dt = info.time_step;
    a.discard = info.discard_first_last_bounces;
    a.hmin[0] = info.hidden_volume_min.x;
    a.hmin[1] = info.hidden_volume_min.y;
    a.hmin[2] = info.hidden_volume_min.z;
    a.vs[0] = info.hidden_volume_voxel_size.x;
    a.vs[1] = info.hidden_volume_voxel_size.y;
    a.vs[2] = info.hidden_volume_voxel_size.z;

    // Offset tables live in their own allocation (kept across calls, rebuilt
    // only when the parameters or the scan points change); the transposed
    // transient and the partial sums use the shared scratch.
    const size_t header = align256(sizeof(TableCache)) + align256(size_t(16384) * sizeof(fbp_vec3));
    const size_t table_bytes = header + align256(p.bytes_entries) + align256(p.bytes_starts) +
                               align256(p.bytes_work);
    if (ctx->aligned_tables_bytes < table_bytes) {
        if (ctx->aligned_tables) {
            FBP_CHECK_HIP(ctx, hipDeviceSynchronize());
            FBP_CHECK_HIP(ctx, hipFree(ctx->aligned_tables));
            ctx->aligned_tables = nullptr;
            ctx->aligned_tables_bytes = 0;
        }
        FBP_CHECK_HIP(ctx, hipMalloc(&ctx->aligned_tables, table_bytes));
        ctx->aligned_tables_bytes = table_bytes;
        // A fresh cache is invalid until a build completes.
        hipLaunchKernelGGL(cache_reset_kernel, dim3(1), dim3(64), 0, s,
                           static_cast<TableCache*>(ctx->aligned_tables));
        FBP_CHECK_HIP(ctx, hipGetLastError());
    }
    char* tcur = static_cast<char*>(ctx->aligned_tables);
    TableCache* cache = reinterpret_cast<TableCache*>(tcur);
    fbp_vec3* cached_laser = reinterpret_cast<fbp_vec3*>(tcur + align256(sizeof(TableCache)));
    tcur += header;
    uint2* entries = reinterpret_cast<uint2*>(tcur);
    tcur += align256(p.bytes_entries);
    uint32_t* starts = reinterpret_cast<uint32_t*>(tcur);
    tcur += align256(p.bytes_starts);
    uint32_t* work = reinterpret_cast<uint32_t*>(tcur);

    char* cur = static_cast<char*>(base);
    float* tm = reinterpret_cast<float*>(cur);
    cur += align256(p.bytes_tm);
    float* partial = reinterpret_cast<float*>(cur);

    hipLaunchKernelGGL(aligned_check_kernel, dim3(1), dim3(1024), 0, s, laser, a, flag, cache,
                       cached_laser);
    FBP_CHECK_HIP(ctx, hipGetLastError());
    FBP_CHECK_HIP(ctx, hipFuncSetAttribute(reinterpret_cast<const void*>(aligned_table_kernel),
                                           hipFuncAttributeMaxDynamicSharedMemorySize,
                                           int(p.sort_n * 8)));
    hipLaunchKernelGGL(aligned_table_kernel, dim3(a.ZD), dim3(1024), size_t(p.sort_n) * 8, s, laser,
                       a, flag, p.sort_n, entries, starts, work, cache, cached_laser);
    FBP_CHECK_HIP(ctx, hipGetLastError());
    const uint32_t L = p.nx * p.nx;
    const float* src = static_cast<const float*>(tr.data);
    if (tr.layout != FBP_LAYOUT_TIME_MAJOR) {
        hipLaunchKernelGGL(transpose_lt_kernel, dim3((a.T + 63) / 64, (L + 63) / 64), dim3(256), 0,
                           s, src, tm, L, a.T, flag);
        FBP_CHECK_HIP(ctx, hipGetLastError());
        src = tm;
    }
    ConvArgs c{};
    c.flag = flag;
    c.tm = src;
    c.entries = entries;
    c.starts = starts;
    c.work = work;
    c.partial = partial;
    c.nx = p.nx;
    c.X = a.X;
    c.VY = a.VY;
    c.ZD = a.ZD;
    c.T = a.T;
    c.n = p.n;
    c.chunks = p.chunks;
    c.stride = p.stride;
    const unsigned blocks = (a.ZD / p.zb) * p.chunks;
    const unsigned threads = 64u * (a.VY / 8u);
    fbp_status st;
    if (a.X == 128)
        st = p.zb == 4   ? launch_conv<16, 4>(ctx, c, blocks, threads, p.lds_bytes, s)
             : p.zb == 2 ? launch_conv<16, 2>(ctx, c, blocks, threads, p.lds_bytes, s)
                         : launch_conv<16, 1>(ctx, c, blocks, threads, p.lds_bytes, s);
    else
        st = p.zb == 4   ? launch_conv<8, 4>(ctx, c, blocks, threads, p.lds_bytes, s)
             : p.zb == 2 ? launch_conv<8, 2>(ctx, c, blocks, threads, p.lds_bytes, s)
                         : launch_conv<8, 1>(ctx, c, blocks, threads, p.lds_bytes, s);
    if (st != FBP_OK) return st;
    const size_t nvox = size_t(a.X) * a.VY * a.ZD;
    hipLaunchKernelGGL(aligned_reduce_kernel, dim3(unsigned((nvox + 255) / 256)), dim3(256), 0, s,
                       partial, volume, nvox, p.chunks, flag, cache);
    FBP_CHECK_HIP(ctx, hipGetLastError());
    return FBP_OK;
}

}  // namespace fbp
