// Confocal voxel backprojection (gather form).
//
// Reference: backprojectConfocalVoxel (FB/transient_reconstruction.cuh:93-136)
// launched by Laser::reconstructAABBConfocal (FB/Laser.cu:482-498) with one
// thread and one float atomicAdd per (voxel, scan point).
//
// Here one wave owns a 16^3 voxel tile (64 voxels per lane in registers) and a
// chunk of scan points.  For every scan point the wave stages the window of
// time bins its tile can reach into LDS (thresholded at I > 1e-5, zero outside
// [0, T)), then every lane gathers from that window for its 64 voxels.
//
// Exact bins without per-voxel IEEE sqrt/div: the reference's bin is a
// monotone step function of the squared distance s (sqrt, *2, -t0, +off, /dt
// and truncation are all monotone), so with S_k = min{s : bin(s) >= k} the
// bin is the k with S_k <= s < S_{k+1}.  A pre-pass computes S_k by binary
// search over float bit patterns using the IEEE chain (one row when t0 = 0
// and the first/last bounces are discarded, one row per scan point
// otherwise).  The window stores {I[k], S[k]} pairs; the fast path takes a
// v_rsq estimate biased low so it is k or k-1, reads entries ka and ka+1 in
// one LDS access and picks I[ka+1] iff s >= S[ka+1].  s itself is exact:
// per-axis tables in glm::dot's order.  Chunk partial sums are reduced in
// FP64 in a fixed order, so results are deterministic.
#include "bp_common.h"
#include "bp_confocal_aligned.h"

namespace fbp {

struct ConfocalArgs {
    const void* data;
    const float* scales;
    const fbp_vec3* laser;
    float* partial;
    VolumeGeom g;
    uint32_t L, T;
    float t0, dt, inv_dt;
    uint32_t discard;
    float lpos[3], spos[3];
    uint32_t chunk_len;
    int batch;
    int W;
    float inv_W;
    const float* thr;      // S_k rows, (T + 1) floats each
    uint32_t thr_rows;     // 1 (pure chain) or L
    float* corr;           // near-voxel corrections, same layout as partial
    uint8_t* corr_used;    // [chunk][tile]: 1 when the block wrote its corr slice
    const int* skip;       // when non-null and *skip != 0 the aligned path ran: return
    // Tile-sampled mode (MIS): each voxel tile visits its own list of
    // `items` scan points sample_l[tile][i] with weights sample_w[tile][i];
    // otherwise items = L and scan point i is visited with weight 1.
    uint32_t items;
    const uint32_t* sample_l;
    const float* sample_w;
};

// Per scan point parameters staged in LDS (kPrmWords words each): position,
// window base, estimate bias, offsets, near flag, scan point index and
// sample weight.
enum { P_LX = 0, P_LY, P_LZ, P_BASE, P_QBIAS, P_COFF, P_NEAR, P_OFF, P_L, P_WL, kPrmWords };
// Window elements per lane whose loads are issued together.
constexpr int kStage = 8;

// Fast gather of one (voxel, scan point) pair.  s is exact (x4 when PURE);
// the estimate ka is biased low so the exact bin is ka or ka + 1, decided by
// the staged threshold S_{ka+1}.  e = {I[k], S[k]} pairs from the window.
template <bool PURE>
__device__ __forceinline__ float confocal_sample(float s, float cb, float coff, float qbias,
                                                 float inv_dt, const float2* win, int base,
                                                 float& w) {
    const float rs = __builtin_amdgcn_rsqf(s);
    int ka;
    if constexpr (PURE) {
        ka = static_cast<int>((s * rs) * cb);     // ~ 2 r / dt, biased low
        w = rs * rs;                              // ~ 1 / (2r)^2
    } else {
        const float dist = fmaf(s * rs, 2.0f, coff);
        ka = static_cast<int>(fmaf(dist, inv_dt, -qbias));
        w = __builtin_amdgcn_rcpf(dist * dist);
    }
    const float2 e0 = win[base + ka];
    const float2 e1 = win[base + ka + 1];
    return (s >= e1.y) ? e1.x : e0.x;
}

// KAHAN: compensated summation (precision form): comp carries the low-order
// bits lost by each add, acc + comp is written out.
template <bool PURE, bool KAHAN>
__device__ __forceinline__ void confocal_update_tile(float (&acc)[64], float (&comp)[64],
                                                     const float (&A)[4][4], const float (&Z2)[4],
                                                     const float2* win, int base, float cb,
                                                     float coff, float qbias, float inv_dt) {
#pragma unroll
    for (int i = 0; i < 4; ++i) {
#pragma unroll
        for (int j = 0; j < 4; ++j) {
#pragma unroll
            for (int m = 0; m < 4; ++m) {
                const int slot = (i * 4 + j) * 4 + m;
                // PURE: tables are pre-scaled by 4, so s = 4 * (exact s) exactly.
                const float s = A[i][j] + Z2[m];
                float w;
                const float v = confocal_sample<PURE>(s, cb, coff, qbias, inv_dt, win, base, w);
                if constexpr (KAHAN) {
                    const float y = fmaf(v, w, -comp[slot]);
                    const float t = acc[slot] + y;
                    comp[slot] = (t - acc[slot]) - y;
                    acc[slot] = t;
                } else {
                    acc[slot] = fmaf(v, w, acc[slot]);
                }
            }
        }
    }
}

__device__ __forceinline__ float pick4(const float (&v)[4], int i) {
    return i == 0 ? v[0] : i == 1 ? v[1] : i == 2 ? v[2] : v[3];
}

template <int STORAGE, int LAYOUT, bool PURE>
__device__ __forceinline__ void confocal_voxel_body(const ConfocalArgs& a) {
    // Quantised (precision-form) transients accumulate with Kahan compensation.
    constexpr bool KAHAN = STORAGE != FBP_STORAGE_F32;
    if (a.skip && *a.skip) return;
    extern __shared__ float smem[];
    float2* win = reinterpret_cast<float2*>(smem);           // batch * W pairs
    float* prm = smem + 2 * a.batch * a.W;                    // batch * kPrmWords

    const int lane = threadIdx.x;
    const uint32_t tile = blockIdx.x % a.g.num_tiles;
    const uint32_t chunk = blockIdx.x / a.g.num_tiles;
    const uint32_t tx = tile % a.g.tiles[0];
    const uint32_t ty = (tile / a.g.tiles[0]) % a.g.tiles[1];
    const uint32_t tz = tile / (a.g.tiles[0] * a.g.tiles[1]);
    const uint32_t X0 = tx * a.g.tile[0], Y0 = ty * a.g.tile[1], Z0 = tz * a.g.tile[2];
    const int lx = lane & 3, lz = (lane >> 2) & 3, ly = lane >> 4;

    // World coordinates of the lane's voxel centres: vx -> world x,
    // z -> world y (depth), vy -> world z (FB/transient_reconstruction.cuh:106).
    float px[4], py[4], pz[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        px[i] = voxel_center(a.g.hmin[0], a.g.vs[0], X0 + lx + 4 * i);
        py[i] = voxel_center(a.g.hmin[1], a.g.vs[1], Z0 + lz + 4 * i);
        pz[i] = voxel_center(a.g.hmin[2], a.g.vs[2], Y0 + ly + 4 * i);
    }
    float lo[3], hi[3];
    tile_box(a.g, X0, Y0, Z0, lo, hi);

    float acc[64], comp[64];
#pragma unroll
    for (int i = 0; i < 64; ++i) acc[i] = comp[i] = 0.f;
    // Near-voxel corrections (rare) go to a lane-owned slice of a.corr;
    // corr_used tells the reduction whether any scan point of the block
    // wrote it.
    float* corr = a.corr + (size_t(chunk) * a.g.num_tiles + tile) * 64 * 64;
#pragma unroll 8
    for (int s = 0; s < 64; ++s) corr[s * 64 + lane] = 0.f;
    bool corr_live = false;

    const uint32_t l_begin = chunk * a.chunk_len;
    const uint32_t l_end = min(a.items, l_begin + a.chunk_len);
    const uint32_t* tile_l = a.sample_l ? a.sample_l + size_t(tile) * a.items : nullptr;
    const float* tile_w = a.sample_l ? a.sample_w + size_t(tile) * a.items : nullptr;
    const float tscale = PURE ? 4.0f : 1.0f;
    // Bias of the pure-chain estimate: 2^-18 relative, far above its error.
    const float cb = a.inv_dt * (1.0f - 3.814697265625e-06f);

    for (uint32_t lb = l_begin; lb < l_end; lb += a.batch) {
        const int nb = int(min(uint32_t(a.batch), l_end - lb));
        __syncthreads();
        if (lane < nb) {
            const uint32_t l = tile_l ? tile_l[lb + lane] : lb + lane;
            const float p[3] = {a.laser[l].x, a.laser[l].y, a.laser[l].z};
            float off = 0.f;
            if (a.discard == 0)
                off = rn_add(distance3(p[0], p[1], p[2], a.lpos[0], a.lpos[1], a.lpos[2]),
                             distance3(p[0], p[1], p[2], a.spos[0], a.spos[1], a.spos[2]));
            const float rmin = box_min_dist(p, lo, hi);
            const float rmax = box_max_dist(p, lo, hi);
            const float dmin = 2.f * rmin - a.t0 + off;
            const float dmax = 2.f * rmax - a.t0 + off;
            const int kmin = int(floorf(dmin * a.inv_dt)) - 3;
            // Low bias of the general-chain estimate (in bins): a generous
            // multiple of the rounding error of the largest term of dist.
            const float mag = 2.f * rmax + fabsf(a.t0) + fabsf(off);
            const float qbias = 32.f * 1.1920929e-7f * mag * a.inv_dt + 1.0e-5f;
            float* P = prm + lane * kPrmWords;
            P[P_LX] = p[0];
            P[P_LY] = p[1];
            P[P_LZ] = p[2];
            P[P_BASE] = __int_as_float(lane * a.W - kmin);
            P[P_QBIAS] = qbias;
            P[P_COFF] = off - a.t0;
            P[P_NEAR] = ((dmin < 0.0045f && dmax > -0.0045f) || rmin < 1.0e-3f) ? 1.f : 0.f;
            P[P_OFF] = off;
            P[P_L] = __uint_as_float(l);
            P[P_WL] = tile_l ? tile_w[lb + lane] : 1.f;
        }
        __syncthreads();
        // Window staging, one 64-bin block of one scan point per step: in
        // step (jl, wb) lane w stages bin kmin + 64 wb + w of scan point jl
        // (the sample and the threshold; lanes past the window, and bins
        // outside [0, T), read element 0 of the row and store zeros).  The
        // scan point's row, offsets and weight are uniform per step; each
        // lane issues the loads of kStage steps before writing any.
        const int nwb = (a.W + 63) / 64;
        const int steps = nb * nwb;
        for (int s0 = 0; s0 < steps; s0 += kStage) {
            float v[kStage], t[kStage];
            int dst[kStage];
#pragma unroll
            for (int u = 0; u < kStage; ++u) {
                const int st = s0 + u;
                const bool live = st < steps;
                const int jl = live ? st / nwb : 0;
                const int w = (st - jl * nwb) * 64 + lane;
                const float* P = prm + jl * kPrmWords;
                const uint32_t l = uint32_t(rfl(__float_as_int(P[P_L])));
                const int kmin = jl * a.W - rfl(__float_as_int(P[P_BASE]));
                const float wl = rfl(P[P_WL]);
                const int k = kmin + w;
                const bool inw = live && w < a.W;
                const bool ok = inw && k >= 0 && k < int(a.T);
                const uint32_t kc = ok ? uint32_t(k) : 0u;
                const size_t idx = (LAYOUT == FBP_LAYOUT_TIME_FASTEST)
                                       ? size_t(l) * a.T + kc
                                       : size_t(kc) * a.L + l;
                float sc = 1.f;
                if constexpr (STORAGE != FBP_STORAGE_F32) sc = a.scales[l];
                float x = load_sample<STORAGE>(a.data, idx, sc);
                // Quantised storage was thresholded before quantisation.
                if constexpr (STORAGE == FBP_STORAGE_F32) x = (x > kEps) ? x : 0.f;
                v[u] = ok ? x * wl : 0.f;
                // Threshold of bin k (only bins 0..T matter; outside, both
                // candidate samples are zero).
                const float* thr = a.thr + size_t(a.thr_rows > 1 ? l : 0) * (a.T + 1);
                const float th = thr[min(max(k, 0), int(a.T))];
                t[u] = (k < 0) ? 0.f : (k > int(a.T) ? __int_as_float(0x7f800000) : th);
                dst[u] = inw ? jl * a.W + w : -1;
            }
#pragma unroll
            for (int u = 0; u < kStage; ++u)
                if (dst[u] >= 0) win[dst[u]] = make_float2(v[u], t[u] * tscale);
        }
        __syncthreads();
        for (int jl = 0; jl < nb; ++jl) {
            const float* P = prm + jl * kPrmWords;
            const float lxp = rfl(P[P_LX]), lyp = rfl(P[P_LY]), lzp = rfl(P[P_LZ]);
            const int base = rfl(__float_as_int(P[P_BASE]));
            const float qbias = rfl(P[P_QBIAS]);
            const float coff = rfl(P[P_COFF]);
            const bool near = rfl(P[P_NEAR]) != 0.f;
            if (__builtin_expect(near, 0)) {
                // Scan point (almost) on a voxel centre, or dist near zero:
                // every voxel of the tile goes through the IEEE chain and
                // lands in the lane-owned correction slots (acc untouched).
                corr_live = true;
                const float off = rfl(P[P_OFF]);
                for (int slot = 0; slot < 64; ++slot) {
                    const float s = rn_add(rn_add(sq(rn_sub(pick4(px, slot >> 4), lxp)),
                                                  sq(rn_sub(pick4(py, (slot >> 2) & 3), lyp))),
                                           sq(rn_sub(pick4(pz, slot & 3), lzp)));
                    const float dist = confocal_dist(s, a.t0, a.discard, off);
                    const int ke = bin_of(dist, a.dt);
                    corr[slot * 64 + lane] += win[base + ke].x * safe_rcp(rn_mul(dist, dist));
                }
                continue;
            }
            float X2[4], Y2[4], Z2[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                X2[i] = sq(rn_sub(px[i], lxp)) * tscale;
                Y2[i] = sq(rn_sub(py[i], lyp)) * tscale;
                Z2[i] = sq(rn_sub(pz[i], lzp)) * tscale;
            }
            float A[4][4];
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j) A[i][j] = rn_add(X2[i], Y2[j]);
            confocal_update_tile<PURE, KAHAN>(acc, comp, A, Z2, win, base, cb, coff, qbias, a.inv_dt);
        }
    }

    if (lane == 0) a.corr_used[size_t(chunk) * a.g.num_tiles + tile] = corr_live ? 1 : 0;
    float* out = a.partial + (size_t(chunk) * a.g.num_tiles + tile) * 64 * 64;
#pragma unroll
    for (int s = 0; s < 64; ++s) out[s * 64 + lane] = KAHAN ? acc[s] - comp[s] : acc[s];
}

// FP32 storage: 2 waves per SIMD (<= 256 VGPRs).  The Kahan variants keep
// 128 more accumulators and run at 1 wave per SIMD.
template <int STORAGE, int LAYOUT, bool PURE>
__global__ __launch_bounds__(64, 2) void confocal_voxel_kernel(ConfocalArgs a) {
    confocal_voxel_body<STORAGE, LAYOUT, PURE>(a);
}
template <int STORAGE, int LAYOUT, bool PURE>
__global__ __launch_bounds__(64, 1) void confocal_voxel_kahan_kernel(ConfocalArgs a) {
    confocal_voxel_body<STORAGE, LAYOUT, PURE>(a);
}

// S_k = min{s >= 0 : bin(s) >= k}, k = 0..T, by binary search over the bit
// patterns of non-negative floats (monotone in the value).  One row per scan
// point, or a single row for the pure chain (t0 = 0, bounces discarded).
__global__ void confocal_thresholds_kernel(const fbp_vec3* __restrict__ laser, uint32_t rows,
                                           uint32_t T, float t0, float dt, uint32_t discard,
                                           fbp_vec3 lpos, fbp_vec3 spos, float* __restrict__ thr,
                                           const int* __restrict__ skip) {
    if (skip && *skip) return;
    const size_t idx = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (idx >= size_t(rows) * (T + 1)) return;
    const uint32_t row = uint32_t(idx / (T + 1));
    const int k = int(idx % (T + 1));
    float off = 0.f;
    if (discard == 0) {
        const float lx = laser[row].x, ly = laser[row].y, lz = laser[row].z;
        off = rn_add(distance3(lx, ly, lz, lpos.x, lpos.y, lpos.z),
                     distance3(lx, ly, lz, spos.x, spos.y, spos.z));
    }
    // bin(s) >= k ?  (bin = (int)(dist / dt), dist from the IEEE chain)
    auto reaches = [&](uint32_t bits) {
        const float s = __uint_as_float(bits);
        const float q = rn_div(confocal_dist(s, t0, discard, off), dt);
        return q >= float(k) || (k <= 0 && q > -1.0f);
    };
    uint32_t lo = 0u, hi = 0x7f800000u;  // +inf: never reached
    if (reaches(0u)) {
        hi = 0u;
    } else {
        while (hi - lo > 1u) {
            const uint32_t mid = lo + (hi - lo) / 2u;
            if (reaches(mid))
                hi = mid;
            else
                lo = mid;
        }
    }
    thr[idx] = __uint_as_float(hi);
}

template <bool CORR>
__global__ void reduce_partials_kernel(const float* __restrict__ partial,
                                       const float* __restrict__ corr,
                                       const uint8_t* __restrict__ corr_used,
                                       float* __restrict__ volume, VolumeGeom g,
                                       uint32_t num_chunks, double scale,
                                       const int* __restrict__ skip) {
    if (skip && *skip) return;
    // One thread per (tile, slot, lane) entry of a chunk's partial block, so
    // the chunk reads are contiguous; the voxel it holds is written once.
    const size_t per = size_t(g.num_tiles) * g.slots * 64;
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= per) return;
    const uint32_t tile = uint32_t(i / (size_t(g.slots) * 64));
    const uint32_t r = uint32_t(i % (size_t(g.slots) * 64));
    const uint32_t slot = r / 64, lane = r % 64;
    const uint32_t tx = tile % g.tiles[0], ty = (tile / g.tiles[0]) % g.tiles[1];
    const uint32_t tz = tile / (g.tiles[0] * g.tiles[1]);
    // Inverse of lane = (ox & 3) + 4 (oz & 3) + 16 (oy & 3) and
    // slot = ((ox >> 2) lt_z + (oz >> 2)) lt_y + (oy >> 2).
    const uint32_t ox = (lane & 3u) + 4u * (slot / (g.lt[2] * g.lt[1]));
    const uint32_t oz = ((lane >> 2) & 3u) + 4u * ((slot / g.lt[1]) % g.lt[2]);
    const uint32_t oy = (lane >> 4) + 4u * (slot % g.lt[1]);
    const uint32_t vx = tx * g.tile[0] + ox, vy = ty * g.tile[1] + oy, z = tz * g.tile[2] + oz;
    if (vx >= g.res[0] || vy >= g.res[1] || z >= g.res[2]) return;
    // Eight chunks per round, every load of the round issued before the sums
    // (same order of additions: chunk by chunk, partial then corr).
    double sum = 0.0;
    constexpr uint32_t R = 8;
    for (uint32_t c0 = 0; c0 < num_chunks; c0 += R) {
        float p[R], q[R];
        bool use[R];
#pragma unroll
        for (uint32_t u = 0; u < R; ++u) {
            const uint32_t c = min(c0 + u, num_chunks - 1);
            p[u] = partial[size_t(c) * per + i];
            if constexpr (CORR) {
                use[u] = corr_used[size_t(c) * g.num_tiles + tile] != 0;
                q[u] = corr[size_t(c) * per + i];
            }
        }
#pragma unroll
        for (uint32_t u = 0; u < R; ++u) {
            if (c0 + u < num_chunks) {
                sum += double(p[u]);
                if constexpr (CORR)
                    if (use[u]) sum += double(q[u]);
            }
        }
    }
    volume[(size_t(z) * g.res[1] + vy) * g.res[0] + vx] = float(sum * scale);
}

fbp_status pure_threshold_row(fbp_context ctx, uint32_t T, float dt, hipStream_t s,
                              const float** thr) {
    if (!ctx->pure_thr || ctx->pure_thr_T != T || ctx->pure_thr_dt != dt) {
        const size_t bytes = (size_t(T) + 1) * sizeof(float);
        if (bytes > ctx->pure_thr_cap) {
            if (ctx->pure_thr) {
                FBP_CHECK_HIP(ctx, hipStreamSynchronize(s));
                FBP_CHECK_HIP(ctx, hipFree(ctx->pure_thr));
                ctx->pure_thr = nullptr;
            }
            FBP_CHECK_HIP(ctx, hipMalloc(reinterpret_cast<void**>(&ctx->pure_thr), bytes));
            ctx->pure_thr_cap = bytes;
        }
        const fbp_vec3 zero{0.f, 0.f, 0.f};
        hipLaunchKernelGGL(confocal_thresholds_kernel, dim3(unsigned((T + 1 + 255) / 256)), dim3(256),
                           0, s, nullptr, 1u, T, 0.f, dt, 1u, zero, zero, ctx->pure_thr, nullptr);
        FBP_CHECK_HIP(ctx, hipGetLastError());
        ctx->pure_thr_T = T;
        ctx->pure_thr_dt = dt;
    }
    *thr = ctx->pure_thr;
    return FBP_OK;
}

fbp_status launch_confocal_thresholds(fbp_context ctx, const fbp_vec3* laser, uint32_t rows,
                                      uint32_t T, float t0, float dt, uint32_t discard,
                                      fbp_vec3 lpos, fbp_vec3 spos, float* thr, hipStream_t s) {
    const size_t n = size_t(rows) * (T + 1);
    hipLaunchKernelGGL(confocal_thresholds_kernel, dim3(unsigned((n + 255) / 256)), dim3(256), 0,
                       s, laser, rows, T, t0, dt, discard, lpos, spos, thr, nullptr);
    FBP_CHECK_HIP(ctx, hipGetLastError());
    return FBP_OK;
}

fbp_status launch_reduce(fbp_context ctx, const float* partial, const float* corr, float* volume,
                         const VolumeGeom& g, uint32_t num_chunks, double scale,
                         hipStream_t stream, const int* skip, const uint8_t* corr_used) {
    const size_t n = size_t(g.num_tiles) * g.slots * 64;
    const unsigned blocks = unsigned((n + 255) / 256);
    if (corr && !corr_used) return fail(ctx, FBP_ERR_INVALID_ARGUMENT, "reduce: corr without flags");
    if (corr)
        hipLaunchKernelGGL(reduce_partials_kernel<true>, dim3(blocks), dim3(256), 0, stream,
                           partial, corr, corr_used, volume, g, num_chunks, scale, skip);
    else
        hipLaunchKernelGGL(reduce_partials_kernel<false>, dim3(blocks), dim3(256), 0, stream,
                           partial, corr, corr_used, volume, g, num_chunks, scale, skip);
    FBP_CHECK_HIP(ctx, hipGetLastError());
    return FBP_OK;
}

// Picks the number of scan-point chunks so the launch has ~16 waves per CU.
static uint32_t pick_chunks(fbp_context ctx, uint32_t tiles, uint32_t units, uint32_t min_per_chunk) {
    const uint32_t target = 16u * uint32_t(ctx->num_cus > 0 ? ctx->num_cus : 256);
    uint32_t nc = (target + tiles - 1) / tiles;
    const uint32_t cap = (units + min_per_chunk - 1) / min_per_chunk;
    if (nc > cap) nc = cap;
    if (nc < 1) nc = 1;
    return nc;
}

template <int STORAGE, int LAYOUT>
static void launch_confocal(const ConfocalArgs& a, bool pure, dim3 grid, size_t shmem,
                            hipStream_t stream) {
    if constexpr (STORAGE == FBP_STORAGE_F32) {
        if (pure)
            hipLaunchKernelGGL((confocal_voxel_kernel<STORAGE, LAYOUT, true>), grid, dim3(64),
                               shmem, stream, a);
        else
            hipLaunchKernelGGL((confocal_voxel_kernel<STORAGE, LAYOUT, false>), grid, dim3(64),
                               shmem, stream, a);
    } else {
        if (pure)
            hipLaunchKernelGGL((confocal_voxel_kahan_kernel<STORAGE, LAYOUT, true>), grid,
                               dim3(64), shmem, stream, a);
        else
            hipLaunchKernelGGL((confocal_voxel_kahan_kernel<STORAGE, LAYOUT, false>), grid,
                               dim3(64), shmem, stream, a);
    }
}

}  // namespace fbp

using namespace fbp;

namespace fbp {

fbp_status confocal_voxel_impl(fbp_context ctx, const fbp_reconstruction_info* info,
                               const fbp_transient* tr, const fbp_vec3* laser_targets,
                               float* volume, void* stream, const uint32_t* sample_l,
                               const float* sample_w, uint32_t samples) {
    if (!ctx || !info || !tr || !tr->data || !laser_targets || !volume)
        return fail(ctx, FBP_ERR_INVALID_ARGUMENT, "confocal: null argument");
    if (info->num_laser_targets == 0 || info->num_time_bins == 0 || info->voxel_resolution.x == 0 ||
        info->voxel_resolution.y == 0 || info->voxel_resolution.z == 0)
        return fail(ctx, FBP_ERR_INVALID_ARGUMENT, "confocal: zero-sized grid");
    if (!(info->time_step > 0.f) || !finite_info(*info))
        return fail(ctx, FBP_ERR_INVALID_ARGUMENT, "confocal: time_step must be > 0, fields finite");
    if (tr->storage < FBP_STORAGE_F32 || tr->storage > FBP_STORAGE_BF16)
        return fail(ctx, FBP_ERR_INVALID_ARGUMENT, "confocal: bad storage %d", tr->storage);
    if (tr->storage != FBP_STORAGE_F32 && !tr->scales)
        return fail(ctx, FBP_ERR_INVALID_ARGUMENT, "confocal: F16/BF16 storage needs scales");
    if (tr->layout != FBP_LAYOUT_TIME_FASTEST &&
        !(tr->layout == FBP_LAYOUT_TIME_MAJOR && tr->storage == FBP_STORAGE_F32))
        return fail(ctx, FBP_ERR_UNSUPPORTED, "confocal: time-major layout is F32 only");
    FBP_CHECK(bind(ctx));
    if (ctx->voxel_scattering) {
        ctx->path_flag_valid = false;
        return confocal_scatter_impl(ctx, info, tr, laser_targets, volume, as_stream(stream),
                                     sample_l, sample_w, samples);
    }

    ConfocalArgs a{};
    a.data = tr->data;
    a.scales = tr->scales;
    a.laser = laser_targets;
    a.g = make_geom(*info, 4, 4, 4);
    a.L = info->num_laser_targets;
    a.items = samples ? samples : a.L;
    a.sample_l = samples ? sample_l : nullptr;
    a.sample_w = samples ? sample_w : nullptr;
    a.T = info->num_time_bins;
    a.t0 = info->time_offset;
    a.dt = info->time_step;
    a.inv_dt = 1.0f / info->time_step;
    a.discard = info->discard_first_last_bounces;
    a.lpos[0] = info->laser_position.x;
    a.lpos[1] = info->laser_position.y;
    a.lpos[2] = info->laser_position.z;
    a.spos[0] = info->sensor_position.x;
    a.spos[1] = info->sensor_position.y;
    a.spos[2] = info->sensor_position.z;
    a.W = window_bins(a.g, a.dt, 2);
    if (size_t(a.W + 8) * 4 > 65536)
        return fail(ctx, FBP_ERR_UNSUPPORTED,
                    "confocal: bin window of %d bins per tile exceeds LDS (voxels too large for "
                    "the bin width)",
                    a.W);
    // LDS per scan point: W {I, S} pairs + the parameters.
    a.batch = 16;
    while (a.batch > 1 && size_t(a.batch) * (2 * a.W + kPrmWords) * 4 > 20480) a.batch /= 2;
    a.inv_W = 1.0f / float(a.W);

    // Tile-sampled lists are short: smaller chunks keep the CUs busy.
    const uint32_t nc = pick_chunks(ctx, a.g.num_tiles, a.items, samples ? 8 : 64);
    a.chunk_len = (a.items + nc - 1) / nc;
    const uint32_t chunks = (a.items + a.chunk_len - 1) / a.chunk_len;
    const bool pure = (a.t0 == 0.f) && (a.discard != 0);
    a.thr_rows = pure ? 1u : a.L;
    const size_t slab = size_t(chunks) * a.g.num_tiles * a.g.slots * 64;
    const size_t thr_n = size_t(a.thr_rows) * (a.T + 1);
    // Scratch holds the general path's buffers or the aligned path's: only
    // one of the two paths writes, so they share the space.
    AlignedPlan ap = plan_aligned_confocal(ctx, *info, *tr);
    if (samples) ap.usable = false;
    const size_t flags_n = size_t(chunks) * a.g.num_tiles;
    const size_t general_bytes = (2 * slab + thr_n) * sizeof(float) + flags_n + 512;
    const size_t body = std::max(general_bytes, ap.usable ? aligned_scratch_bytes(ap) : size_t(0));
    FBP_CHECK(ensure_scratch(ctx, body));
    if (ap.usable && !ctx->path_flag)
        FBP_CHECK_HIP(ctx, hipMalloc(reinterpret_cast<void**>(&ctx->path_flag), 256));
    int* flag = ctx->path_flag;
    char* base = static_cast<char*>(ctx->scratch);
    a.partial = reinterpret_cast<float*>(base);
    a.corr = a.partial + slab;
    a.thr = a.corr + slab;
    a.corr_used = reinterpret_cast<uint8_t*>(const_cast<float*>(a.thr) + thr_n);
    a.skip = ap.usable ? flag : nullptr;

    hipStream_t s = as_stream(stream);
    ctx->path_flag_valid = ap.usable;
    if (ap.usable)
        FBP_CHECK(launch_aligned_confocal(ctx, ap, *info, *tr, laser_targets, volume, flag, base, s));
    if (pure) {
        FBP_CHECK(pure_threshold_row(ctx, a.T, a.dt, s, &a.thr));
    } else {
        hipLaunchKernelGGL(confocal_thresholds_kernel, dim3(unsigned((thr_n + 255) / 256)),
                           dim3(256), 0, s, laser_targets, a.thr_rows, a.T, a.t0, a.dt, a.discard,
                           info->laser_position, info->sensor_position, const_cast<float*>(a.thr),
                           a.skip);
        FBP_CHECK_HIP(ctx, hipGetLastError());
    }
    const dim3 grid(chunks * a.g.num_tiles);
    const size_t shmem = size_t(a.batch) * (2 * a.W + kPrmWords) * sizeof(float);
    if (tr->layout == FBP_LAYOUT_TIME_MAJOR) {
        launch_confocal<FBP_STORAGE_F32, FBP_LAYOUT_TIME_MAJOR>(a, pure, grid, shmem, s);
    } else if (tr->storage == FBP_STORAGE_F32) {
        launch_confocal<FBP_STORAGE_F32, FBP_LAYOUT_TIME_FASTEST>(a, pure, grid, shmem, s);
    } else if (tr->storage == FBP_STORAGE_F16) {
        launch_confocal<FBP_STORAGE_F16, FBP_LAYOUT_TIME_FASTEST>(a, pure, grid, shmem, s);
    } else {
        launch_confocal<FBP_STORAGE_BF16, FBP_LAYOUT_TIME_FASTEST>(a, pure, grid, shmem, s);
    }
    FBP_CHECK_HIP(ctx, hipGetLastError());
    return launch_reduce(ctx, a.partial, a.corr, volume, a.g, chunks, 1.0, s, a.skip, a.corr_used);
}

}  // namespace fbp

extern "C" fbp_status fbp_backproject_confocal_voxel(fbp_context ctx,
                                                     const fbp_reconstruction_info* info,
                                                     const fbp_transient* tr,
                                                     const fbp_vec3* laser_targets, float* volume,
                                                     void* stream) {
    return confocal_voxel_impl(ctx, info, tr, laser_targets, volume, stream, nullptr, nullptr, 0);
}
