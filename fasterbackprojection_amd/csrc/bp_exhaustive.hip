// Exhaustive (non-confocal) voxel backprojection: time of flight
// |l - v| + |v - s| for every (laser, sensor) pair.
//
// Reference: backprojectExhaustiveVoxel (FB/transient_reconstruction.cuh:138-166)
// + Laser::reconstructAABBExhaustive (FB/Laser.cu:500-516).  The reference
// splits ls as l = ls / L, s = ls % L and launches with the grid axes swapped,
// so only ls < ceil(XY/16)*16 pairs are visited and L != S reads out of
// bounds; here every pair is visited with l = ls / S, s = ls % S (identical
// indexing when L == S).
//
// A workgroup of four waves owns a 16 (vx) x 16 (vy) x 32 (z) tile (each
// wave a z-slab of 8, 32 voxels per lane) and a contiguous chunk of pair
// indices.  Pairs are staged through LDS in double-buffered batches: the bin
// window of each pair (I[l][s][kmin .. kmin+W)) is loaded once per tile,
// the next batch's windows while the current one is processed.  Laser distances of the lane's voxels are cached while l is
// unchanged (S consecutive pairs share it) in bin units (rl / dt).  The
// fast estimate of the bin quotient is q = rl/dt + sqrt(ss/dt^2) with the
// squared sensor distance from per-pair axis tables scaled by 1/dt^2, and
// the weight is 1/q^2 (dt^2 / dist^2; the reduction multiplies by 1/dt^2 in
// FP64).  Voxels whose q lies within the estimate's error bound of a bin
// boundary add nothing in the fast loop (a select, no branch); after the
// pair's 32 updates, the lanes that flagged any take the reference's IEEE
// chain (bin and weight) for those voxels into per-thread LDS correction
// slots, behind one wave-uniform test.  Pairs whose tile may come within a
// few millimetres of the wall or of zero distance take the whole exact chain.
#include <mutex>

#include "bp_common.h"

namespace fbp {

// Partial sums of fbp_backproject_exhaustive_voxel_noalloc: one array per
// GPU (each device holds its own copy of this module's variable until the
// process ends), shared by every no-alloc call on that device.  Host code
// never uses the name itself for device memory: its per-device address comes
// from hipGetSymbolAddress (noalloc_scratch_address below).
static __device__ __align__(256) float scratch[FBP_NOALLOC_SCRATCH_BYTES / sizeof(float)];

struct ExhaustiveArgs {
    const void* data;
    const float* scales;
    const fbp_vec3* laser;
    const fbp_vec3* sensor;
    float* partial;
    VolumeGeom g;
    uint32_t L, S, T;
    float t0, dt, inv_dt, inv_dt2, dt2;
    uint32_t discard;
    float lpos[3], spos[3];
    uint64_t num_pairs;
    uint64_t chunk_len;
    int batch;
    int W;
    float inv_W;
};

// Laser position of the cached laser distances (the exact fallback
// recomputes the exact squared laser distance from it).
struct LaserPos {
    float x, y, z;
};

// Exact IEEE chain (reference order) for the voxel (x, y, z) and sensor s:
// the squared laser and sensor distances and the time-of-flight distance.
__device__ __forceinline__ float exact_dist(float x, float y, float z, LaserPos lp, float sx,
                                            float sy, float sz, const ExhaustiveArgs& a,
                                            float off) {
    const float sl = dot3(sq(rn_sub(x, lp.x)), sq(rn_sub(y, lp.y)), sq(rn_sub(z, lp.z)));
    const float ss = dot3(sq(rn_sub(x, sx)), sq(rn_sub(y, sy)), sq(rn_sub(z, sz)));
    return exhaustive_dist(sl, ss, a.t0, a.discard, off);
}

// One pair applied to the lane's 32 voxels.  rq: laser distances / dt;
// AQ, ZQ: the pair's per-axis squared sensor distances (x+y, z) / dt^2.
// NEAR: every voxel takes the whole exact chain (bin and weight).
__device__ __forceinline__ float pick4f(const float (&v)[4], int i) {
    return i == 0 ? v[0] : i == 1 ? v[1] : i == 2 ? v[2] : v[3];
}

template <bool NEAR, bool COFF>
__device__ __forceinline__ void exhaustive_update_tile(float (&acc)[32], const float (&rq)[32],
                                                       const float (&px)[4], const float (&py)[2],
                                                       const float (&pz)[4], LaserPos lp,
                                                       const float (&AQ)[4][2], const float (&ZQ)[4],
                                                       const float* win,
                                                       int base, float hme, float coffq,
                                                       float off, float sx, float sy, float sz,
                                                       const ExhaustiveArgs& a, float* corr) {
    uint32_t band_mask = 0;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
#pragma unroll
        for (int j = 0; j < 2; ++j) {
#pragma unroll
            for (int m = 0; m < 4; ++m) {
                const int slot = (i * 2 + j) * 4 + m;
                const float s2 = AQ[i][j] + ZQ[m];
                float q = fmaf(s2, __builtin_amdgcn_rsqf(s2), rq[slot]);
                if constexpr (COFF) q = q + coffq;
                int k = static_cast<int>(q);
                float w = __builtin_amdgcn_rcpf(q * q);
                if constexpr (NEAR) {
                    const float de = exact_dist(px[i], py[j], pz[m], lp, sx, sy, sz, a, off);
                    k = bin_of(de, a.dt);
                    w = safe_rcp(rn_mul(de, de)) * a.dt2;
                    acc[slot] = fmaf(win[base + k], w, acc[slot]);
                } else {
                    // q within the error bound of a bin boundary: left out
                    // here, added exactly below.
                    const float f = __builtin_amdgcn_fractf(q);
                    const bool band = fabsf(f - 0.5f) > hme;
                    band_mask |= uint32_t(band) << slot;
                    const float v = win[base + k];
                    acc[slot] = fmaf(band ? 0.f : v, w, acc[slot]);
                }
            }
        }
    }
    if constexpr (!NEAR) {
        if (__any(band_mask != 0)) {
            // The voxels next to a boundary take the exact chain (bin and
            // weight) into the lane's correction slots.
            while (band_mask) {
                const int slot = __builtin_ctz(band_mask);
                band_mask &= band_mask - 1u;
                const int i = slot >> 3, j = (slot >> 2) & 1, m = slot & 3;
                const float de = exact_dist(pick4f(px, i), j ? py[1] : py[0], pick4f(pz, m), lp,
                                            sx, sy, sz, a, off);
                const int k = bin_of(de, a.dt);
                if (k >= 0 && k < int(a.T))
                    corr[slot * 256] += win[base + k] * (safe_rcp(rn_mul(de, de)) * a.dt2);
            }
        }
    }
}

// Window prefetch registers per thread: batch * W <= kPrefetch * 256.
constexpr int kPrefetch = 16;
constexpr int kWaves = 4;

// Pair parameters of one batch slot (computed by one thread, read by all).
struct PairParams {
    float sx, sy, sz;
    int base;     // slot * W - kmin: LDS index of bin k is base + k
    float hme;    // 0.5 - fast-path tolerance (in bins)
    float coffq;  // (off - t0) / dt
    float near;   // 1: every voxel of the tile takes the exact chain
    float off;
    int kmin;
};

template <int STORAGE>
__device__ __forceinline__ PairParams pair_params(const ExhaustiveArgs& a, uint64_t pair, int slot,
                                                  const float (&lo)[3], const float (&hi)[3]) {
    const uint32_t l = uint32_t(pair / a.S), s = uint32_t(pair % a.S);
    const float lp[3] = {a.laser[l].x, a.laser[l].y, a.laser[l].z};
    const float sp[3] = {a.sensor[s].x, a.sensor[s].y, a.sensor[s].z};
    float off = 0.f;
    if (a.discard == 0)
        off = rn_add(distance3(lp[0], lp[1], lp[2], a.lpos[0], a.lpos[1], a.lpos[2]),
                     distance3(sp[0], sp[1], sp[2], a.spos[0], a.spos[1], a.spos[2]));
    const float rlmin = box_min_dist(lp, lo, hi), rlmax = box_max_dist(lp, lo, hi);
    const float rsmin = box_min_dist(sp, lo, hi), rsmax = box_max_dist(sp, lo, hi);
    const float dmin = rlmin + rsmin - a.t0 + off;
    const float dmax = rlmax + rsmax - a.t0 + off;
    PairParams P;
    P.kmin = int(floorf(dmin * a.inv_dt)) - 3;
    // Error bound of the fast quotient against fl(dist / dt), in bins: with
    // 1-ulp v_rsq the pure chain stays within 4.75 * 2^-23 * q and the
    // offset terms add 3 * 2^-23 * mag / dt; 10 * 2^-23 * mag / dt covers both.
    const float mag = rlmax + rsmax + fabsf(a.t0) + fabsf(off);
    const float eps = 10.f * 1.1920929e-7f * mag * a.inv_dt + 1.0e-6f;
    P.sx = sp[0];
    P.sy = sp[1];
    P.sz = sp[2];
    P.base = slot * a.W - P.kmin;
    P.hme = 0.5f - eps;
    P.coffq = (off - a.t0) * a.inv_dt;
    P.near = ((dmin < 0.0045f && dmax > -0.0045f) || rlmin < 1.0e-3f || rsmin < 1.0e-3f) ? 1.f : 0.f;
    P.off = off;
    return P;
}

// Workgroup: kWaves waves own one 16 (vx) x 16 (vy) x 32 (z) tile; wave w
// holds the z-slab 8w .. 8w+7 (32 voxels per lane) and all waves share the
// staged bin windows of a batch of (laser, sensor) pairs.  Batches are
// double-buffered: the windows of batch b+1 are fetched into registers while
// batch b is processed.
// COFF: the chain has a time offset or per-pair path offsets (q needs the
// pair's (off - t0) / dt).
template <int STORAGE, bool COFF>
__global__ __launch_bounds__(256, 2) void exhaustive_voxel_kernel(ExhaustiveArgs a) {
    extern __shared__ float smem[];
    // [2][batch * W] windows, [2][batch] PairParams, then the correction
    // slots of the boundary voxels [32][256 threads] (thread-fastest: no
    // bank conflicts).
    float* winbuf = smem;
    PairParams* prmbuf = reinterpret_cast<PairParams*>(smem + 2 * a.batch * a.W);
    float* corrl = reinterpret_cast<float*>(prmbuf + 2 * a.batch) + threadIdx.x;

    const int tid = threadIdx.x;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int lane = tid & 63;
    const uint32_t tile = blockIdx.x % a.g.num_tiles;
    const uint32_t chunk = blockIdx.x / a.g.num_tiles;
    const uint32_t tx = tile % a.g.tiles[0];
    const uint32_t ty = (tile / a.g.tiles[0]) % a.g.tiles[1];
    const uint32_t tz = tile / (a.g.tiles[0] * a.g.tiles[1]);
    const uint32_t X0 = tx * a.g.tile[0], Y0 = ty * a.g.tile[1], Z0 = tz * a.g.tile[2];
    const int lx = lane & 3, lz = (lane >> 2) & 3, ly = lane >> 4;
    const uint32_t Zw = Z0 + 8u * uint32_t(wave);

    float px[4], py[2], pz[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        px[i] = voxel_center(a.g.hmin[0], a.g.vs[0], X0 + lx + 4 * i);
        pz[i] = voxel_center(a.g.hmin[2], a.g.vs[2], Y0 + ly + 4 * i);
    }
#pragma unroll
    for (int j = 0; j < 2; ++j) py[j] = voxel_center(a.g.hmin[1], a.g.vs[1], Zw + lz + 4 * j);
    float lo[3], hi[3];
    tile_box(a.g, X0, Y0, Z0, lo, hi);

    float acc[32];
#pragma unroll
    for (int i = 0; i < 32; ++i) {
        acc[i] = 0.f;
        corrl[i * 256] = 0.f;
    }
    float rq[32];
    LaserPos lp{0.f, 0.f, 0.f};
    uint32_t cur_l = 0xffffffffu;

    const uint64_t p_begin = uint64_t(chunk) * a.chunk_len;
    const uint64_t p_end = min(a.num_pairs, p_begin + a.chunk_len);
    const int per = a.batch * a.W;

    // Every register's load is issued (unused ones read element 0), so the
    // loads of a batch are in flight together and waited for once, at the
    // store.
    auto fetch = [&](uint64_t pb, int nb, const PairParams* prm, float (&reg)[kPrefetch]) {
#pragma unroll
        for (int i = 0; i < kPrefetch; ++i) {
            const int f = tid + i * 256;
            const bool in = f < nb * a.W;
            // f / W through the reciprocal: exact for f < 2^22.
            const int jl = in ? int((float(f) + 0.5f) * a.inv_W) : 0;
            const int w = f - jl * a.W;
            const uint64_t pair = pb + jl;
            const int k = prm[jl].kmin + w;
            const bool ok = in && k >= 0 && k < int(a.T);
            float sc = 1.f;
            if constexpr (STORAGE != FBP_STORAGE_F32) sc = a.scales[ok ? pair : 0];
            float v = load_sample<STORAGE>(a.data, ok ? pair * a.T + uint32_t(k) : 0, sc);
            if constexpr (STORAGE == FBP_STORAGE_F32) v = (v > kEps) ? v : 0.f;
            reg[i] = ok ? v : 0.f;
        }
    };
    auto store = [&](float* win, int nb, const float (&reg)[kPrefetch]) {
#pragma unroll
        for (int i = 0; i < kPrefetch; ++i) {
            const int f = tid + i * 256;
            if (f < nb * a.W) win[f] = reg[i];
        }
    };

    float reg[kPrefetch];
    // Prologue: batch 0 into buffer 0.
    if (p_begin < p_end) {
        const int nb = int(min(uint64_t(a.batch), p_end - p_begin));
        if (tid < nb) prmbuf[tid] = pair_params<STORAGE>(a, p_begin + tid, tid, lo, hi);
        __syncthreads();
        fetch(p_begin, nb, prmbuf, reg);
        store(winbuf, nb, reg);
        __syncthreads();
    }
    int cur = 0;
    for (uint64_t pb = p_begin; pb < p_end; pb += a.batch) {
        const int nb = int(min(uint64_t(a.batch), p_end - pb));
        const uint64_t pn = pb + a.batch;
        const bool has_next = pn < p_end;
        const int nbn = has_next ? int(min(uint64_t(a.batch), p_end - pn)) : 0;
        float* win = winbuf + cur * per;
        const PairParams* prm = prmbuf + cur * a.batch;
        PairParams* prm_next = prmbuf + (cur ^ 1) * a.batch;
        if (has_next) {
            if (tid < nbn) prm_next[tid] = pair_params<STORAGE>(a, pn + tid, tid, lo, hi);
            __syncthreads();
            fetch(pn, nbn, prm_next, reg);
        }
        for (int jl = 0; jl < nb; ++jl) {
            const uint64_t pair = pb + jl;
            const uint32_t l = rfl(int(uint32_t(pair / a.S)));
            if (l != cur_l) {
                cur_l = l;
                lp = LaserPos{rfl(a.laser[l].x), rfl(a.laser[l].y), rfl(a.laser[l].z)};
                float XL2[4], YL2[2], ZL2[4];
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    XL2[i] = sq(rn_sub(px[i], lp.x));
                    ZL2[i] = sq(rn_sub(pz[i], lp.z));
                }
#pragma unroll
                for (int j = 0; j < 2; ++j) YL2[j] = sq(rn_sub(py[j], lp.y));
#pragma unroll
                for (int i = 0; i < 4; ++i)
#pragma unroll
                    for (int j = 0; j < 2; ++j)
#pragma unroll
                        for (int m = 0; m < 4; ++m) {
                            const float sl = dot3(XL2[i], YL2[j], ZL2[m]);
                            rq[(i * 2 + j) * 4 + m] =
                                sl * __builtin_amdgcn_rsqf(sl) * a.inv_dt;
                        }
            }
            const PairParams& P = prm[jl];
            const float sxp = rfl(P.sx), syp = rfl(P.sy), szp = rfl(P.sz);
            const int base = rfl(P.base) + cur * per;
            const float hme = rfl(P.hme);
            const float coffq = rfl(P.coffq);
            const bool near = rfl(P.near) != 0.f;
            const float off = rfl(P.off);
            // Per-axis squared sensor distances / dt^2 (x + y, z).
            float XS2[4], YS2[2], ZQ[4], AQ[4][2];
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                XS2[i] = sq(rn_sub(px[i], sxp));
                ZQ[i] = sq(rn_sub(pz[i], szp)) * a.inv_dt2;
            }
#pragma unroll
            for (int j = 0; j < 2; ++j) YS2[j] = sq(rn_sub(py[j], syp));
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 2; ++j) AQ[i][j] = (XS2[i] + YS2[j]) * a.inv_dt2;
            if (near)
                exhaustive_update_tile<true, COFF>(acc, rq, px, py, pz, lp, AQ, ZQ, winbuf, base,
                                                   hme, coffq, off, sxp, syp, szp, a, corrl);
            else
                exhaustive_update_tile<false, COFF>(acc, rq, px, py, pz, lp, AQ, ZQ, winbuf, base,
                                                    hme, coffq, off, sxp, syp, szp, a, corrl);
        }
        if (has_next) store(winbuf + (cur ^ 1) * per, nbn, reg);
        __syncthreads();
        cur ^= 1;
    }

    // Slots of the workgroup tile (lt = 4 x 4 x 8): wave w's z-slab holds
    // oz >> 2 = 2w + j.
    float* out = a.partial + (size_t(chunk) * a.g.num_tiles + tile) * a.g.slots * 64;
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int m = 0; m < 4; ++m)
                out[size_t((i * 8 + 2 * wave + j) * 4 + m) * 64 + lane] =
                    acc[(i * 2 + j) * 4 + m] + corrl[((i * 2 + j) * 4 + m) * 256];
}

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_noalloc_mu;
void* g_noalloc_ptr[kMaxDevices] = {};
}  // namespace

// Device address of `scratch` on the current device (ctx->device), looked up
// once per device and kept; called by fbp_context_create with the device
// current.  The module variable exists separately on every device, so the
// address is per device.
fbp_status noalloc_scratch_address(fbp_context ctx, void** out) {
    *out = nullptr;
    if (ctx->device < 0 || ctx->device >= kMaxDevices)
        return fail(ctx, FBP_ERR_UNSUPPORTED, "no-alloc scratch: device %d beyond %d", ctx->device,
                    kMaxDevices);
    std::lock_guard<std::mutex> lock(g_noalloc_mu);
    if (!g_noalloc_ptr[ctx->device]) {
        void* p = nullptr;
        FBP_CHECK_HIP(ctx, hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        g_noalloc_ptr[ctx->device] = p;
    }
    *out = g_noalloc_ptr[ctx->device];
    return FBP_OK;
}

// Shared body of the two exhaustive entry points.  fixed_partial == nullptr:
// partial sums in the context's scratch arena (grown on demand).  Otherwise
// they go to fixed_partial (fixed_floats floats) and the number of pair
// chunks is capped so they fit; nothing is allocated.
static fbp_status exhaustive_voxel_impl(fbp_context ctx, const fbp_reconstruction_info* info,
                                        const fbp_transient* tr, const fbp_vec3* laser_targets,
                                        const fbp_vec3* sensor_targets, float* volume,
                                        void* stream, float* fixed_partial, size_t fixed_floats) {
    if (!ctx || !info || !tr || !tr->data || !laser_targets || !sensor_targets || !volume)
        return fail(ctx, FBP_ERR_INVALID_ARGUMENT, "exhaustive: null argument");
    if (info->num_laser_targets == 0 || info->num_sensor_targets == 0 ||
        info->num_time_bins == 0 || info->voxel_resolution.x == 0 ||
        info->voxel_resolution.y == 0 || info->voxel_resolution.z == 0)
        return fail(ctx, FBP_ERR_INVALID_ARGUMENT, "exhaustive: zero-sized grid");
    if (!(info->time_step > 0.f) || !finite_info(*info))
        return fail(ctx, FBP_ERR_INVALID_ARGUMENT, "exhaustive: time_step must be > 0, fields finite");
    if (tr->storage < FBP_STORAGE_F32 || tr->storage > FBP_STORAGE_BF16)
        return fail(ctx, FBP_ERR_INVALID_ARGUMENT, "exhaustive: bad storage %d", tr->storage);
    if (tr->storage != FBP_STORAGE_F32 && !tr->scales)
        return fail(ctx, FBP_ERR_INVALID_ARGUMENT, "exhaustive: F16/BF16 storage needs scales");
    if (tr->layout != FBP_LAYOUT_TIME_FASTEST)
        return fail(ctx, FBP_ERR_UNSUPPORTED, "exhaustive: transient must be [l][s][t]");
    FBP_CHECK(bind(ctx));

    ExhaustiveArgs a{};
    a.data = tr->data;
    a.scales = tr->scales;
    a.laser = laser_targets;
    a.sensor = sensor_targets;
    // Workgroup tiles of 16 x 16 x 32 voxels (four waves of 16 x 16 x 8).
    a.g = make_geom(*info, 4, 4, 2 * kWaves);
    a.L = info->num_laser_targets;
    a.S = info->num_sensor_targets;
    a.T = info->num_time_bins;
    a.t0 = info->time_offset;
    a.dt = info->time_step;
    a.inv_dt = 1.0f / info->time_step;
    a.inv_dt2 = a.inv_dt * a.inv_dt;
    a.dt2 = info->time_step * info->time_step;
    a.discard = info->discard_first_last_bounces;
    a.lpos[0] = info->laser_position.x;
    a.lpos[1] = info->laser_position.y;
    a.lpos[2] = info->laser_position.z;
    a.spos[0] = info->sensor_position.x;
    a.spos[1] = info->sensor_position.y;
    a.spos[2] = info->sensor_position.z;
    a.num_pairs = uint64_t(a.L) * a.S;
    a.W = window_bins(a.g, a.dt, 2);
    if (a.W > kPrefetch * 256)
        return fail(ctx, FBP_ERR_UNSUPPORTED, "exhaustive: bin window of %d bins per tile too long "
                    "(voxels too large for the bin width)", a.W);
    // Pairs per batch: the window prefetch registers hold one batch.
    a.batch = 16;
    while (a.batch > 1 && a.batch * a.W > kPrefetch * 256) a.batch /= 2;
    a.inv_W = 1.0f / float(a.W);
    // ~8 workgroups (32 waves) per CU over the launch.
    const uint32_t target = 8u * uint32_t(ctx->num_cus > 0 ? ctx->num_cus : 256);
    uint64_t nc = (target + a.g.num_tiles - 1) / a.g.num_tiles;
    const uint64_t cap = (a.num_pairs + 63) / 64;
    if (nc > cap) nc = cap;
    const size_t chunk_floats = size_t(a.g.num_tiles) * a.g.slots * 64;
    if (fixed_partial) {
        const uint64_t fit = fixed_floats / chunk_floats;
        if (fit == 0)
            return fail(ctx, FBP_ERR_UNSUPPORTED,
                        "exhaustive no-alloc: %zu voxels (tile-padded) need %zu bytes of partial "
                        "sums, beyond the %zu-byte scratch",
                        chunk_floats, chunk_floats * sizeof(float), fixed_floats * sizeof(float));
        if (nc > fit) nc = fit;
    }
    if (nc < 1) nc = 1;
    a.chunk_len = (a.num_pairs + nc - 1) / nc;
    const uint64_t chunks = (a.num_pairs + a.chunk_len - 1) / a.chunk_len;
    if (fixed_partial) {
        a.partial = fixed_partial;
    } else {
        FBP_CHECK(ensure_scratch(ctx, size_t(chunks) * chunk_floats * sizeof(float)));
        a.partial = static_cast<float*>(ctx->scratch);
    }
    hipStream_t s = as_stream(stream);

    const dim3 grid(unsigned(chunks * a.g.num_tiles));
    const size_t shmem = 2 * (size_t(a.batch) * a.W * sizeof(float) + a.batch * sizeof(PairParams)) +
                         size_t(32) * 64 * kWaves * sizeof(float);
    // At most 2 * 16 KiB of windows + 32 KiB of correction slots (+ pair
    // parameters): beyond the default 64 KiB for the longest windows.
    if (shmem > size_t(ctx->max_lds_per_block))
        return fail(ctx, FBP_ERR_UNSUPPORTED, "exhaustive: %zu B of LDS per workgroup needed", shmem);
    const bool coff = !(a.t0 == 0.f && a.discard != 0);
    void (*kern)(ExhaustiveArgs) =
        tr->storage == FBP_STORAGE_F32
            ? (coff ? exhaustive_voxel_kernel<FBP_STORAGE_F32, true>
                    : exhaustive_voxel_kernel<FBP_STORAGE_F32, false>)
        : tr->storage == FBP_STORAGE_F16
            ? (coff ? exhaustive_voxel_kernel<FBP_STORAGE_F16, true>
                    : exhaustive_voxel_kernel<FBP_STORAGE_F16, false>)
            : (coff ? exhaustive_voxel_kernel<FBP_STORAGE_BF16, true>
                    : exhaustive_voxel_kernel<FBP_STORAGE_BF16, false>);
    FBP_CHECK_HIP(ctx, hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                                           hipFuncAttributeMaxDynamicSharedMemorySize, int(shmem)));
    hipLaunchKernelGGL(kern, grid, dim3(64 * kWaves), shmem, s, a);
    FBP_CHECK_HIP(ctx, hipGetLastError());
    // The fast weights carry dt^2 (1 / q^2): scale by 1 / dt^2 in FP64.
    return launch_reduce(ctx, a.partial, nullptr, volume, a.g, uint32_t(chunks),
                         1.0 / (double(a.dt) * double(a.dt)), s);
}

}  // namespace fbp

using namespace fbp;

extern "C" fbp_status fbp_backproject_exhaustive_voxel(fbp_context ctx,
                                                       const fbp_reconstruction_info* info,
                                                       const fbp_transient* tr,
                                                       const fbp_vec3* laser_targets,
                                                       const fbp_vec3* sensor_targets,
                                                       float* volume, void* stream) {
    return exhaustive_voxel_impl(ctx, info, tr, laser_targets, sensor_targets, volume, stream,
                                 nullptr, 0);
}

extern "C" fbp_status fbp_backproject_exhaustive_voxel_noalloc(fbp_context ctx,
                                                               const fbp_reconstruction_info* info,
                                                               const fbp_transient* tr,
                                                               const fbp_vec3* laser_targets,
                                                               const fbp_vec3* sensor_targets,
                                                               float* volume, void* stream) {
    if (!ctx) return FBP_ERR_INVALID_ARGUMENT;
    if (!ctx->noalloc_scratch)
        return fail(ctx, FBP_ERR_DEVICE, "exhaustive no-alloc: the scratch address of device %d "
                    "could not be resolved at context creation", ctx->device);
    // Every partial-sum slot the reduction reads is written by this call's
    // kernel first, so the array needs no clearing between calls.
    return exhaustive_voxel_impl(ctx, info, tr, laser_targets, sensor_targets, volume, stream,
                                 static_cast<float*>(ctx->noalloc_scratch),
                                 FBP_NOALLOC_SCRATCH_BYTES / sizeof(float));
}
