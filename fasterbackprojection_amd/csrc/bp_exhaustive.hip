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
// One wave owns a 16 (vx) x 16 (vy) x 8 (z) tile, 32 voxels per lane, and a
// contiguous chunk of pair indices.  Pairs are staged through LDS in batches:
// the bin window of each pair (I[l][s][kmin .. kmin+W)) is loaded once per
// tile.  Laser distances of the lane's voxels are cached while l is
// unchanged (S consecutive pairs share it); sensor distances come from exact
// per-axis tables.  Bins use the same fast-quotient / exact-fallback scheme
// as the confocal kernel.
#include "bp_common.h"

namespace fbp {

struct ExhaustiveArgs {
    const void* data;
    const float* scales;
    const fbp_vec3* laser;
    const fbp_vec3* sensor;
    float* partial;
    VolumeGeom g;
    uint32_t L, S, T;
    float t0, dt, inv_dt;
    uint32_t discard;
    float lpos[3], spos[3];
    uint64_t num_pairs;
    uint64_t chunk_len;
    int batch;
    int W;
};

enum { Q_SX = 0, Q_SY, Q_SZ, Q_BASE, Q_HME, Q_COFF, Q_NEAR, Q_OFF };

template <bool NEAR>
__device__ __forceinline__ void exhaustive_update_tile(float (&acc)[32], const float (&rl)[32],
                                                       const float (&AL)[4][2], const float (&ZL2)[4],
                                                       const float (&AS)[4][2], const float (&ZS2)[4],
                                                       const float* win, int base, float hme,
                                                       float coff, float off,
                                                       const ExhaustiveArgs& a) {
#pragma unroll
    for (int i = 0; i < 4; ++i) {
#pragma unroll
        for (int j = 0; j < 2; ++j) {
#pragma unroll
            for (int m = 0; m < 4; ++m) {
                const int slot = (i * 2 + j) * 4 + m;
                const float ss = AS[i][j] + ZS2[m];
                const float rs = ss * __builtin_amdgcn_rsqf(ss);
                const float dist = (rl[slot] + rs) + coff;
                const float q = dist * a.inv_dt;
                int k = static_cast<int>(q);
                const float f = __builtin_amdgcn_fractf(q);
                const float d2 = dist * dist;
                float w = __builtin_amdgcn_rcpf(d2);
                bool flag = fabsf(f - 0.5f) > hme;
                if constexpr (NEAR) flag = flag || !(d2 >= 2.0e-5f) || !(ss >= 1.0e-6f);
                if (__builtin_expect(flag, 0)) {
                    const float sl = AL[i][j] + ZL2[m];  // exact (unscaled) tables
                    const float de = exhaustive_dist(sl, ss, a.t0, a.discard, off);
                    k = bin_of(de, a.dt);
                    w = safe_rcp(rn_mul(de, de));
                }
                acc[slot] = fmaf(win[base + k], w, acc[slot]);
            }
        }
    }
}

template <int STORAGE>
__global__ __launch_bounds__(64, 2) void exhaustive_voxel_kernel(ExhaustiveArgs a) {
    extern __shared__ float smem[];
    float* win = smem;
    float* prm = smem + a.batch * a.W;

    const int lane = threadIdx.x;
    const uint32_t tile = blockIdx.x % a.g.num_tiles;
    const uint32_t chunk = blockIdx.x / a.g.num_tiles;
    const uint32_t tx = tile % a.g.tiles[0];
    const uint32_t ty = (tile / a.g.tiles[0]) % a.g.tiles[1];
    const uint32_t tz = tile / (a.g.tiles[0] * a.g.tiles[1]);
    const uint32_t X0 = tx * a.g.tile[0], Y0 = ty * a.g.tile[1], Z0 = tz * a.g.tile[2];
    const int lx = lane & 3, lz = (lane >> 2) & 3, ly = lane >> 4;

    float px[4], py[2], pz[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        px[i] = voxel_center(a.g.hmin[0], a.g.vs[0], X0 + lx + 4 * i);
        pz[i] = voxel_center(a.g.hmin[2], a.g.vs[2], Y0 + ly + 4 * i);
    }
#pragma unroll
    for (int j = 0; j < 2; ++j) py[j] = voxel_center(a.g.hmin[1], a.g.vs[1], Z0 + lz + 4 * j);
    float lo[3], hi[3];
    tile_box(a.g, X0, Y0, Z0, lo, hi);

    float acc[32];
#pragma unroll
    for (int i = 0; i < 32; ++i) acc[i] = 0.f;
    float rl[32];
    float AL[4][2], ZL2[4];
    uint32_t cur_l = 0xffffffffu;

    const uint64_t p_begin = uint64_t(chunk) * a.chunk_len;
    const uint64_t p_end = min(a.num_pairs, p_begin + a.chunk_len);

    for (uint64_t pb = p_begin; pb < p_end; pb += a.batch) {
        const int nb = int(min(uint64_t(a.batch), p_end - pb));
        __syncthreads();
        if (lane < nb) {
            const uint64_t pair = pb + lane;
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
            const int kmin = int(floorf(dmin * a.inv_dt)) - 3;
            const float mag = rlmax + rsmax + fabsf(a.t0) + fabsf(off);
            const float eps = 16.f * 1.1920929e-7f * mag * a.inv_dt + 1.0e-6f;
            float* P = prm + lane * 8;
            P[Q_SX] = sp[0];
            P[Q_SY] = sp[1];
            P[Q_SZ] = sp[2];
            P[Q_BASE] = __int_as_float(lane * a.W - kmin);
            P[Q_HME] = 0.5f - eps;
            P[Q_COFF] = off - a.t0;
            P[Q_NEAR] = ((dmin < 0.0045f && dmax > -0.0045f) || rlmin < 1.0e-3f ||
                         rsmin < 1.0e-3f)
                            ? 1.f
                            : 0.f;
            P[Q_OFF] = off;
        }
        __syncthreads();
        for (int jl = 0; jl < nb; ++jl) {
            const uint64_t pair = pb + jl;
            const int kmin = jl * a.W - __float_as_int(prm[jl * 8 + Q_BASE]);
            float sc = 1.f;
            if constexpr (STORAGE != FBP_STORAGE_F32) sc = a.scales[pair];
            for (int w = lane; w < a.W; w += 64) {
                const int k = kmin + w;
                float v = 0.f;
                if (k >= 0 && k < int(a.T)) {
                    v = load_sample<STORAGE>(a.data, pair * a.T + uint32_t(k), sc);
                    if constexpr (STORAGE == FBP_STORAGE_F32) v = (v > kEps) ? v : 0.f;
                }
                win[jl * a.W + w] = v;
            }
        }
        __syncthreads();
        for (int jl = 0; jl < nb; ++jl) {
            const uint64_t pair = pb + jl;
            const uint32_t l = rfl(int(uint32_t(pair / a.S)));
            if (l != cur_l) {
                cur_l = l;
                const float lxp = a.laser[l].x, lyp = a.laser[l].y, lzp = a.laser[l].z;
                float XL2[4], YL2[2];
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    XL2[i] = sq(rn_sub(px[i], lxp));
                    ZL2[i] = sq(rn_sub(pz[i], lzp));
                }
#pragma unroll
                for (int j = 0; j < 2; ++j) YL2[j] = sq(rn_sub(py[j], lyp));
#pragma unroll
                for (int i = 0; i < 4; ++i)
#pragma unroll
                    for (int j = 0; j < 2; ++j) AL[i][j] = rn_add(XL2[i], YL2[j]);
#pragma unroll
                for (int i = 0; i < 4; ++i)
#pragma unroll
                    for (int j = 0; j < 2; ++j)
#pragma unroll
                        for (int m = 0; m < 4; ++m) {
                            const float sl = AL[i][j] + ZL2[m];  // exact (unscaled) tables
                            rl[(i * 2 + j) * 4 + m] = sl * __builtin_amdgcn_rsqf(sl);
                        }
            }
            const float* P = prm + jl * 8;
            const float sxp = rfl(P[Q_SX]), syp = rfl(P[Q_SY]), szp = rfl(P[Q_SZ]);
            const int base = rfl(__float_as_int(P[Q_BASE]));
            const float hme = rfl(P[Q_HME]);
            const float coff = rfl(P[Q_COFF]);
            const bool near = rfl(P[Q_NEAR]) != 0.f;
            const float off = rfl(P[Q_OFF]);
            float XS2[4], YS2[2], ZS2[4], AS[4][2];
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                XS2[i] = sq(rn_sub(px[i], sxp));
                ZS2[i] = sq(rn_sub(pz[i], szp));
            }
#pragma unroll
            for (int j = 0; j < 2; ++j) YS2[j] = sq(rn_sub(py[j], syp));
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 2; ++j) AS[i][j] = rn_add(XS2[i], YS2[j]);
            if (near)
                exhaustive_update_tile<true>(acc, rl, AL, ZL2, AS, ZS2, win, base, hme, coff, off, a);
            else
                exhaustive_update_tile<false>(acc, rl, AL, ZL2, AS, ZS2, win, base, hme, coff, off, a);
        }
    }

    float* out = a.partial + (size_t(chunk) * a.g.num_tiles + tile) * 32 * 64;
#pragma unroll
    for (int s = 0; s < 32; ++s) out[s * 64 + lane] = acc[s];
}

}  // namespace fbp

using namespace fbp;

extern "C" fbp_status fbp_backproject_exhaustive_voxel(fbp_context ctx,
                                                       const fbp_reconstruction_info* info,
                                                       const fbp_transient* tr,
                                                       const fbp_vec3* laser_targets,
                                                       const fbp_vec3* sensor_targets,
                                                       float* volume, void* stream) {
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
    a.g = make_geom(*info, 4, 4, 2);
    a.L = info->num_laser_targets;
    a.S = info->num_sensor_targets;
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
    a.num_pairs = uint64_t(a.L) * a.S;
    a.W = window_bins(a.g, a.dt, 2);
    if (size_t(a.W + 8) * 4 > 65536)
        return fail(ctx, FBP_ERR_UNSUPPORTED, "exhaustive: bin window of %d bins exceeds LDS", a.W);
    a.batch = 16;
    while (a.batch > 1 && size_t(a.batch) * (a.W + 8) * 4 > 32768) a.batch /= 2;

    const uint32_t target = 16u * uint32_t(ctx->num_cus > 0 ? ctx->num_cus : 256);
    uint64_t nc = (target + a.g.num_tiles - 1) / a.g.num_tiles;
    const uint64_t cap = (a.num_pairs + 63) / 64;
    if (nc > cap) nc = cap;
    if (nc < 1) nc = 1;
    a.chunk_len = (a.num_pairs + nc - 1) / nc;
    const uint64_t chunks = (a.num_pairs + a.chunk_len - 1) / a.chunk_len;
    const size_t partial_bytes = size_t(chunks) * a.g.num_tiles * a.g.slots * 64 * sizeof(float);
    FBP_CHECK(ensure_scratch(ctx, partial_bytes));
    a.partial = static_cast<float*>(ctx->scratch);

    const dim3 grid(unsigned(chunks * a.g.num_tiles));
    const size_t shmem = size_t(a.batch) * (a.W + 8) * sizeof(float);
    hipStream_t s = as_stream(stream);
    if (tr->storage == FBP_STORAGE_F32)
        hipLaunchKernelGGL(exhaustive_voxel_kernel<FBP_STORAGE_F32>, grid, dim3(64), shmem, s, a);
    else if (tr->storage == FBP_STORAGE_F16)
        hipLaunchKernelGGL(exhaustive_voxel_kernel<FBP_STORAGE_F16>, grid, dim3(64), shmem, s, a);
    else
        hipLaunchKernelGGL(exhaustive_voxel_kernel<FBP_STORAGE_BF16>, grid, dim3(64), shmem, s, a);
    FBP_CHECK_HIP(ctx, hipGetLastError());
    return launch_reduce(ctx, a.partial, nullptr, volume, a.g, uint32_t(chunks), 1.0, s);
}
