// Confocal voxel backprojection, voxel-scattering variant.
//
// Reference: the ACCUMULATE_VOXEL_SCATTERING branch of backprojectConfocalVoxel
// (FB/transient_reconstruction.cuh:112-126, a compile-time switch there; a
// context option here).  Every (voxel, scan point) pair adds the samples of
// all bins its voxel can reach: the walk d = max(0, dist - e),
// d += (Delta + EPS) while d < dist + e, with e half the voxel diagonal,
// each sample weighted by safeRCP(dist^2) of the voxel centre.  The walk and
// the distance are the reference's IEEE chain (scatter_walk in chain.h), so
// the visited bins are exactly the reference's.
//
// One thread per voxel sums every scan point in order with a Kahan-compensated
// FP32 sum (deterministic; no partial buffers).  The samples of one scan point
// seen by the 64 neighbouring voxels of a wave lie a few bins apart, so the
// loads of a wave share cache lines.  The tile-sampled MIS mode passes
// per-16^3-tile scan-point lists and weights, as for the general kernel.
#include "bp_common.h"

namespace fbp {

namespace {

struct ScatterArgs {
    const void* data;
    const float* scales;
    const fbp_vec3* laser;
    float* volume;
    uint32_t X, Y, Z, L, T;
    float hmin[3], vs[3];
    float t0, dt, ext;
    uint32_t discard;
    fbp_vec3 lpos, spos;
    // Tile-sampled mode: tile (vx/16, vy/16, z/16), x fastest, visits
    // sample_l[tile][0..items) with weights sample_w; else every scan point.
    uint32_t items;
    const uint32_t* sample_l;
    const float* sample_w;
    uint32_t tiles_x, tiles_y;
};

template <int STORAGE, int LAYOUT>
__global__ __launch_bounds__(256) void confocal_scatter_kernel(ScatterArgs a) {
    const size_t vid = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    const size_t slice = size_t(a.X) * a.Y;
    if (vid >= slice * a.Z) return;
    const uint32_t vx = uint32_t(vid % a.X), vy = uint32_t((vid / a.X) % a.Y),
                   z = uint32_t(vid / slice);
    const float px = voxel_center(a.hmin[0], a.vs[0], vx);
    const float py = voxel_center(a.hmin[1], a.vs[1], z);
    const float pz = voxel_center(a.hmin[2], a.vs[2], vy);
    const bool sampled = a.sample_l != nullptr;
    const size_t list = sampled ? (size_t(vx / 16) + size_t(vy / 16) * a.tiles_x +
                                   size_t(z / 16) * a.tiles_x * a.tiles_y) * a.items
                                : 0;
    const uint32_t n = sampled ? a.items : a.L;
    float sum = 0.f, comp = 0.f;
    for (uint32_t i = 0; i < n; ++i) {
        const uint32_t l = sampled ? a.sample_l[list + i] : i;
        const float sw = sampled ? a.sample_w[list + i] : 1.f;
        const fbp_vec3 lp = a.laser[l];
        const float r = distance3(lp.x, lp.y, lp.z, px, py, pz);
        float dist = rn_sub(rn_mul(r, 2.0f), a.t0);
        if (a.discard == 0)
            dist = rn_add(dist, rn_add(distance3(lp.x, lp.y, lp.z, a.lpos.x, a.lpos.y, a.lpos.z),
                                       distance3(lp.x, lp.y, lp.z, a.spos.x, a.spos.y, a.spos.z)));
        const float w = rn_mul(safe_rcp(rn_mul(dist, dist)), sw);
        const float sc = STORAGE == FBP_STORAGE_F32 ? 1.f : a.scales[l];
        scatter_walk(dist, a.ext, a.dt, a.T, [&](int bin) {
            const size_t idx = LAYOUT == FBP_LAYOUT_TIME_MAJOR ? size_t(bin) * a.L + l
                                                                : size_t(l) * a.T + uint32_t(bin);
            const float v = load_sample<STORAGE>(a.data, idx, sc);
            // Quantised storage was thresholded before quantisation.
            if (STORAGE == FBP_STORAGE_F32 && !(v > kEps)) return;
            const float y = rn_sub(rn_mul(v, w), comp);
            const float t = rn_add(sum, y);
            comp = rn_sub(rn_sub(t, sum), y);
            sum = t;
        });
    }
    a.volume[vid] = sum;
}

template <int STORAGE, int LAYOUT>
void launch(const ScatterArgs& a, hipStream_t s) {
    const size_t nv = size_t(a.X) * a.Y * a.Z;
    hipLaunchKernelGGL((confocal_scatter_kernel<STORAGE, LAYOUT>), dim3(unsigned((nv + 255) / 256)),
                       dim3(256), 0, s, a);
}

}  // namespace

fbp_status confocal_scatter_impl(fbp_context ctx, const fbp_reconstruction_info* info,
                                 const fbp_transient* tr, const fbp_vec3* laser_targets,
                                 float* volume, hipStream_t s, const uint32_t* sample_l,
                                 const float* sample_w, uint32_t samples) {
    ScatterArgs a{};
    a.data = tr->data;
    a.scales = tr->scales;
    a.laser = laser_targets;
    a.volume = volume;
    a.X = info->voxel_resolution.x;
    a.Y = info->voxel_resolution.y;
    a.Z = info->voxel_resolution.z;
    a.L = info->num_laser_targets;
    a.T = info->num_time_bins;
    a.hmin[0] = info->hidden_volume_min.x;
    a.hmin[1] = info->hidden_volume_min.y;
    a.hmin[2] = info->hidden_volume_min.z;
    a.vs[0] = info->hidden_volume_voxel_size.x;
    a.vs[1] = info->hidden_volume_voxel_size.y;
    a.vs[2] = info->hidden_volume_voxel_size.z;
    a.t0 = info->time_offset;
    a.dt = info->time_step;
    a.ext = voxel_extent(a.vs[0], a.vs[1], a.vs[2]);
    a.discard = info->discard_first_last_bounces;
    a.lpos = info->laser_position;
    a.spos = info->sensor_position;
    a.items = samples;
    a.sample_l = samples ? sample_l : nullptr;
    a.sample_w = samples ? sample_w : nullptr;
    a.tiles_x = (a.X + 15) / 16;
    a.tiles_y = (a.Y + 15) / 16;
    if (tr->layout == FBP_LAYOUT_TIME_MAJOR)
        launch<FBP_STORAGE_F32, FBP_LAYOUT_TIME_MAJOR>(a, s);
    else if (tr->storage == FBP_STORAGE_F32)
        launch<FBP_STORAGE_F32, FBP_LAYOUT_TIME_FASTEST>(a, s);
    else if (tr->storage == FBP_STORAGE_F16)
        launch<FBP_STORAGE_F16, FBP_LAYOUT_TIME_FASTEST>(a, s);
    else
        launch<FBP_STORAGE_BF16, FBP_LAYOUT_TIME_FASTEST>(a, s);
    FBP_CHECK_HIP(ctx, hipGetLastError());
    return FBP_OK;
}

}  // namespace fbp
