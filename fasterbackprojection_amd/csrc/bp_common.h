// Pieces shared by the voxel backprojection kernels (confocal, exhaustive).
#pragma once

#include <hip/hip_fp16.h>

#include <cmath>

#include "chain.h"
#include "fbp_internal.h"

namespace fbp {

// A wave owns a voxel tile of (4*NI) x (4*NM) x (4*NJ) voxels along (vx, vy, z);
// lane (lx, lz, ly) in 4x4x4 holds the NI*NJ*NM voxels
//   vx = X0 + lx + 4i,  z = Z0 + lz + 4j,  vy = Y0 + ly + 4m,
// accumulated in registers at slot (i*NJ + j)*NM + m.
struct VolumeGeom {
    float hmin[3];      // hidden_volume_min (world x, y, z)
    float vs[3];        // voxel size (world x, y, z)
    uint32_t res[3];    // voxel_resolution: X (vx), Y (vy), Z (z = depth)
    uint32_t lt[3];     // lane-tile counts along vx, vy, z (NI, NM, NJ)
    uint32_t tile[3];   // wave-tile extents along vx, vy, z (= 4 * lt)
    uint32_t tiles[3];  // number of wave tiles along vx, vy, z
    uint32_t num_tiles;
    uint32_t slots;     // voxels per lane
};

inline VolumeGeom make_geom(const fbp_reconstruction_info& info, uint32_t ni, uint32_t nm,
                            uint32_t nj) {
    VolumeGeom g;
    g.hmin[0] = info.hidden_volume_min.x;
    g.hmin[1] = info.hidden_volume_min.y;
    g.hmin[2] = info.hidden_volume_min.z;
    g.vs[0] = info.hidden_volume_voxel_size.x;
    g.vs[1] = info.hidden_volume_voxel_size.y;
    g.vs[2] = info.hidden_volume_voxel_size.z;
    g.res[0] = info.voxel_resolution.x;
    g.res[1] = info.voxel_resolution.y;
    g.res[2] = info.voxel_resolution.z;
    g.lt[0] = ni;
    g.lt[1] = nm;
    g.lt[2] = nj;
    for (int a = 0; a < 3; ++a) {
        g.tile[a] = 4 * g.lt[a];
        g.tiles[a] = (g.res[a] + g.tile[a] - 1) / g.tile[a];
    }
    g.num_tiles = g.tiles[0] * g.tiles[1] * g.tiles[2];
    g.slots = ni * nm * nj;
    return g;
}

// Window length (bins) that covers every bin a tile can produce for one
// (laser[, sensor]) pair: the dist range over a tile is at most
// `paths` x the tile's centre-box diagonal; +8 bins of margin.
inline int window_bins(const VolumeGeom& g, float dt, int paths) {
    double ex = double(g.tile[0] - 1) * std::fabs(g.vs[0]);  // world x <- vx
    double ey = double(g.tile[2] - 1) * std::fabs(g.vs[1]);  // world y <- z
    double ez = double(g.tile[1] - 1) * std::fabs(g.vs[2]);  // world z <- vy
    double diag = std::sqrt(ex * ex + ey * ey + ez * ez);
    double w = std::ceil(paths * diag / double(dt)) + 8.0;
    int wi = int(w);
    return (wi + 3) & ~3;
}

// Centre box of a wave tile per world axis.
__device__ __forceinline__ void tile_box(const VolumeGeom& g, uint32_t X0, uint32_t Y0, uint32_t Z0,
                                         float lo[3], float hi[3]) {
    const uint32_t first[3] = {X0, Z0, Y0};
    const uint32_t ext[3] = {g.tile[0], g.tile[2], g.tile[1]};
    for (int ax = 0; ax < 3; ++ax) {
        float c0 = voxel_center(g.hmin[ax], g.vs[ax], first[ax]);
        float c1 = voxel_center(g.hmin[ax], g.vs[ax], first[ax] + ext[ax] - 1);
        lo[ax] = fminf(c0, c1);
        hi[ax] = fmaxf(c0, c1);
    }
}

inline bool finite_info(const fbp_reconstruction_info& i) {
    const float vals[] = {i.sensor_position.x, i.sensor_position.y, i.sensor_position.z,
                          i.laser_position.x, i.laser_position.y, i.laser_position.z,
                          i.time_offset, i.time_step,
                          i.hidden_volume_min.x, i.hidden_volume_min.y, i.hidden_volume_min.z,
                          i.hidden_volume_voxel_size.x, i.hidden_volume_voxel_size.y,
                          i.hidden_volume_voxel_size.z};
    for (float v : vals)
        if (!std::isfinite(v)) return false;
    return true;
}

// Loads one transient sample as float from any storage (scale applied).
template <int STORAGE>
__device__ __forceinline__ float load_sample(const void* data, size_t idx, float scale) {
    if constexpr (STORAGE == FBP_STORAGE_F32) {
        return static_cast<const float*>(data)[idx];
    } else if constexpr (STORAGE == FBP_STORAGE_F16) {
        uint16_t b = static_cast<const uint16_t*>(data)[idx];
        return __half2float(__ushort_as_half(b)) * scale;
    } else {
        uint16_t b = static_cast<const uint16_t*>(data)[idx];
        return __uint_as_float(uint32_t(b) << 16) * scale;
    }
}

__device__ __forceinline__ float rfl(float v) {
    return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v)));
}
__device__ __forceinline__ int rfl(int v) { return __builtin_amdgcn_readfirstlane(v); }

// Distance from point p to the axis-aligned box [lo, hi] (float, ~1 ulp).
__device__ __forceinline__ float box_min_dist(const float p[3], const float lo[3],
                                              const float hi[3]) {
    float s = 0.f;
    for (int a = 0; a < 3; ++a) {
        float d = fmaxf(fmaxf(lo[a] - p[a], p[a] - hi[a]), 0.f);
        s += d * d;
    }
    return sqrtf(s);
}
__device__ __forceinline__ float box_max_dist(const float p[3], const float lo[3],
                                              const float hi[3]) {
    float s = 0.f;
    for (int a = 0; a < 3; ++a) {
        float d = fmaxf(fabsf(lo[a] - p[a]), fabsf(hi[a] - p[a]));
        s += d * d;
    }
    return sqrtf(s);
}

// Sums per-chunk partial tiles in FP64 (fixed order) and scatters them to
// the [Z][Y][X] volume.  partial layout: [chunk][tile][slot][lane(64)].
// corr (may be null) is a second buffer of the same layout added in, for
// the (chunk, tile) blocks whose corr_used[chunk][tile] is set (corr_used is
// required with corr).
// skip (may be null): when *skip != 0 on the device the kernel returns.
// The pure chain's threshold row S_k = min{s >= 0 : bin(s) >= k}, k = 0..T
// (bp_confocal.hip), computed on stream s when (T, dt) differs from the
// context's cached row; device pointer in *thr.  Calls on one context are
// stream-ordered, so a later call on the same stream sees the row.
fbp_status pure_threshold_row(fbp_context ctx, uint32_t T, float dt, hipStream_t s,
                              const float** thr);

// Threshold rows S_k of the confocal chain, one per scan point (rows = L,
// laser required) or the pure chain's single row (rows = 1, laser unused):
// thr[row * (T + 1) + k], k = 0..T.
fbp_status launch_confocal_thresholds(fbp_context ctx, const fbp_vec3* laser, uint32_t rows,
                                      uint32_t T, float t0, float dt, uint32_t discard,
                                      fbp_vec3 lpos, fbp_vec3 spos, float* thr, hipStream_t s);

fbp_status launch_reduce(fbp_context ctx, const float* partial, const float* corr, float* volume,
                         const VolumeGeom& g, uint32_t num_chunks, double scale,
                         hipStream_t stream, const int* skip = nullptr,
                         const uint8_t* corr_used = nullptr);

// Confocal voxel backprojection (bp_confocal.hip).  With samples != 0 every
// 16^3 voxel tile visits its own scan-point list sample_l[tile][0..samples)
// with weights sample_w[tile][..] (tile-sampled MIS); otherwise all scan
// points with weight 1 (the aligned-grid path may then run).
fbp_status confocal_voxel_impl(fbp_context ctx, const fbp_reconstruction_info* info,
                               const fbp_transient* tr, const fbp_vec3* laser_targets,
                               float* volume, void* stream, const uint32_t* sample_l,
                               const float* sample_w, uint32_t samples);

// Voxel-scattering variant of the confocal voxel backprojection
// (bp_scatter.hip); the same sample-list convention as confocal_voxel_impl.
fbp_status confocal_scatter_impl(fbp_context ctx, const fbp_reconstruction_info* info,
                                 const fbp_transient* tr, const fbp_vec3* laser_targets,
                                 float* volume, hipStream_t s, const uint32_t* sample_l,
                                 const float* sample_w, uint32_t samples);

}  // namespace fbp
