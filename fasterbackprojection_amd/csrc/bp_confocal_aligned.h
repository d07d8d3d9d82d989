// Aligned-grid (ring convolution) path of the confocal voxel backprojection.
#pragma once

#include "fbp_internal.h"

namespace fbp {

struct AlignedPlan {
    bool usable;            // host-side conditions hold (the grid check runs on the device)
    uint32_t nx;            // scan grid edge
    uint32_t R;             // scan points per voxel pitch (1 or 2)
    uint32_t pn;            // phase plane edge nx / R
    uint32_t classes;       // offset sign classes (1 for R = 1, 4 for R = 2)
    uint32_t stride;        // LDS row stride (floats)
    size_t lds_bytes;       // dynamic LDS of the convolution kernel
    uint32_t kq, mq, n;     // quadrant extents, n = kq * mq
    uint32_t sort_n;        // n rounded up to a power of two
    uint32_t chunks;        // bin chunks per (depth, phase)
    size_t bytes_entries, bytes_starts, bytes_work, bytes_tm, bytes_partial;
};

// Host-side eligibility and sizes: pure chain; F32 storage (either layout)
// or F16/BF16 time-fastest; a square scan grid of nx^2 points with X = 64 or
// 128 voxels along x and nx = X or 2 X (the voxel lattice on every first or
// second scan point: checked on the device); a multiple of 16 voxels (at
// most X) along y; enough LDS per workgroup.
AlignedPlan plan_aligned_confocal(fbp_context ctx, const fbp_reconstruction_info& info,
                                  const fbp_transient& tr);

// Depth-plane mode (backprojectConfocal, FB/transient_reconstruction.cuh:41-63):
// pure chain, F32 time-fastest input, a 64^2 or 128^2 scan grid, and a wall
// normal with zero x and z (the planes move the scan points along y).  The
// outputs are the scan points themselves at every depth (X = VY = nx,
// ZD = num_depths), weight 1.
AlignedPlan plan_aligned_depth(fbp_context ctx, const fbp_reconstruction_info& info,
                               const fbp_transient& tr, uint32_t num_depths);

size_t aligned_scratch_bytes(const AlignedPlan& p);

// Launches the grid check (writes *flag), the offset tables, the reorder
// into phase planes (unless the input is F32 time-major with R = 1) and the
// convolution + FP64 reduction into volume.
// Every kernel after the check returns at once when *flag == 0.
fbp_status launch_aligned_confocal(fbp_context ctx, const AlignedPlan& p,
                                   const fbp_reconstruction_info& info, const fbp_transient& tr,
                                   const fbp_vec3* laser, float* volume, int* flag, void* base,
                                   hipStream_t s);

// The same for the depth-plane activation [nx^2][num_depths] (depths: device
// doubles).  The tables are kept while the parameters, the scan points and
// the depth values stay the same (compared on the device).
fbp_status launch_aligned_depth(fbp_context ctx, const AlignedPlan& p,
                                const fbp_reconstruction_info& info, const fbp_transient& tr,
                                const fbp_vec3* laser, const double* depths, uint32_t num_depths,
                                float* activation, int* flag, void* base, hipStream_t s);

}  // namespace fbp
