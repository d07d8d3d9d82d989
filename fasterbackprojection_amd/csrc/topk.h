// k smallest of 64-bit keys held by a workgroup of 256 threads (4 waves),
// by k rounds of wave-wide minima on DPP lane moves (no LDS round trips
// inside a round), then one merge round in wave 0.  Keys are unique
// (callers put an index in the low word), so every round has one winner.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace fbp {

constexpr uint64_t kTopkPad = ~uint64_t(0);

// Minimum over the 64 lanes of a full wave (every lane gets it).  DPP:
// quad_perm [1,0,3,2] and [2,3,0,1], row_half_mirror, row_mirror give the
// minimum of each 16-lane row; the four rows are read out by lane index.
__device__ __forceinline__ uint32_t wave_min_u32(uint32_t v) {
    v = min(v, uint32_t(__builtin_amdgcn_mov_dpp(int(v), 0xB1, 0xF, 0xF, false)));
    v = min(v, uint32_t(__builtin_amdgcn_mov_dpp(int(v), 0x4E, 0xF, 0xF, false)));
    v = min(v, uint32_t(__builtin_amdgcn_mov_dpp(int(v), 0x141, 0xF, 0xF, false)));
    v = min(v, uint32_t(__builtin_amdgcn_mov_dpp(int(v), 0x140, 0xF, 0xF, false)));
    const uint32_t a = __builtin_amdgcn_readlane(v, 0), b = __builtin_amdgcn_readlane(v, 16);
    const uint32_t c = __builtin_amdgcn_readlane(v, 32), d = __builtin_amdgcn_readlane(v, 48);
    return min(min(a, b), min(c, d));
}

__device__ __forceinline__ uint64_t wave_min_u64(uint64_t v) {
    const uint32_t hi = wave_min_u32(uint32_t(v >> 32));
    const uint32_t lo = wave_min_u32(uint32_t(v >> 32) == hi ? uint32_t(v) : 0xffffffffu);
    return (uint64_t(hi) << 32) | lo;
}

// The k smallest of the wave's 64 x E keys, ascending, into out[0..k)
// (written by lane 0).  Called by a full wave.
template <int E>
__device__ __forceinline__ void wave_topk(uint64_t (&key)[E], uint32_t k, uint64_t* out) {
    const uint32_t lane = threadIdx.x & 63;
    for (uint32_t r = 0; r < k; ++r) {
        uint64_t best = key[0];
#pragma unroll
        for (int j = 1; j < E; ++j) best = key[j] < best ? key[j] : best;
        const uint64_t w = wave_min_u64(best);
        if (lane == 0) out[r] = w;
#pragma unroll
        for (int j = 0; j < E; ++j)
            if (key[j] == w) key[j] = kTopkPad;  // pads stay pads
    }
}

// Block-wide: the k smallest (k <= KMAX) of the block's 256 x E keys and of
// prev[0..k) when has_prev, ascending, into res[0..k) (LDS).  res may be
// prev.  Called by all 256 threads.
template <int E, int KMAX>
__device__ void block_topk(uint64_t (&key)[E], uint32_t k, const uint64_t* prev, bool has_prev,
                           uint64_t* res) {
    __shared__ uint64_t part[4 * KMAX + KMAX];
    const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    wave_topk<E>(key, k, part + wave * KMAX);
    if (wave == 0 && lane < k) part[4 * KMAX + lane] = has_prev ? prev[lane] : kTopkPad;
    __syncthreads();
    if (wave == 0) {
        constexpr int M = (5 * KMAX + 63) / 64;
        uint64_t m[M];
#pragma unroll
        for (int j = 0; j < M; ++j) {
            const uint32_t i = lane + 64u * uint32_t(j);
            const uint32_t src = i / KMAX, r = i % KMAX;
            m[j] = (i < 5u * KMAX && r < k) ? part[src * KMAX + r] : kTopkPad;
        }
        wave_topk<M>(m, k, res);
    }
    __syncthreads();
}

}  // namespace fbp
