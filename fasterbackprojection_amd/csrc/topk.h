// k smallest of 64-bit keys held by a workgroup of 256 threads.  Keys are
// unique (callers put an index in the low word) apart from pads.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace fbp {

constexpr uint64_t kTopkPad = ~uint64_t(0);

// Block-wide (256 threads): the k smallest of the block's 256 x E unique
// keys (pads kTopkPad may repeat; they are chosen only when fewer than k
// other keys exist), ascending, into res[0..k) (LDS), k <= 256.  Radix selection
// of the k-th smallest key, 8 bits per pass from the top (a 256-bin LDS
// histogram of the keys that match the digits chosen so far; wave 0 finds
// the bin holding the rank), then the keys <= that key are gathered and
// ranked.  No per-candidate rounds: 8 passes whatever k is.
template <int E>
__device__ void block_select_k(const uint64_t (&key)[E], uint32_t k, uint64_t* res) {
    __shared__ uint32_t hist[256];
    __shared__ uint64_t s_pref;
    __shared__ uint32_t s_rank, s_cnt;
    __shared__ uint64_t gath[256];
    const uint32_t t = threadIdx.x, lane = t & 63;
    if (t == 0) {
        s_pref = 0;
        s_rank = k;  // 1-based rank of the wanted key among the matching keys
        s_cnt = 0;
    }
    uint64_t mask = 0;
    for (int shift = 56; shift >= 0; shift -= 8) {
        hist[t] = 0;
        __syncthreads();
        const uint64_t pref = s_pref;
#pragma unroll
        for (int j = 0; j < E; ++j)
            if ((key[j] & mask) == pref) atomicAdd(&hist[uint32_t(key[j] >> shift) & 255u], 1u);
        __syncthreads();
        if (t < 64) {
            // Lane owns bins 4 lane .. 4 lane + 3; inclusive scan of the lane
            // totals across the wave.
            const uint32_t b0 = hist[4 * lane], b1 = hist[4 * lane + 1], b2 = hist[4 * lane + 2],
                           b3 = hist[4 * lane + 3];
            const uint32_t tot = b0 + b1 + b2 + b3;
            uint32_t inc = tot;
            for (int off = 1; off < 64; off <<= 1) {
                const uint32_t o = __shfl_up(inc, off, 64);
                if (int(lane) >= off) inc += o;
            }
            const uint32_t exc = inc - tot, r = s_rank;
            if (r > exc && r <= inc) {
                uint32_t d = 4 * lane, before = exc;
                const uint32_t bs[4] = {b0, b1, b2, b3};
                for (int q = 0; q < 4; ++q) {
                    if (r <= before + bs[q]) {
                        d = 4 * lane + uint32_t(q);
                        break;
                    }
                    before += bs[q];
                }
                s_pref = pref | (uint64_t(d) << shift);
                s_rank = r - before;
            }
        }
        mask |= uint64_t(255) << shift;
        __syncthreads();
    }
    // Keys below the k-th first (fewer than k), then copies of the k-th
    // itself (one, or pads when fewer than k other keys exist).
    const uint64_t kth = s_pref;
#pragma unroll
    for (int j = 0; j < E; ++j)
        if (key[j] < kth) gath[atomicAdd(&s_cnt, 1u)] = key[j];
    __syncthreads();
#pragma unroll
    for (int j = 0; j < E; ++j)
        if (key[j] == kth) {
            const uint32_t slot = atomicAdd(&s_cnt, 1u);
            if (slot < k) gath[slot] = key[j];
        }
    __syncthreads();
    const uint32_t got = min(s_cnt, k);
    if (t < k) {
        // Fewer than k keys at or below the k-th (only when pads fill in):
        // the rest of res is pads.
        const uint64_t v = t < got ? gath[t] : kTopkPad;
        uint32_t rank = t;
        if (t < got) {
            rank = 0;
            for (uint32_t j = 0; j < got; ++j)
                rank += (gath[j] < v || (gath[j] == v && j < t)) ? 1u : 0u;
        }
        res[rank] = v;
    }
    __syncthreads();
}

}  // namespace fbp
