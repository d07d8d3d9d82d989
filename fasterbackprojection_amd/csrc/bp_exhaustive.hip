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
#include "group_common.h"

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

__device__ __forceinline__ float pick4f(const float (&v)[4], int i) {
    return i == 0 ? v[0] : i == 1 ? v[1] : i == 2 ? v[2] : v[3];
}

// Window entry of bin k.  Linear-weight pairs (LIN, bins >= kLinMinBin
// only): {I / m^2, -2 I / m^3} with m = k + 1/2, so I / q^2 for q = m + d is
// a + b d up to 3 d^2 / m^2 <= 0.75 / m^2 relative (<= 3e-5 at m >= 160).
// Other pairs: {I, 0} and the weight is 1 / q^2.  Bins in units of dt: the
// reduction scales by 1 / dt^2.
constexpr int kLinMinBin = 160;

// Fast update of one pair on the lane's 32 voxels.  rq: laser distances /
// dt; AQ, ZQ: the pair's per-axis squared sensor distances (x + y, z) /
// dt^2.  q = rl/dt + sqrt(ss/dt^2) (+ (off - t0)/dt), bin (int)q.  A voxel
// whose q lies within the estimate's error bound of a bin boundary still
// adds its fast contribution but sets its bit in the returned mask (bit
// 31 - slot); exhaustive_fixup later replaces that contribution with the
// reference's exact chain.  No select and no division per update.
template <bool LIN, bool COFF>
__device__ __forceinline__ uint32_t exhaustive_fast(float (&acc)[32], const float (&rq)[32],
                                                    const float (&AQ)[4][2], const float (&ZQ)[4],
                                                    const float2* win, int base, float hme,
                                                    float coffq) {
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
                const int k = static_cast<int>(q);
                const float d = __builtin_amdgcn_fractf(q) - 0.5f;
                band_mask = band_mask + band_mask + (fabsf(d) > hme ? 1u : 0u);
                const float2 e = win[base + k];
                if constexpr (LIN)
                    acc[slot] += fmaf(d, e.y, e.x);
                else
                    acc[slot] += e.x * __builtin_amdgcn_rcpf(q * q);
            }
        }
    }
    return band_mask;
}

// Pairs near the wall or near zero distance: every voxel takes the
// reference's exact chain (bin by IEEE division, weight safeRCP(dist^2)).
// Window entries {I, 0}.
__device__ __forceinline__ void exhaustive_near(float (&acc)[32], const float (&px)[4],
                                                const float (&py)[2], const float (&pz)[4],
                                                LaserPos lp, const float2* win, int base,
                                                float off, float sx, float sy, float sz,
                                                const ExhaustiveArgs& a) {
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int m = 0; m < 4; ++m) {
                const float de = exact_dist(px[i], py[j], pz[m], lp, sx, sy, sz, a, off);
                const int k = bin_of(de, a.dt);
                const float w = safe_rcp(rn_mul(de, de)) * a.dt2;
                acc[(i * 2 + j) * 4 + m] = fmaf(win[base + k].x, w, acc[(i * 2 + j) * 4 + m]);
            }
}

// Replaces the fast contribution of one flagged (voxel, pair) by the exact
// one: recomputes the fast path's q, bin and contribution with the same
// operations (so the same value) and the reference's IEEE chain, and adds
// the difference to the voxel's correction slot.
template <bool COFF>
__device__ __forceinline__ void exhaustive_fixup(int slot, float x, float y, float z, LaserPos lp,
                                                 const float2* win, int base, float coffq,
                                                 float off, float sx, float sy, float sz, bool lin,
                                                 const ExhaustiveArgs& a, float* corr) {
    const float sl = dot3(sq(rn_sub(x, lp.x)), sq(rn_sub(y, lp.y)), sq(rn_sub(z, lp.z)));
    const float rqv = sl * __builtin_amdgcn_rsqf(sl) * a.inv_dt;
    const float aq = (sq(rn_sub(x, sx)) + sq(rn_sub(y, sy))) * a.inv_dt2;
    const float zq = sq(rn_sub(z, sz)) * a.inv_dt2;
    const float s2 = aq + zq;
    float q = fmaf(s2, __builtin_amdgcn_rsqf(s2), rqv);
    if constexpr (COFF) q = q + coffq;
    const int kf = static_cast<int>(q);
    const float d = __builtin_amdgcn_fractf(q) - 0.5f;
    const float2 ef = win[base + kf];
    const float fast = lin ? fmaf(d, ef.y, ef.x) : ef.x * __builtin_amdgcn_rcpf(q * q);
    const float de = exact_dist(x, y, z, lp, sx, sy, sz, a, off);
    const int ke = bin_of(de, a.dt);
    float exact = 0.f;
    if (ke >= 0 && ke < int(a.T)) {
        float I = win[base + ke].x;
        if (lin) {
            const float mk = float(ke) + 0.5f;
            I = I * (mk * mk);
        }
        exact = I * (safe_rcp(rn_mul(de, de)) * a.dt2);
    }
    corr[slot * 256] += exact - fast;
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
    int lin;      // 1: window entries carry linear weights (kmin >= kLinMinBin)
    uint32_t l;   // laser index of the pair
};

template <int STORAGE, bool LINK>
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
    P.lin = (LINK && P.near == 0.f && P.kmin >= kLinMinBin) ? 1 : 0;
    P.l = l;
    return P;
}

// Workgroup: kWaves waves own one 16 (vx) x 16 (vy) x 32 (z) tile; wave w
// holds the z-slab 8w .. 8w+7 (32 voxels per lane) and all waves share the
// staged bin windows of a batch of (laser, sensor) pairs.  Batches are
// double-buffered: the windows of batch b+1 are fetched into registers while
// batch b is processed.
// COFF: the chain has a time offset or per-pair path offsets (q needs the
// pair's (off - t0) / dt).  LINK: the fast body uses linear-weight windows;
// pairs that do not qualify (bins below kLinMinBin) take the exact body.
// Without LINK the fast body divides (rcp) per update.  One fast body per
// kernel keeps its q computation next to its window reads.
template <int STORAGE, bool COFF, bool LINK>
__global__ __launch_bounds__(256, 2) void exhaustive_voxel_kernel(ExhaustiveArgs a) {
    extern __shared__ float smem[];
    // [2][batch * W] window entries, [2][batch] PairParams, the correction
    // slots of the boundary voxels [32][256 threads] and the band masks of
    // the batch's pairs [batch][256 threads] (thread-fastest: no bank
    // conflicts).
    float2* winbuf = reinterpret_cast<float2*>(smem);
    PairParams* prmbuf = reinterpret_cast<PairParams*>(winbuf + 2 * a.batch * a.W);
    float* corrl = reinterpret_cast<float*>(prmbuf + 2 * a.batch) + threadIdx.x;
    uint32_t* maskl = reinterpret_cast<uint32_t*>(corrl - threadIdx.x + 32 * 256) + threadIdx.x;

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
    uint32_t pending = 0;   // pairs of the current batch with flagged voxels (bit jl)

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
    // Window entries from the fetched samples: {I / m^2, -2 I / m^3} (m = k
    // + 1/2) for linear-weight pairs, {I, 0} otherwise.
    auto store = [&](float2* win, const PairParams* prm, int nb, const float (&reg)[kPrefetch]) {
#pragma unroll
        for (int i = 0; i < kPrefetch; ++i) {
            const int f = tid + i * 256;
            if (f < nb * a.W) {
                const int jl = int((float(f) + 0.5f) * a.inv_W);
                const float v = reg[i];
                float2 e = make_float2(v, 0.f);
                if (prm[jl].lin) {
                    const float mk = float(prm[jl].kmin + (f - jl * a.W)) + 0.5f;
                    const float r = __builtin_amdgcn_rcpf(mk);
                    const float wa = v * (r * r);
                    e = make_float2(wa, -2.f * wa * r);
                }
                win[f] = e;
            }
        }
    };

    float reg[kPrefetch];
    // Prologue: batch 0 into buffer 0.
    if (p_begin < p_end) {
        const int nb = int(min(uint64_t(a.batch), p_end - p_begin));
        if (tid < nb) prmbuf[tid] = pair_params<STORAGE, LINK>(a, p_begin + tid, tid, lo, hi);
        __syncthreads();
        fetch(p_begin, nb, prmbuf, reg);
        store(winbuf, prmbuf, nb, reg);
        __syncthreads();
    }
    int cur = 0;
    for (uint64_t pb = p_begin; pb < p_end; pb += a.batch) {
        const int nb = int(min(uint64_t(a.batch), p_end - pb));
        const uint64_t pn = pb + a.batch;
        const bool has_next = pn < p_end;
        const int nbn = has_next ? int(min(uint64_t(a.batch), p_end - pn)) : 0;
        const PairParams* prm = prmbuf + cur * a.batch;
        PairParams* prm_next = prmbuf + (cur ^ 1) * a.batch;
        if (has_next) {
            if (tid < nbn) prm_next[tid] = pair_params<STORAGE, LINK>(a, pn + tid, tid, lo, hi);
            __syncthreads();
            fetch(pn, nbn, prm_next, reg);
        }
        for (int jl = 0; jl < nb; ++jl) {
            const PairParams& P = prm[jl];
            const uint32_t l = uint32_t(rfl(int(P.l)));
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
            uint32_t mk = 0;
            if (near || (LINK && rfl(P.lin) == 0))
                exhaustive_near(acc, px, py, pz, lp, winbuf, base, off, sxp, syp, szp, a);
            else
                mk = exhaustive_fast<LINK, COFF>(acc, rq, AQ, ZQ, winbuf, base, hme, coffq);
            maskl[jl * 256] = mk;
            pending |= (mk != 0u ? 1u : 0u) << jl;
        }
        // The next batch's windows go to the other buffer (its previous batch
        // was finished by every thread before this batch began), which frees
        // the prefetch registers before the boundary voxels are handled.
        if (has_next) store(winbuf + (cur ^ 1) * per, prm_next, nbn, reg);
        // Deferred boundary voxels of this batch: each lane replaces the fast
        // contributions of its own flagged voxels, pair by pair in order
        // (deterministic), while this batch's windows are still staged.
        if (__any(pending != 0u)) {
            while (pending) {
                const int jl = __builtin_ctz(pending);
                pending &= pending - 1u;
                uint32_t m = maskl[jl * 256];
                const PairParams& P = prm[jl];
                const LaserPos lq{a.laser[P.l].x, a.laser[P.l].y, a.laser[P.l].z};
                const int base = P.base + cur * per;
                while (m) {
                    const int slot = 31 - __builtin_ctz(m);
                    m &= m - 1u;
                    const int i = slot >> 3, j = (slot >> 2) & 1, mm = slot & 3;
                    exhaustive_fixup<COFF>(slot, pick4f(px, i), j ? py[1] : py[0], pick4f(pz, mm),
                                           lq, winbuf, base, P.coffq, P.off, P.sx, P.sy, P.sz,
                                           P.lin != 0, a, corrl);
                }
            }
        }
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

// ---------------------------------------------------------------------------
// Laser-group kernel: linear-weight captures with many sensors.
//
// A workgroup of eight waves owns the same 16 (vx) x 16 (vy) x 32 (z) tile
// with 16 voxels per lane (wave w holds z = Z0 + lz + 4w) and a range of
// (laser group, sensor) batches; a laser group is kGroup = 8 consecutive
// lasers.  Each lane keeps the group's eight exact laser distances of its 16
// voxels in registers while the group is unchanged, and computes its 16
// exact sensor distances once per sensor for the eight pairs (l, s) of the
// group.  Both are the reference's own IEEE values (correctly rounded square
// roots of the same squared sums), so an update starts from the reference's
// dist = fl(rl + rs) (- t0 and + off when the chain has them) and needs no
// square root:
//   q = fma(dist, c_hi, fl(dist * c_lo)),  c_hi + c_lo = 1 / dt;
//   k = (int)q;  acc[v] += A_k + B_k q.
// The two-term quotient q truncates like the reference's fl(dist / dt): the
// host checks every float within 16 ulps of each bin boundary k dt (the only
// places where two quotients rounded from values within 2^-47 of dist / dt
// can truncate differently) before it chooses this kernel, so the bins are
// the reference's exactly.  The window entry of bin k is the linear
// expansion of I / q^2 about m = k + 1/2: {3 I r^2, -2 I r^3}, r = 1/m, from a
// per-bin table (bins >= kLinMinBin; other pairs take the reference's chain
// for every voxel).  One batch = one sensor and the group's eight pairs;
// wave g stages pair g's window (from a 4-bin aligned start, 16-byte loads
// and stores), the next batch's raw loads in flight while the current batch
// runs.
constexpr int kGThreads = 512;
constexpr int kGVec = 2;           // 4-bin loads per lane per batch: window <= 4 * 64 * kGVec bins
// Pair parameters are computed kGPrmBatches batches at a time, by the first
// kGPrmBatches * kGroup threads, into one half of a two-block LDS ring.
constexpr int kGPrmBatches = 16;

struct GroupArgs {
    const void* data;
    const float* scales;
    const fbp_vec3* laser;
    const fbp_vec3* sensor;
    float* partial;
    VolumeGeom g;
    uint32_t L, S, T;
    float t0, dt, inv_dt, inv_dt_lo, dt2;  // inv_dt + inv_dt_lo = 1 / dt
    uint32_t discard;
    float lpos[3], spos[3];
    int W;                    // window length in bins (multiple of 4): bins kmin4 .. kmin4 + W - 1
    uint32_t groups;          // ceil(L / kGroup)
    uint32_t groups_per_chunk;
    uint32_t sensor_parts;    // chunks along the sensors
    uint32_t sensors_per_part;
    uint32_t num_blocks;      // grid size (XCD remap)
};

// Pair parameters of one batch slot, written by one thread and read by all.
struct __align__(16) GroupPair {
    int base;      // g * W - kmin4: index of bin k in the batch's window buffer
    int kmin;      // first bin the tile can reach (- 3)
    int kmin4;     // kmin rounded down to a multiple of 4 (window start)
    int fast;      // 1: linear-weight fast body, 0: exact chain, -1: no pair (l >= L)
    float lx, ly, lz, off;     // laser; path offset fl(|l - Lpos| + |s - Spos|) (0 with discard)
    float sx, sy, sz, pad;     // sensor
};

// Laser-side parameters of the chunk's lasers against the tile (set-up).
struct __align__(16) GroupLaser {
    float x, y, z, offl;       // position, |l - Lpos| (0 with discard)
    float rmin, rmax, pad[2];  // box distances
};

template <bool COFF>
__device__ __forceinline__ void group_fast(float (&acc)[16], const float (&rl)[16],
                                           const float (&rs)[16], const float2* win, int base,
                                           float c_hi, float c_lo, float t0, float off) {
#pragma unroll
    for (int h = 0; h < 16; h += 8) {
        float q[8];
        float2 e[8];
#pragma unroll
        for (int v = 0; v < 8; ++v) {
            float d = rl[h + v] + rs[h + v];
            if constexpr (COFF) d = (d - t0) + off;
            q[v] = group_quotient(d, c_hi, c_lo);
            e[v] = win[base + static_cast<int>(q[v])];
        }
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int v = 0; v < 8; ++v) acc[h + v] += fmaf(q[v], e[v].y, e[v].x);
    }
}

// Raw window samples of 4 bins (converted and thresholded when stored).
template <int STORAGE>
struct GroupRaw {
    using type = float4;
};
template <>
struct GroupRaw<FBP_STORAGE_F16> {
    using type = uint2;
};
template <>
struct GroupRaw<FBP_STORAGE_BF16> {
    using type = uint2;
};

template <int STORAGE>
__device__ __forceinline__ void group_convert(const typename GroupRaw<STORAGE>::type& r, float sc,
                                              float (&v)[4]) {
    if constexpr (STORAGE == FBP_STORAGE_F32) {
        v[0] = r.x;
        v[1] = r.y;
        v[2] = r.z;
        v[3] = r.w;
#pragma unroll
        for (int c = 0; c < 4; ++c) v[c] = (v[c] > kEps) ? v[c] : 0.f;
    } else {
        const uint32_t w[2] = {r.x, r.y};
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const uint16_t b = uint16_t(w[c >> 1] >> (16 * (c & 1)));
            if constexpr (STORAGE == FBP_STORAGE_F16)
                v[c] = __half2float(__ushort_as_half(b)) * sc;
            else
                v[c] = __uint_as_float(uint32_t(b) << 16) * sc;
        }
    }
}

template <int STORAGE, bool COFF>
__global__ __launch_bounds__(kGThreads, 1) void exhaustive_group_kernel(GroupArgs a) {
    using Raw = typename GroupRaw<STORAGE>::type;
    extern __shared__ float smem[];
    // [2][kGroup][W] window entries, [2 * kGPrmBatches][kGroup] pair
    // parameters, the chunk's lasers [groups_per_chunk * kGroup] and sensors
    // [sensors_per_part], [T] weight factors.
    float2* winbuf = reinterpret_cast<float2*>(smem);
    GroupPair* prmbuf = reinterpret_cast<GroupPair*>(winbuf + 2 * kGroup * a.W);
    // Per batch slot: bits 0-7 pairs on the fast body, bits 8-15 pairs on
    // the exact chain.
    uint32_t* bmask = reinterpret_cast<uint32_t*>(prmbuf + 2 * kGPrmBatches * kGroup);
    GroupLaser* lasl = reinterpret_cast<GroupLaser*>(bmask + 2 * kGPrmBatches);
    float4* senl = reinterpret_cast<float4*>(lasl + a.groups_per_chunk * kGroup);
    float2* tab = reinterpret_cast<float2*>(senl + a.sensors_per_part);

    const int tid = threadIdx.x;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int lane = tid & 63;
    // Blocks are dealt to the XCDs round-robin; consecutive indices on one XCD
    // take the tiles of one chunk, so its L2 serves their shared windows.
    const uint32_t nb8 = a.num_blocks / 8u;
    const uint32_t bid = blockIdx.x < nb8 * 8u ? (blockIdx.x % 8u) * nb8 + blockIdx.x / 8u
                                               : blockIdx.x;
    const uint32_t tile = bid % a.g.num_tiles;
    const uint32_t chunk = bid / a.g.num_tiles;
    const uint32_t tx = tile % a.g.tiles[0];
    const uint32_t ty = (tile / a.g.tiles[0]) % a.g.tiles[1];
    const uint32_t tz = tile / (a.g.tiles[0] * a.g.tiles[1]);
    const uint32_t X0 = tx * a.g.tile[0], Y0 = ty * a.g.tile[1], Z0 = tz * a.g.tile[2];
    const int lx = lane & 3, lz = (lane >> 2) & 3, ly = lane >> 4;

    float px[4], pz[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        px[i] = voxel_center(a.g.hmin[0], a.g.vs[0], X0 + lx + 4 * i);
        pz[i] = voxel_center(a.g.hmin[2], a.g.vs[2], Y0 + ly + 4 * i);
    }
    const float py = voxel_center(a.g.hmin[1], a.g.vs[1], Z0 + lz + 4 * uint32_t(wave));
    float lo[3], hi[3];
    tile_box(a.g, X0, Y0, Z0, lo, hi);

    const uint32_t lgc = chunk / a.sensor_parts, sp = chunk % a.sensor_parts;
    const uint32_t g_begin = lgc * a.groups_per_chunk;
    const uint32_t g_end = min(a.groups, g_begin + a.groups_per_chunk);
    const uint32_t s_begin = sp * a.sensors_per_part;
    const uint32_t s_end = min(a.S, s_begin + a.sensors_per_part);
    const uint32_t ns = s_end > s_begin ? s_end - s_begin : 0u;
    const uint32_t nbatch = (g_end > g_begin ? g_end - g_begin : 0u) * ns;
    const float c = a.inv_dt, cl = a.inv_dt_lo;

    // Weight factors of bin k (m = k + 1/2): {3 r^2, -2 r^3}.
    for (uint32_t k = tid; k < a.T; k += kGThreads) {
        const double r = 1.0 / (double(k) + 0.5);
        tab[k] = make_float2(float(3.0 * r * r), float(-2.0 * r * r * r));
    }
    // The chunk's lasers (with their bounds against the tile) and sensors
    // (with |s - Spos|), so the per-batch parameters read LDS only.
    const uint32_t l_begin = g_begin * kGroup;
    const uint32_t nl = min(a.L, g_end * kGroup) - min(a.L, l_begin);
    for (uint32_t i = tid; i < nl; i += kGThreads) {
        const uint32_t l = l_begin + i;
        GroupLaser G;
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
    for (uint32_t i = tid; i < ns; i += kGThreads) {
        const uint32_t s = s_begin + i;
        const float x = a.sensor[s].x, y = a.sensor[s].y, z = a.sensor[s].z;
        const float offs =
            a.discard == 0 ? distance3(x, y, z, a.spos[0], a.spos[1], a.spos[2]) : 0.f;
        senl[i] = make_float4(x, y, z, offs);
    }
    __syncthreads();

    // Parameters of batch b's pair g from the staged lasers and sensors.
    auto params = [&](uint32_t b, int g, GroupPair* out) {
        GroupPair P;
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
        const GroupLaser G = lasl[li];
        P.lx = G.x;
        P.ly = G.y;
        P.lz = G.z;
        P.off = a.discard == 0 ? rn_add(G.offl, offs) : 0.f;
        const float dmin = G.rmin + rsmin - a.t0 + P.off;
        const float dmax = G.rmax + rsmax - a.t0 + P.off;
        P.kmin = int(floorf(dmin * a.inv_dt)) - 3;
        P.kmin4 = P.kmin & ~3;
        P.base = g * a.W - P.kmin4;
        const bool near = (dmin < 0.0045f && dmax > -0.0045f) || G.rmin < 1.0e-3f || rsmin < 1.0e-3f;
        P.fast = (!near && P.kmin >= kLinMinBin) ? 1 : 0;
        *out = P;
    };

    // Parameters of batch bn's pair g for the threads of a block (lanes
    // 8j .. 8j + 7 of a wave hold one batch; every lane of the wave takes
    // part, with valid = false past the last batch) and the batch's masks.
    auto params_block = [&](uint32_t bn, int g, bool valid) {
        int fast = -1;
        if (valid) {
            GroupPair* o = prmbuf + (bn % (2u * kGPrmBatches)) * kGroup + g;
            params(bn, g, o);
            fast = o->fast;
        }
        const uint64_t fm = __ballot(fast == 1), em = __ballot(fast == 0);
        const int sh = lane & ~(kGroup - 1);
        if (valid && g == 0)
            bmask[bn % (2u * kGPrmBatches)] =
                uint32_t((fm >> sh) & 0xffu) | (uint32_t((em >> sh) & 0xffu) << 8);
    };

    // Raw loads of batch b: wave g fetches pair g's bins kmin4 + 4 j .. + 3,
    // j = lane + 64 i.  Nothing reads them before the store, so they stay in
    // flight across the batch.  The pair's row is a scalar address; lanes whose
    // bins lie outside [0, T) read bins inside the row (the store zeroes
    // them), and a batch slot without a pair (l >= L) reads row 0.
    auto fetch = [&](uint32_t gi, uint32_t si, const GroupPair* prm, Raw (&reg)[kGVec]) {
        const int fast = rfl(prm[wave].fast);
        const int kmin4 = rfl(prm[wave].kmin4);
        const uint32_t grp = g_begin + gi;
        const uint32_t s = s_begin + si;
        const uint64_t pair = fast >= 0 ? uint64_t(grp * kGroup + uint32_t(wave)) * a.S + s : 0u;
#pragma unroll
        for (int i = 0; i < kGVec; ++i) {
            const int k0 = kmin4 + 4 * (lane + 64 * i);
            // T <= 8192 and a multiple of 4 for this kernel.
            const uint32_t kc = uint32_t(min(max(k0, 0), int(a.T) - 4));
            __builtin_assume(kc < 8192u);
            if constexpr (STORAGE == FBP_STORAGE_F32) {
                const float* row = static_cast<const float*>(a.data) + pair * a.T;
                reg[i] = *reinterpret_cast<const float4*>(row + kc);
            } else {
                const uint16_t* row = static_cast<const uint16_t*>(a.data) + pair * a.T;
                reg[i] = *reinterpret_cast<const uint2*>(row + kc);
            }
        }
    };
    auto store = [&](uint32_t gi, uint32_t si, float2* win, const GroupPair* prm,
                     const Raw (&reg)[kGVec]) {
        const int fast = rfl(prm[wave].fast);
        if (fast < 0) return;
        const int kmin4 = rfl(prm[wave].kmin4);
        float sc = 1.f;
        if constexpr (STORAGE != FBP_STORAGE_F32) {
            const uint32_t grp = g_begin + gi;
            const uint32_t s = s_begin + si;
            sc = a.scales[uint64_t(grp * kGroup + uint32_t(wave)) * a.S + s];
        }
#pragma unroll
        for (int i = 0; i < kGVec; ++i) {
            const int j = lane + 64 * i;
            if (4 * j < a.W) {
                const int k0 = kmin4 + 4 * j;
                const bool ok = k0 >= 0 && k0 < int(a.T);
                float v[4];
                group_convert<STORAGE>(reg[i], sc, v);
                // Bins outside [0, T) hold zero samples (their loads read
                // element 0).
#pragma unroll
                for (int c = 0; c < 4; ++c) v[c] = ok ? v[c] : 0.f;
                float4 e01, e23;
                if (fast == 1) {
                    // Factors of bins k0 .. k0 + 3, the index clamped into
                    // the table (T is a multiple of 4) for the zero samples.
                    const int kt = min(max(k0, 0), int(a.T) - 4);
                    const float4 t01 = *reinterpret_cast<const float4*>(tab + kt);
                    const float4 t23 = *reinterpret_cast<const float4*>(tab + kt + 2);
                    e01 = make_float4(v[0] * t01.x, v[0] * t01.y, v[1] * t01.z, v[1] * t01.w);
                    e23 = make_float4(v[2] * t23.x, v[2] * t23.y, v[3] * t23.z, v[3] * t23.w);
                } else {
                    e01 = make_float4(v[0], 0.f, v[1], 0.f);
                    e23 = make_float4(v[2], 0.f, v[3], 0.f);
                }
                // Four 8-byte entries (two paired LDS writes: no register
                // moves to make 16-byte sources).
                float2* dst = win + wave * a.W + 4 * j;
                dst[0] = make_float2(e01.x, e01.y);
                dst[1] = make_float2(e01.z, e01.w);
                dst[2] = make_float2(e23.x, e23.y);
                dst[3] = make_float2(e23.z, e23.w);
            }
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

    const int per = kGroup * a.W;
    Raw reg[kGVec];
    if (nbatch > 0) {
        // Parameters of the first two blocks of batches.
        if (tid < 2 * kGPrmBatches * kGroup)
            params_block(uint32_t(tid / kGroup), tid % kGroup, uint32_t(tid / kGroup) < nbatch);
        __syncthreads();  // parameters and the weight table
        fetch(0, 0, prmbuf, reg);
        store(0, 0, winbuf, prmbuf, reg);
        __syncthreads();
    }
    // Batch b = gi * ns + si (laser group gi of the chunk, sensor si of the
    // part), kept as running counters.
    uint32_t gi = 0, si = 0;
    for (uint32_t b = 0; b < nbatch; ++b) {
        const int cur = int(b & 1u);
        uint32_t gn = gi, sn = si + 1;
        if (sn == ns) {
            sn = 0;
            ++gn;
        }
        const GroupPair* prm = prmbuf + (b % (2u * kGPrmBatches)) * kGroup;
        const bool has_next = b + 1 < nbatch;
        const GroupPair* prm_next = prmbuf + ((b + 1) % (2u * kGPrmBatches)) * kGroup;
        if (has_next) fetch(gn, sn, prm_next, reg);

        if (si == 0) {
            // New laser group: the eight exact laser distances of the lane's voxels.
#pragma unroll
            for (int g = 0; g < kGroup; ++g) {
                const float lxg = rfl(prm[g].lx), lyg = rfl(prm[g].ly), lzg = rfl(prm[g].lz);
                float XL2[4], ZL2[4];
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    XL2[i] = sq(rn_sub(px[i], lxg));
                    ZL2[i] = sq(rn_sub(pz[i], lzg));
                }
                const float YL2 = sq(rn_sub(py, lyg));
#pragma unroll
                for (int i = 0; i < 4; ++i)
#pragma unroll
                    for (int m = 0; m < 4; ++m)
                        rl[g][i * 4 + m] = rn_sqrt(dot3(XL2[i], YL2, ZL2[m]));
            }
        }
        // Exact sensor distances of the batch's sensor.
        float rs[16];
        const float sxp = rfl(prm[0].sx), syp = rfl(prm[0].sy), szp = rfl(prm[0].sz);
        {
            float XS2[4], ZS2[4];
            const float YS2 = sq(rn_sub(py, syp));
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                XS2[i] = sq(rn_sub(px[i], sxp));
                ZS2[i] = sq(rn_sub(pz[i], szp));
            }
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int m = 0; m < 4; ++m)
                    rs[i * 4 + m] = sqrt_rn_normal(dot3(XS2[i], YS2, ZS2[m]));
        }
        const uint32_t masks = uint32_t(rfl(int(bmask[b % (2u * kGPrmBatches)])));
        const int fast_pairs = int(masks & 0xffu);
        int exact_pairs = int(masks >> 8);
        if (fast_pairs == (1 << kGroup) - 1) {
            // Common case, straight-line: all eight pairs take the fast body,
            // their window offsets and path offsets read up front.
            int bases[kGroup];
            float offs[kGroup];
#pragma unroll
            for (int g = 0; g < kGroup; ++g) {
                bases[g] = rfl(prm[g].base) + cur * per;
                offs[g] = rfl(prm[g].off);
            }
#pragma unroll
            for (int g = 0; g < kGroup; ++g)
                group_fast<COFF>(acc, rl[g], rs, winbuf, bases[g], c, cl, a.t0, offs[g]);
        } else {
#pragma unroll
            for (int g = 0; g < kGroup; ++g) {
                if (fast_pairs & (1 << g))
                    group_fast<COFF>(acc, rl[g], rs, winbuf, rfl(prm[g].base) + cur * per, c,
                                     cl, a.t0, rfl(prm[g].off));
            }
        }
        // Pairs near the wall or zero distance, or with bins below the linear
        // range: the reference's chain and weight for every voxel (entries
        // {I, 0}).
        while (exact_pairs) {
            const int g = __builtin_ctz(exact_pairs);
            exact_pairs &= exact_pairs - 1;
            const GroupPair& P = prm[g];
            const LaserPos lp{rfl(P.lx), rfl(P.ly), rfl(P.lz)};
            const int base = rfl(P.base) + cur * per;
            const float off = rfl(P.off);
            ExhaustiveArgs ea{};
            ea.t0 = a.t0;
            ea.discard = a.discard;
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int m = 0; m < 4; ++m) {
                    const float de = exact_dist(px[i], py, pz[m], lp, sxp, syp, szp, ea, off);
                    const int k = bin_of(de, a.dt);
                    const float w = safe_rcp(rn_mul(de, de)) * a.dt2;
                    acc[i * 4 + m] = fmaf(winbuf[base + k].x, w, acc[i * 4 + m]);
                }
        }
        if (has_next) {
            // The other window buffer was last read in batch b - 1.
            store(gn, sn, winbuf + (cur ^ 1) * per, prm_next, reg);
        }
        // First batch of a block: the block after next goes to the slots of
        // the block before, whose last batch ended with the previous barrier
        // (and is first needed by the fetch at the end of the next block).
        if (b > 0 && b % kGPrmBatches == 0 && tid < kGPrmBatches * kGroup) {
            const uint32_t bn = b + kGPrmBatches + uint32_t(tid / kGroup);
            params_block(bn, tid % kGroup, bn < nbatch);
        }
        __syncthreads();
        gi = gn;
        si = sn;
    }

    // Slots of the workgroup tile (lt = 4 x 4 x 8): wave w holds oz >> 2 = w.
    float* out = a.partial + (size_t(chunk) * a.g.num_tiles + tile) * a.g.slots * 64;
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int m = 0; m < 4; ++m)
            out[size_t((i * 8 + wave) * 4 + m) * 64 + lane] = acc[i * 4 + m];
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

template <int STORAGE>
static void (*pick_kernel_s(bool coff, bool link))(ExhaustiveArgs) {
    if (coff) return link ? exhaustive_voxel_kernel<STORAGE, true, true>
                          : exhaustive_voxel_kernel<STORAGE, true, false>;
    return link ? exhaustive_voxel_kernel<STORAGE, false, true>
                : exhaustive_voxel_kernel<STORAGE, false, false>;
}

static void (*pick_kernel(int storage, bool coff, bool link))(ExhaustiveArgs) {
    return storage == FBP_STORAGE_F32   ? pick_kernel_s<FBP_STORAGE_F32>(coff, link)
           : storage == FBP_STORAGE_F16 ? pick_kernel_s<FBP_STORAGE_F16>(coff, link)
                                        : pick_kernel_s<FBP_STORAGE_BF16>(coff, link);
}

// Host-side choice of the linear-weight kernel: when the scan targets lie on
// the relay wall plane, every leg is at least the hidden box's distance from
// that plane, so every pair's first window bin is at least
// floor((2 gap - t0) / dt) - 3 (the bounce offsets only add).  If that bound
// reaches kLinMinBin, (almost) every pair qualifies for linear weights.  The
// choice is performance only: pairs that do not qualify take the exact body.
static bool linear_weights_apply(const fbp_reconstruction_info& in) {
    const double n[3] = {in.relay_wall_normal.x, in.relay_wall_normal.y, in.relay_wall_normal.z};
    const double nn = std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
    if (!(nn > 0.0)) return false;
    const double w0[3] = {in.relay_wall_min_position.x, in.relay_wall_min_position.y,
                          in.relay_wall_min_position.z};
    const double lo[3] = {in.hidden_volume_min.x, in.hidden_volume_min.y, in.hidden_volume_min.z};
    const double hi[3] = {in.hidden_volume_max.x, in.hidden_volume_max.y, in.hidden_volume_max.z};
    // Signed distance of the box from the plane along the normal: its
    // smallest corner distance (negative when the box crosses the plane).
    double gap = 1e30;
    for (int c = 0; c < 8; ++c) {
        double dot = 0.0;
        for (int k = 0; k < 3; ++k) dot += ((((c >> k) & 1) ? hi[k] : lo[k]) - w0[k]) * n[k];
        gap = std::min(gap, dot / nn);
    }
    if (!(gap > 0.0)) return false;
    const double kmin = std::floor((2.0 * gap - double(in.time_offset)) / double(in.time_step)) - 3.0;
    return kmin >= double(kLinMinBin);
}

template <int STORAGE>
static void (*pick_group_s(bool coff))(GroupArgs) {
    return coff ? exhaustive_group_kernel<STORAGE, true> : exhaustive_group_kernel<STORAGE, false>;
}

// Laser-group kernel for linear-weight captures with at least kGroup lasers
// and 64 sensors whose window fits the staging registers and whose bin
// quotient check passes; *used = false (nothing launched) otherwise.  Chunks
// are (laser-group range, sensor range) blocks, reduced in a fixed order like
// the per-pair kernel's.
static fbp_status exhaustive_group_launch(fbp_context ctx, const fbp_reconstruction_info* info,
                                          const fbp_transient* tr, const fbp_vec3* laser_targets,
                                          const fbp_vec3* sensor_targets, float* volume,
                                          void* stream, float* fixed_partial, size_t fixed_floats,
                                          bool coff, bool* used) {
    *used = false;
    GroupArgs a{};
    a.g = make_geom(*info, 4, 4, 8);
    a.L = info->num_laser_targets;
    a.S = info->num_sensor_targets;
    a.T = info->num_time_bins;
    // Window of the 4-bin aligned start: the per-pair window plus 4 bins.
    a.W = window_bins(a.g, info->time_step, 2) + 4;
    // 4-bin loads: T a multiple of 4 and the samples 16-byte (F32) or 8-byte
    // (F16/BF16) aligned.
    const uintptr_t align = tr->storage == FBP_STORAGE_F32 ? 16u : 8u;
    if (a.L < uint32_t(kGroup) || a.S < 64u || a.T < 4u || a.T > 8192u || (a.T & 3u) != 0u ||
        a.W > 4 * 64 * kGVec || (reinterpret_cast<uintptr_t>(tr->data) % align) != 0u)
        return FBP_OK;
    // The quotient check depends on (dt, T) only: kept for the next call.
    if (!group_quotient_cached(ctx, info->time_step, a.T)) return FBP_OK;

    a.data = tr->data;
    a.scales = tr->scales;
    a.laser = laser_targets;
    a.sensor = sensor_targets;
    a.t0 = info->time_offset;
    a.dt = info->time_step;
    a.inv_dt = 1.0f / info->time_step;
    a.inv_dt_lo = group_inv_dt_lo(info->time_step);
    a.dt2 = info->time_step * info->time_step;
    a.discard = info->discard_first_last_bounces;
    a.lpos[0] = info->laser_position.x;
    a.lpos[1] = info->laser_position.y;
    a.lpos[2] = info->laser_position.z;
    a.spos[0] = info->sensor_position.x;
    a.spos[1] = info->sensor_position.y;
    a.spos[2] = info->sensor_position.z;
    a.groups = (a.L + kGroup - 1) / kGroup;

    // About four workgroups (one per CU at a time) per CU over the launch:
    // laser-group chunks first, then sensor parts of at least 64 sensors.
    const uint64_t cus = uint64_t(ctx->num_cus > 0 ? ctx->num_cus : 256);
    uint64_t want = (4 * cus + a.g.num_tiles - 1) / a.g.num_tiles;
    const size_t chunk_floats = size_t(a.g.num_tiles) * a.g.slots * 64;
    if (fixed_partial) {
        const uint64_t fit = fixed_floats / chunk_floats;
        if (fit == 0) return FBP_OK;  // the per-pair kernel reports the size error
        want = std::min<uint64_t>(want, fit);
    }
    if (want < 1) want = 1;
    uint64_t lgc = std::min<uint64_t>(a.groups, want);
    a.groups_per_chunk = uint32_t((a.groups + lgc - 1) / lgc);
    lgc = (a.groups + a.groups_per_chunk - 1) / a.groups_per_chunk;
    uint64_t parts = std::max<uint64_t>(1, want / lgc);
    parts = std::min<uint64_t>(parts, std::max<uint64_t>(1, a.S / 64));
    a.sensors_per_part = uint32_t((a.S + parts - 1) / parts);
    // LDS: windows, pair parameters, the chunk's lasers and sensors, weight
    // factors.  Sensor ranges shrink (more parts) to fit.
    auto lds_bytes = [&](uint32_t spp) {
        return size_t(2) * kGroup * a.W * sizeof(float2) +
               size_t(2) * kGPrmBatches * (kGroup * sizeof(GroupPair) + sizeof(uint32_t)) +
               size_t(a.groups_per_chunk) * kGroup * sizeof(GroupLaser) + size_t(spp) * sizeof(float4) +
               size_t(a.T) * sizeof(float2);
    };
    while (a.sensors_per_part > 64 && lds_bytes(a.sensors_per_part) > size_t(ctx->max_lds_per_block))
        a.sensors_per_part = (a.sensors_per_part + 1) / 2;
    const size_t shmem = lds_bytes(a.sensors_per_part);
    if (shmem > size_t(ctx->max_lds_per_block)) return FBP_OK;
    a.sensor_parts = uint32_t((a.S + a.sensors_per_part - 1) / a.sensors_per_part);
    const uint64_t chunks = lgc * a.sensor_parts;
    if (fixed_partial) {
        if (chunks * chunk_floats > fixed_floats) return FBP_OK;
        a.partial = fixed_partial;
    } else {
        FBP_CHECK(ensure_scratch(ctx, size_t(chunks) * chunk_floats * sizeof(float)));
        a.partial = static_cast<float*>(ctx->scratch);
    }
    a.num_blocks = uint32_t(chunks * a.g.num_tiles);

    void (*kern)(GroupArgs) = tr->storage == FBP_STORAGE_F32   ? pick_group_s<FBP_STORAGE_F32>(coff)
                              : tr->storage == FBP_STORAGE_F16 ? pick_group_s<FBP_STORAGE_F16>(coff)
                                                               : pick_group_s<FBP_STORAGE_BF16>(coff);
    hipStream_t s = as_stream(stream);
    FBP_CHECK_HIP(ctx, hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                                           hipFuncAttributeMaxDynamicSharedMemorySize, int(shmem)));
    hipLaunchKernelGGL(kern, dim3(a.num_blocks), dim3(kGThreads), shmem, s, a);
    FBP_CHECK_HIP(ctx, hipGetLastError());
    *used = true;
    return launch_reduce(ctx, a.partial, nullptr, volume, a.g, uint32_t(chunks),
                         1.0 / (double(a.dt) * double(a.dt)), s);
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

    const bool coff = !(info->time_offset == 0.f && info->discard_first_last_bounces != 0);
    const bool link = linear_weights_apply(*info);
    ctx->last_exhaustive_group = 0;
    if (link && ctx->exhaustive_group_kernel) {
        bool used = false;
        FBP_CHECK(exhaustive_group_launch(ctx, info, tr, laser_targets, sensor_targets, volume,
                                          stream, fixed_partial, fixed_floats, coff, &used));
        if (used) {
            ctx->last_exhaustive_group = 1;
            return FBP_OK;
        }
    }

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
    // Pairs per batch: the window prefetch registers hold one batch, and the
    // LDS of two workgroups per CU (2 x 80 KiB) holds their windows, pair
    // parameters, correction slots and band masks when the window allows.
    auto lds_bytes = [&](int b) {
        return 2 * (size_t(b) * a.W * sizeof(float2) + b * sizeof(PairParams)) +
               size_t(32) * 64 * kWaves * sizeof(float) + size_t(b) * 64 * kWaves * sizeof(uint32_t);
    };
    a.batch = 16;
    while (a.batch > 1 && (a.batch * a.W > kPrefetch * 256 || lds_bytes(a.batch) > (80u << 10)))
        --a.batch;
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
    const size_t shmem = lds_bytes(a.batch);
    // Up to 2 x 32 KiB of window entries + 32 KiB of correction slots (+ pair
    // parameters and masks): beyond the default 64 KiB for the longest windows.
    if (shmem > size_t(ctx->max_lds_per_block))
        return fail(ctx, FBP_ERR_UNSUPPORTED, "exhaustive: %zu B of LDS per workgroup needed", shmem);
    void (*kern)(ExhaustiveArgs) = pick_kernel(tr->storage, coff, link);
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
