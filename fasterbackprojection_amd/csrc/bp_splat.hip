// Confocal backprojection in scatter form ("splat"): every lit sample (l, t)
// adds I[l][t] * safeRCP(dist^2) into the voxels of its shell.
//
// Reference: DrawEllipsoids::solveBackprojection (FB/DrawEllipsoids.cpp:332-440)
// rasterises one half-ellipsoid instance per lit sample (instances built for
// I > epsilon at FB/DrawEllipsoids.cpp:29-56) into a 64^3 r32f image with
// imageAtomicAdd (FB/assets/shaders/shading/backprojection-fs.glsl:24).  That
// image depends on the raster density of the tessellated mesh; the shell here
// is the exact set of voxel centres whose bin for scan point l is t under the
// gather's arithmetic (FB/transient_reconstruction.cuh:102-133), so the volume
// is backprojectConfocalVoxel's, summed in another order.
//
// Membership without per-voxel sqrt or division: the bin is a monotone step
// function of the exact squared distance s, so voxel v lies in shell t of
// scan point l iff S_t <= s(v, l) < S_{t+1}, with S_k = min{s : bin(s) >= k}
// (the gather's threshold rows, bp_confocal.hip).
//
// Work split.  A workgroup owns a 16 x 16 x 32 voxel tile, one thread per
// (vx, vy) column of 32 depths, and a chunk of scan points.  Per scan point
// it stages the tile's window of bins in LDS (sample, 0 unless I > 1e-5;
// threshold; a bitmap and count of lit bins) while the next scan point's
// loads are in flight.  A window with no lit bin is skipped by the whole
// workgroup.  Otherwise each thread splats the lit shells into its column:
//  * sparse windows (under 1/64 of the bins lit): along a column s is
//    monotone in depth on each side of the scan point's depth, so the thread
//    walks the column once, jumping (square-root estimate, checked exactly)
//    to the first voxel of the next lit shell and adding the sample into
//    every voxel of that shell;
//  * denser windows: each voxel's shell is found directly (bin estimate
//    biased low, two threshold compares), eight independent voxels per pass,
//    unlit shells adding 0, the sums in registers;
//  * a voxel within 4.5 mm of zero distance: the walk with the reference's
//    IEEE weights (safeRCP's clamp included).
// Columns are thread-private (LDS or registers, no atomics); per-chunk
// partial volumes are summed in FP64 in chunk order, so results are
// deterministic.  Chunks are sized for two full rounds of resident
// workgroups.
#include "bp_common.h"

namespace fbp {
namespace {

constexpr int kSplatTZ = 32;        // depths per tile
constexpr int kSplatThreads = 256;  // one per column of a 16 x 16 (vx, vy) tile
constexpr int kSplatMaxCap = 4096;  // window entries (bins) per scan point

struct SplatArgs {
    const float* data;
    const fbp_vec3* laser;
    const float* thr;    // S_k rows, (T + 1) floats each
    uint32_t thr_rows;   // 1 (pure chain) or L
    float* out;          // [chunk][Z][Y][X] partials, or the volume with one chunk
    float hmin[3], vs[3];
    uint32_t res[3];     // X (vx), Y (vy), Z (depth)
    uint32_t tiles[3];   // tile counts along vx, vy, depth
    uint32_t L, T;
    float t0, dt, inv_dt;
    uint32_t discard;
    float lpos[3], spos[3];
    uint32_t chunk_len;
    int cap;             // window entries, 256 * R
};

// Window of one scan point over a tile: bins tlo .. tlo + W - 1 (W <= cap - 2).
// In LDS, wI[k + 1] holds bin k's sample (0 unless lit; wI[0] = 0 and every
// entry past W is 0) and wS[k] its threshold (+inf past W), so lookups at
// k = -1 .. W need no range checks.
struct SplatScan {
    float lx, ly, lz, off, coff;
    float qbias;  // low bias of the bin estimate, bins
    int tlo, W;
    bool exact;   // near the scan point or zero distance: IEEE weights, exact bin search
};

// Treatment of a column per scan point: kExact (walk with exact bin search
// and IEEE weights; a voxel near zero distance), kStep (splat_dense: every
// voxel's shell found directly), kJump (walk that jumps from lit shell to lit
// shell; sparse windows).
enum SplatMode { kExact = 0, kStep = 1, kJump = 2 };

__device__ __forceinline__ SplatScan splat_scan(const SplatArgs& a, uint32_t l, const float lo[3],
                                                const float hi[3]) {
    SplatScan p;
    p.lx = a.laser[l].x;
    p.ly = a.laser[l].y;
    p.lz = a.laser[l].z;
    p.off = 0.f;
    if (a.discard == 0)
        p.off = rn_add(distance3(p.lx, p.ly, p.lz, a.lpos[0], a.lpos[1], a.lpos[2]),
                       distance3(p.lx, p.ly, p.lz, a.spos[0], a.spos[1], a.spos[2]));
    p.coff = p.off - a.t0;
    const float pt[3] = {p.lx, p.ly, p.lz};
    const float rmin = box_min_dist(pt, lo, hi), rmax = box_max_dist(pt, lo, hi);
    const float dmin = 2.f * rmin + p.coff;
    const float dmax = 2.f * rmax + p.coff;
    // Estimate bias: a generous multiple of the rounding error of dist's
    // largest term (pure chain: 2^-18 relative), at most 0.9 bins so that two
    // threshold compares reach the exact bin; beyond that, or with a voxel
    // within 4.5 mm of zero distance, the exact walk runs.
    const float mag = 2.f * rmax + fabsf(a.t0) + fabsf(p.off);
    p.qbias = 32.f * 1.1920929e-7f * mag * a.inv_dt + 1.0e-5f;
    p.exact = (dmin < 0.0045f && dmax > -0.0045f) || rmin < 1.0e-3f || !(p.qbias < 0.9f);
    // Bins of the tile's distance range, 3 bins of margin for float error.
    const float qmin = fminf(fmaxf(floorf(dmin * a.inv_dt), -4.f), 1.0e9f);
    const float qmax = fminf(fmaxf(floorf(dmax * a.inv_dt), -4.f), 1.0e9f);
    const int tlo = max(int(qmin) - 3, 0);
    const int thi = min(int(qmax) + 3, int(a.T) - 1);
    // An empty window starts at bin 0, so its one threshold load stays in the row.
    p.tlo = thi >= tlo ? tlo : 0;
    p.W = thi >= tlo ? min(thi - tlo + 1, a.cap - 2) : 0;
    return p;
}

// Per scan point parameters in LDS (kPrmWords words each).
enum { Q_LX = 0, Q_LY, Q_LZ, Q_OFF, Q_COFF, Q_BIAS, Q_TLO, Q_W, Q_EXACT, Q_L, kPrmWords };

__device__ __forceinline__ void put_scan(float* q, const SplatScan& p, uint32_t l) {
    q[Q_LX] = p.lx;
    q[Q_LY] = p.ly;
    q[Q_LZ] = p.lz;
    q[Q_OFF] = p.off;
    q[Q_COFF] = p.coff;
    q[Q_BIAS] = p.qbias;
    q[Q_TLO] = __int_as_float(p.tlo);
    q[Q_W] = __int_as_float(p.W);
    q[Q_EXACT] = p.exact ? 1.f : 0.f;
    q[Q_L] = __uint_as_float(l);
}

__device__ __forceinline__ SplatScan get_scan(const float* q) {
    SplatScan p;
    p.lx = rfl(q[Q_LX]);
    p.ly = rfl(q[Q_LY]);
    p.lz = rfl(q[Q_LZ]);
    p.off = rfl(q[Q_OFF]);
    p.coff = rfl(q[Q_COFF]);
    p.qbias = rfl(q[Q_BIAS]);
    p.tlo = rfl(__float_as_int(q[Q_TLO]));
    p.W = rfl(__float_as_int(q[Q_W]));
    p.exact = rfl(q[Q_EXACT]) != 0.f;
    return p;
}

// Loads of one batch: round u of the thread is entry k = tid + 256 (u % R) of
// scan u / R's window (the sample, 0 past W, and the threshold, 0 past W + 1).
template <int LAYOUT, int R, int NU>
__device__ __forceinline__ void splat_load(const SplatArgs& a, const float* prm, int tid,
                                           float (&vI)[NU], float (&vS)[NU]) {
#pragma unroll
    for (int u = 0; u < NU; ++u) {
        const float* q = prm + (u / R) * kPrmWords;
        const int W = __float_as_int(q[Q_W]);
        const uint32_t l = __float_as_uint(q[Q_L]);
        const int k = tid + kSplatThreads * (u % R);
        const uint32_t t = uint32_t(__float_as_int(q[Q_TLO]) + k);
        const size_t iI = k < W ? (LAYOUT == FBP_LAYOUT_TIME_FASTEST ? size_t(l) * a.T + t
                                                                      : size_t(t) * a.L + l)
                                : 0;
        const size_t iS = k <= W ? size_t(a.thr_rows > 1 ? l : 0) * (a.T + 1) + t : 0;
        vI[u] = a.data[iI];
        vS[u] = a.thr[iS];
    }
}

// Stores a batch into the window (entry u * 256 + tid), the lit bitmap (word
// 4u + wave) and the lit count of each wave's 64 entries (cnt[u][wave]).
template <int R, int NU>
__device__ __forceinline__ void splat_store(const float* prm, int tid, const float (&vI)[NU],
                                            const float (&vS)[NU], float* wI, float* wS,
                                            uint64_t* bits, int* cnt) {
    const int wave = tid >> 6, lane = tid & 63;
#pragma unroll
    for (int u = 0; u < NU; ++u) {
        const int W = __float_as_int(prm[(u / R) * kPrmWords + Q_W]);
        const int k = tid + kSplatThreads * (u % R);
        const bool lit = k < W && vI[u] > kEps;
        const int base = (u / R) * (kSplatThreads * R);
        wI[base + min(k + 1, kSplatThreads * R - 1)] = lit ? vI[u] : 0.f;
        if (k == 0) wI[base] = 0.f;
        wS[base + k] = k <= W ? vS[u] : __int_as_float(0x7f800000);
        const uint64_t m = __ballot(lit);
        if (lane == 0) {
            bits[4 * u + wave] = m;
            cnt[4 * u + wave] = __builtin_popcountll(m);
        }
    }
}

// One monotone run of a column: depths zs, zs + dir, ... (n of them), along
// which s increases.  d0 = |depth centre of zs - scan point depth|.
template <bool PURE, int MODE>
__device__ __forceinline__ void splat_segment(const SplatArgs& a, const SplatScan& p, int zs,
                                              int dir, int n, float d0, float inv_vsy, float X2,
                                              float Z2, const float* cyl, const float* wI,
                                              const float* wS, const uint64_t* bits, float* acc,
                                              int tid) {
    const int W = p.W;
    auto s_at = [&](int pos) {
        return rn_add(rn_add(X2, sq(rn_sub(cyl[zs + dir * pos], p.ly))), Z2);
    };
    // Window bin k with S[k] <= s < S[k + 1]; -1 below the window, W above.
    auto find_bin = [&](float s) {
        const float rt = __builtin_amdgcn_sqrtf(s);
        if constexpr (MODE == kExact) {
            const float q = PURE ? (rt * 2.f) * a.inv_dt : fmaf(rt, 2.f, p.coff) * a.inv_dt;
            int k = int(fminf(fmaxf(floorf(q), -4.f), 1.0e9f)) - p.tlo;
            k = min(max(k, -1), W);
            while (k < W && wS[k + 1] <= s) ++k;
            while (k >= 0 && wS[k] > s) --k;
            return k;
        } else {
            // Biased low by at most ~2 bins: two compares, no loop.
            const float q = PURE ? (rt * 2.f) * (a.inv_dt * (1.0f - 3.814697265625e-06f))
                                 : fmaf(fmaf(rt, 2.f, p.coff), a.inv_dt, -p.qbias);
            int k = int(fminf(fmaxf(floorf(q), -4.f), 1.0e9f)) - p.tlo;
            k = min(max(k, -1), W - 1);
            // Low by at most two bins; the two thresholds load together.
            const float S1 = wS[k + 1], S2 = wS[k + 2];
            return k + int(S1 <= s) + int(S2 <= s);
        }
    };
    // First lit bin >= start, or W; m = the bits of word wm from bin start on
    // (start may be the first bin of word wm + 1, m then 0).
    auto next_lit = [&](int start, uint64_t m, int wm) {
        if (m != 0) return min(start + int(__builtin_ctzll(m)), W);
        const int nw = (W + 63) >> 6;
        for (int w = wm + 1; w < nw; ++w) {
            const uint64_t b = bits[w];
            if (b != 0) return min((w << 6) + int(__builtin_ctzll(b)), W);
        }
        return W;
    };
    // A pos in [p0, n] at or before the first voxel of [p0, n) with s >= Sa:
    // an estimate biased one voxel low, checked exactly (else p0).
    auto jump = [&](int p0, float Sa) {
        const float r2 = Sa - X2 - Z2;
        const float r = r2 > 0.f ? __builtin_amdgcn_sqrtf(r2) : 0.f;
        const float e = fminf(fmaxf(floorf((r - d0) * inv_vsy) - 1.f, -1.f), float(n));
        int q = min(max(int(e), p0), n);
        if constexpr (MODE == kExact) {
            while (q > p0 && s_at(q - 1) >= Sa) --q;
        } else {
            q = (q > p0 && s_at(q - 1) >= Sa) ? p0 : q;
        }
        return q;
    };
    auto weight = [&](float s) {
        if constexpr (MODE == kExact) {
            const float d = confocal_dist(s, a.t0, a.discard, p.off);
            return safe_rcp(rn_mul(d, d));
        } else if constexpr (PURE) {
            return __builtin_amdgcn_rcpf(s * 4.f);
        } else {
            const float d = fmaf(__builtin_amdgcn_sqrtf(s), 2.f, p.coff);
            return __builtin_amdgcn_rcpf(d * d);
        }
    };

    {
        // One voxel per pass: its lit shell's contribution, or a jump to the
        // next lit shell.
        int pos = 0;
        while (pos < n) {
            const float s = s_at(pos);
            const int k = find_bin(s);
            if (k >= W) break;
            const int kk = max(k, 0);
            const uint64_t wd = bits[kk >> 6] >> (kk & 63);
            if (k >= 0 && (wd & 1u)) {
                float* c = acc + (zs + dir * pos) * kSplatThreads + tid;
                *c = fmaf(wI[kk + 1], weight(s), *c);
                ++pos;
                continue;
            }
            // Below the window (k = -1) bin tlo itself may be lit.
            const int t = k < 0 ? next_lit(0, wd, 0) : next_lit(k + 1, wd >> 1, kk >> 6);
            if (t >= W) break;
            pos = jump(pos + 1, wS[t]);
        }
    }
}

// Dense windows: every voxel of the column in turn, eight independent
// chains per pass, unlit bins (stored as 0) adding 0.  The column's sums stay
// in registers (racc; the other walks add into the LDS column).
template <bool PURE>
__device__ __forceinline__ void splat_dense(const SplatArgs& a, const SplatScan& p, int nz,
                                            const float* cyl, float X2, float Z2,
                                            const float* wI, const float* wS,
                                            float (&racc)[kSplatTZ]) {
    const int W = p.W;
    const float c2 = PURE ? 2.f * (a.inv_dt * (1.0f - 3.814697265625e-06f)) : 2.f * a.inv_dt;
    // Window bin estimate q - tlo in one FMA.
    const float c0 = (PURE ? 0.f : fmaf(p.coff, a.inv_dt, -p.qbias)) - float(p.tlo);
    constexpr int U = 8;
#pragma unroll
    for (int z0 = 0; z0 < kSplatTZ; z0 += U) {
        if (z0 >= nz) break;
        float sv[U], S1[U], S2[U];
        int ka[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            // s in glm::dot's order (x^2 + y^2) + z^2; y is the depth axis.
            sv[u] = rn_add(rn_add(X2, sq(rn_sub(cyl[z0 + u], p.ly))), Z2);
            const int k = int(floorf(fmaf(__builtin_amdgcn_sqrtf(sv[u]), c2, c0)));
            ka[u] = min(max(k, -1), W - 1);
            S1[u] = wS[ka[u] + 1];
            S2[u] = wS[ka[u] + 2];
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            // The estimate is low by at most two bins; thresholds are
            // monotone (+inf past W).  Depths past nz are never written out.
            const int k = ka[u] + int(S1[u] <= sv[u]) + int(S2[u] <= sv[u]);
            const float I = wI[k + 1];
            float w;
            if constexpr (PURE) {
                w = __builtin_amdgcn_rcpf(sv[u] * 4.f);
            } else {
                const float d = fmaf(__builtin_amdgcn_sqrtf(sv[u]), 2.f, p.coff);
                w = __builtin_amdgcn_rcpf(d * d);
            }
            racc[z0 + u] = fmaf(I, w, racc[z0 + u]);
        }
    }
}

// R = window rounds per scan point (cap = 256 R), B = scan points per batch.
template <int LAYOUT, bool PURE, int R, int B>
__global__ __launch_bounds__(kSplatThreads) void splat_kernel(SplatArgs a) {
    constexpr int cap = kSplatThreads * R;
    constexpr int NU = R * B;  // load rounds per batch
    extern __shared__ float smem[];
    const int tid = threadIdx.x;
    float* acc = smem;                                   // [TZ][256]
    float* cyl = acc + kSplatTZ * kSplatThreads;         // depth centres of the tile
    float* wI = cyl + kSplatTZ;                          // [B][cap] samples
    float* wS = wI + B * cap;                            // [B][cap] thresholds
    uint64_t* bits = reinterpret_cast<uint64_t*>(wS + B * cap);  // [B][cap / 64]
    int* cnt = reinterpret_cast<int*>(bits + B * cap / 64);      // [NU][4]
    float* prm = reinterpret_cast<float*>(cnt + 4 * NU);         // [2][B][kPrmWords]

    const uint32_t ntiles = a.tiles[0] * a.tiles[1] * a.tiles[2];
    const uint32_t tile = blockIdx.x % ntiles, chunk = blockIdx.x / ntiles;
    const uint32_t tx = tile % a.tiles[0], ty = (tile / a.tiles[0]) % a.tiles[1];
    const uint32_t tz = tile / (a.tiles[0] * a.tiles[1]);
    const uint32_t X0 = tx * 16, Y0 = ty * 16, Z0 = tz * kSplatTZ;
    const uint32_t vx = X0 + (tid & 15), vy = Y0 + (tid >> 4);
    const bool col_ok = vx < a.res[0] && vy < a.res[1];
    const int nz = int(min(uint32_t(kSplatTZ), a.res[2] - Z0));
    const uint32_t X1 = min(X0 + 15, a.res[0] - 1), Y1 = min(Y0 + 15, a.res[1] - 1);

    // World coordinates: vx -> x, depth -> y, vy -> z (FB/transient_reconstruction.cuh:106).
    const float cx = voxel_center(a.hmin[0], a.vs[0], min(vx, a.res[0] - 1));
    const float cz = voxel_center(a.hmin[2], a.vs[2], min(vy, a.res[1] - 1));
    if (tid < kSplatTZ) cyl[tid] = voxel_center(a.hmin[1], a.vs[1], Z0 + min(tid, nz - 1));
    float racc[kSplatTZ];
#pragma unroll
    for (int z = 0; z < kSplatTZ; ++z) racc[z] = 0.f;
#pragma unroll 8
    for (int z = 0; z < kSplatTZ; ++z) acc[z * kSplatThreads + tid] = 0.f;
    float lo[3], hi[3];
    {
        const float e[3][2] = {
            {voxel_center(a.hmin[0], a.vs[0], X0), voxel_center(a.hmin[0], a.vs[0], X1)},
            {voxel_center(a.hmin[1], a.vs[1], Z0), voxel_center(a.hmin[1], a.vs[1], Z0 + nz - 1)},
            {voxel_center(a.hmin[2], a.vs[2], Y0), voxel_center(a.hmin[2], a.vs[2], Y1)}};
        for (int ax = 0; ax < 3; ++ax) {
            lo[ax] = fminf(e[ax][0], e[ax][1]);
            hi[ax] = fmaxf(e[ax][0], e[ax][1]);
        }
    }
    const float vsy = fabsf(a.vs[1]);
    const float inv_vsy = vsy > 0.f ? 1.f / vsy : 0.f;

    const uint32_t l_begin = chunk * a.chunk_len;
    const uint32_t l_end = min(a.L, l_begin + a.chunk_len);
    const uint32_t nbatch = (l_end > l_begin) ? (l_end - l_begin + B - 1) / B : 0;
    // Parameters of batch i live in prm[i & 1]; scan points past the chunk get
    // an empty window.
    auto scan_params = [&](uint32_t i) {
        if (tid < B) {
            const uint32_t l = l_begin + i * B + uint32_t(tid);
            SplatScan p{};
            if (l < l_end) p = splat_scan(a, l, lo, hi);
            put_scan(prm + ((i & 1) * B + tid) * kPrmWords, p, l < l_end ? l : 0u);
        }
    };
    float vI[NU], vS[NU];
    if (nbatch > 0) {
        scan_params(0);
        __syncthreads();
        splat_load<LAYOUT, R, NU>(a, prm, tid, vI, vS);
    }
    for (uint32_t i = 0; i < nbatch; ++i) {
        const float* cur = prm + (i & 1) * B * kPrmWords;
        splat_store<R, NU>(cur, tid, vI, vS, wI, wS, bits, cnt);
        if (i + 1 < nbatch) scan_params(i + 1);
        __syncthreads();
        if (i + 1 < nbatch)
            splat_load<LAYOUT, R, NU>(a, prm + ((i + 1) & 1) * B * kPrmWords, tid, vI, vS);
        if (col_ok) {
            for (int j = 0; j < B; ++j) {
                const SplatScan p = get_scan(cur + j * kPrmWords);
                int nlit = 0;
#pragma unroll
                for (int u = 0; u < 4 * R; ++u) nlit += cnt[4 * R * j + u];
                if (p.W <= 0 || nlit == 0) continue;  // no lit shell reaches the tile
                const float* jI = wI + j * cap;
                const float* jS = wS + j * cap;
                const uint64_t* jb = bits + j * (cap / 64);
                const float X2 = sq(rn_sub(cx, p.lx));
                const float Z2 = sq(rn_sub(cz, p.lz));
                const int mode = p.exact ? kExact : (64 * nlit >= p.W ? kStep : kJump);
                if (mode == kStep) {
                    splat_dense<PURE>(a, p, nz, cyl, X2, Z2, jI, jS, racc);
                    continue;
                }
                // zsplit = first depth whose centre is at or beyond the scan
                // point's depth; s increases away from it on both sides.
                int zsplit = 0;
                if (p.ly > cyl[0]) {
                    int lo_i = 1, hi_i = nz;
                    while (lo_i < hi_i) {
                        const int mid = (lo_i + hi_i) >> 1;
                        if (cyl[mid] >= p.ly)
                            hi_i = mid;
                        else
                            lo_i = mid + 1;
                    }
                    zsplit = lo_i;
                }
                auto run = [&](int zs, int dir, int n, float d0) {
                    if (mode == kJump)
                        splat_segment<PURE, kJump>(a, p, zs, dir, n, d0, inv_vsy, X2, Z2, cyl,
                                                   jI, jS, jb, acc, tid);
                    else
                        splat_segment<PURE, kExact>(a, p, zs, dir, n, d0, inv_vsy, X2, Z2, cyl,
                                                    jI, jS, jb, acc, tid);
                };
                if (zsplit < nz) run(zsplit, 1, nz - zsplit, cyl[zsplit] - p.ly);
                if (zsplit > 0) run(zsplit - 1, -1, zsplit, p.ly - cyl[zsplit - 1]);
            }
        }
        __syncthreads();
    }

    if (!col_ok) return;
    const size_t V = size_t(a.res[0]) * a.res[1] * a.res[2];
    float* out = a.out + size_t(chunk) * V;
#pragma unroll
    for (int z = 0; z < kSplatTZ; ++z)
        if (z < nz)
            out[(size_t(Z0 + z) * a.res[1] + vy) * a.res[0] + vx] =
                racc[z] + acc[z * kSplatThreads + tid];
}

// volume[i] = sum over chunks of partial[c][i], in FP64, chunk order.
__global__ void splat_reduce_kernel(const float* __restrict__ partial, float* __restrict__ volume,
                                    size_t V, uint32_t chunks) {
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= V) return;
    double sum = 0.0;
    constexpr uint32_t B = 8;
    for (uint32_t c0 = 0; c0 < chunks; c0 += B) {
        float v[B];
#pragma unroll
        for (uint32_t u = 0; u < B; ++u) v[u] = partial[size_t(min(c0 + u, chunks - 1)) * V + i];
#pragma unroll
        for (uint32_t u = 0; u < B; ++u)
            if (c0 + u < chunks) sum += double(v[u]);
    }
    volume[i] = float(sum);
}

template <int R, int B>
constexpr size_t splat_shmem() {
    return (size_t(kSplatTZ) * kSplatThreads + kSplatTZ) * sizeof(float) +
           size_t(B) * kSplatThreads * R * 2 * sizeof(float) +
           size_t(B) * kSplatThreads * R / 64 * sizeof(uint64_t) + size_t(4) * R * B * sizeof(int) +
           size_t(2) * B * kPrmWords * sizeof(float);
}

// One scan point per staging round: deeper batches cost occupancy (LDS) and
// measured slower.
constexpr int kSplatBatch = 1;

// Sets the LDS limit of the instance for cap and returns its resident
// workgroups per CU (0 on error).
template <int LAYOUT, bool PURE, int R>
int prepare_splat_r() {
    constexpr size_t shmem = splat_shmem<R, kSplatBatch>();
    const void* k = reinterpret_cast<const void*>(&splat_kernel<LAYOUT, PURE, R, kSplatBatch>);
    if (shmem > 65536 &&
        hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, int(shmem)) !=
            hipSuccess)
        return 0;
    int n = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &n, splat_kernel<LAYOUT, PURE, R, kSplatBatch>, kSplatThreads, shmem) != hipSuccess)
        return 0;
    return n;
}

template <int LAYOUT, bool PURE, int R>
void launch_splat_r(const SplatArgs& a, dim3 grid, hipStream_t s) {
    constexpr size_t shmem = splat_shmem<R, kSplatBatch>();
    hipLaunchKernelGGL((splat_kernel<LAYOUT, PURE, R, kSplatBatch>), grid, dim3(kSplatThreads),
                       shmem, s, a);
}

template <int LAYOUT, bool PURE>
int prepare_splat(int cap) {
    switch (cap / kSplatThreads) {
        case 1: return prepare_splat_r<LAYOUT, PURE, 1>();
        case 2: return prepare_splat_r<LAYOUT, PURE, 2>();
        case 4: return prepare_splat_r<LAYOUT, PURE, 4>();
        case 8: return prepare_splat_r<LAYOUT, PURE, 8>();
        default: return prepare_splat_r<LAYOUT, PURE, 16>();
    }
}

template <int LAYOUT, bool PURE>
void launch_splat(const SplatArgs& a, dim3 grid, hipStream_t s) {
    switch (a.cap / kSplatThreads) {
        case 1: return launch_splat_r<LAYOUT, PURE, 1>(a, grid, s);
        case 2: return launch_splat_r<LAYOUT, PURE, 2>(a, grid, s);
        case 4: return launch_splat_r<LAYOUT, PURE, 4>(a, grid, s);
        case 8: return launch_splat_r<LAYOUT, PURE, 8>(a, grid, s);
        default: return launch_splat_r<LAYOUT, PURE, 16>(a, grid, s);
    }
}

}  // namespace
}  // namespace fbp

using namespace fbp;

extern "C" fbp_status fbp_backproject_confocal_splat(fbp_context ctx,
                                                     const fbp_reconstruction_info* info,
                                                     const fbp_transient* tr,
                                                     const fbp_vec3* laser_targets, float* volume,
                                                     void* stream) {
    if (!ctx || !info || !tr || !tr->data || !laser_targets || !volume)
        return fail(ctx, FBP_ERR_INVALID_ARGUMENT, "splat: null argument");
    if (info->num_laser_targets == 0 || info->num_time_bins == 0 || info->voxel_resolution.x == 0 ||
        info->voxel_resolution.y == 0 || info->voxel_resolution.z == 0)
        return fail(ctx, FBP_ERR_INVALID_ARGUMENT, "splat: zero-sized grid");
    if (!(info->time_step > 0.f) || !finite_info(*info))
        return fail(ctx, FBP_ERR_INVALID_ARGUMENT, "splat: time_step must be > 0, fields finite");
    if (tr->storage != FBP_STORAGE_F32)
        return fail(ctx, FBP_ERR_UNSUPPORTED, "splat: F32 storage only");
    if (tr->layout != FBP_LAYOUT_TIME_FASTEST && tr->layout != FBP_LAYOUT_TIME_MAJOR)
        return fail(ctx, FBP_ERR_INVALID_ARGUMENT, "splat: bad layout %d", tr->layout);
    if (info->num_time_bins > (1u << 17))
        return fail(ctx, FBP_ERR_UNSUPPORTED, "splat: more than 2^17 time bins");
    if (!(info->hidden_volume_voxel_size.y >= 0.f))
        return fail(ctx, FBP_ERR_UNSUPPORTED,
                    "splat: the depth voxel size must be >= 0 (depth centres in order)");
    if (ctx->voxel_scattering)
        return fail(ctx, FBP_ERR_UNSUPPORTED,
                    "splat: the voxel-scattering option has no scatter form; use "
                    "fbp_backproject_confocal_voxel");
    FBP_CHECK(bind(ctx));

    SplatArgs a{};
    a.data = static_cast<const float*>(tr->data);
    a.laser = laser_targets;
    a.hmin[0] = info->hidden_volume_min.x;
    a.hmin[1] = info->hidden_volume_min.y;
    a.hmin[2] = info->hidden_volume_min.z;
    a.vs[0] = info->hidden_volume_voxel_size.x;
    a.vs[1] = info->hidden_volume_voxel_size.y;
    a.vs[2] = info->hidden_volume_voxel_size.z;
    a.res[0] = info->voxel_resolution.x;
    a.res[1] = info->voxel_resolution.y;
    a.res[2] = info->voxel_resolution.z;
    a.tiles[0] = (a.res[0] + 15) / 16;
    a.tiles[1] = (a.res[1] + 15) / 16;
    a.tiles[2] = (a.res[2] + kSplatTZ - 1) / kSplatTZ;
    a.L = info->num_laser_targets;
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

    // Window bound: the dist range over a tile's centre box is at most twice
    // its diagonal; the kernel adds 3 bins of margin on each side.
    const double ex = double(std::min(a.res[0], 16u) - 1) * std::fabs(double(a.vs[0]));
    const double ey = double(std::min(a.res[2], uint32_t(kSplatTZ)) - 1) * std::fabs(double(a.vs[1]));
    const double ez = double(std::min(a.res[1], 16u) - 1) * std::fabs(double(a.vs[2]));
    const double wbins = std::ceil(2.0 * std::sqrt(ex * ex + ey * ey + ez * ez) / a.dt) + 16.0;
    if (!(wbins <= double(kSplatMaxCap - 2)))
        return fail(ctx, FBP_ERR_UNSUPPORTED,
                    "splat: a tile spans %.0f bins (more than %d): voxels too large for the bin "
                    "width",
                    wbins, kSplatMaxCap - 2);
    a.cap = kSplatThreads;
    while (a.cap < int(wbins) + 2) a.cap *= 2;

    const bool pure = (a.t0 == 0.f) && (a.discard != 0);
    const bool major = tr->layout == FBP_LAYOUT_TIME_MAJOR;
    const int per_cu = major ? (pure ? prepare_splat<FBP_LAYOUT_TIME_MAJOR, true>(a.cap)
                                     : prepare_splat<FBP_LAYOUT_TIME_MAJOR, false>(a.cap))
                             : (pure ? prepare_splat<FBP_LAYOUT_TIME_FASTEST, true>(a.cap)
                                     : prepare_splat<FBP_LAYOUT_TIME_FASTEST, false>(a.cap));
    if (per_cu < 1) return fail(ctx, FBP_ERR_DEVICE, "splat: kernel set-up failed");

    const uint32_t ntiles = a.tiles[0] * a.tiles[1] * a.tiles[2];
    const size_t V = size_t(a.res[0]) * a.res[1] * a.res[2];
    // Two full rounds of resident workgroups (equal chunks finish together),
    // at least 32 scan points per chunk, at most 512 MiB of partial volumes.
    const uint32_t slots = uint32_t(per_cu) * uint32_t(ctx->num_cus > 0 ? ctx->num_cus : 256);
    uint32_t nc = (2 * slots) / ntiles;
    nc = std::min(nc, (a.L + 31) / 32);
    nc = std::min<uint64_t>(nc, std::max<uint64_t>(1, (size_t(512) << 20) / (V * 4)));
    nc = std::max(nc, 1u);
    a.chunk_len = (a.L + nc - 1) / nc;
    const uint32_t chunks = (a.L + a.chunk_len - 1) / a.chunk_len;

    const size_t thr_n = pure ? 0 : size_t(a.L) * (a.T + 1);
    const size_t part_n = chunks > 1 ? size_t(chunks) * V : 0;
    FBP_CHECK(ensure_scratch(ctx, (part_n + thr_n) * sizeof(float) + 256));
    float* part = static_cast<float*>(ctx->scratch);
    hipStream_t s = as_stream(stream);
    if (pure) {
        FBP_CHECK(pure_threshold_row(ctx, a.T, a.dt, s, &a.thr));
        a.thr_rows = 1;
    } else {
        float* thr = part + part_n;
        FBP_CHECK(launch_confocal_thresholds(ctx, laser_targets, a.L, a.T, a.t0, a.dt, a.discard,
                                             info->laser_position, info->sensor_position, thr, s));
        a.thr = thr;
        a.thr_rows = a.L;
    }
    a.out = chunks > 1 ? part : volume;
    const dim3 grid(chunks * ntiles);
    if (major) {
        if (pure)
            launch_splat<FBP_LAYOUT_TIME_MAJOR, true>(a, grid, s);
        else
            launch_splat<FBP_LAYOUT_TIME_MAJOR, false>(a, grid, s);
    } else {
        if (pure)
            launch_splat<FBP_LAYOUT_TIME_FASTEST, true>(a, grid, s);
        else
            launch_splat<FBP_LAYOUT_TIME_FASTEST, false>(a, grid, s);
    }
    FBP_CHECK_HIP(ctx, hipGetLastError());
    if (chunks > 1) {
        hipLaunchKernelGGL(splat_reduce_kernel, dim3(unsigned((V + 255) / 256)), dim3(256), 0, s,
                           part, volume, V, chunks);
        FBP_CHECK_HIP(ctx, hipGetLastError());
    }
    return FBP_OK;
}
