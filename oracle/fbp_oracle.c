/*
 * fbp_oracle.c - CPU oracle of the NLOS reconstruction hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: tests/, the smoke
 * check and bench.py's CPU baseline may load it; the product library never
 * does.  It restates the reference's formulas as plainly as possible:
 *
 *   - bins and validity: the reference's FP32 expression evaluated as
 *     separately rounded IEEE operations (compiled with -ffp-contract=off,
 *     SSE2 scalar), i.e. the specification the GPU must match bit for bit;
 *   - weights and sums: FP64 (1 / dist^2 from the FP32 dist, summed in double).
 *
 * Parity status per function is stated in the header of each block and in
 * the test that pins it.  Paths are relative to the reference repository;
 * FB/ = FasterBackprojection/.
 */
#include <float.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include <omp.h>

typedef struct { float x, y, z; } vec3;
typedef struct { uint32_t x, y, z; } uvec3;

/* FB/GpuStructs.cuh:8-38, 148 bytes. */
typedef struct {
    vec3 sensor_position;
    uint32_t capture_system;
    vec3 laser_position;
    float time_offset;
    uint32_t num_laser_targets, num_sensor_targets, num_time_bins;
    float time_step;
    vec3 hidden_volume_min;
    uint32_t discard_first_last_bounces;
    vec3 hidden_volume_max, hidden_volume_size, hidden_volume_voxel_size;
    vec3 relay_wall_min_position, relay_wall_size, relay_wall_normal;
    uvec3 voxel_resolution;
} RecInfo;

#define OR_EPS 0.00001f /* FB/math.cuh:11 */

int oracle_recinfo_size(void) { return (int)sizeof(RecInfo); }

/* ------------------------------------------------------------ FP32 chain */

/* (hmin + float(i) * vs) + 0.5f * vs   (FB/transient_reconstruction.cuh:106) */
static float vcenter(float hmin, float vs, uint32_t i) {
    float a = (float)i * vs;
    float b = hmin + a;
    float c = 0.5f * vs;
    return b + c;
}

/* glm::distance(a, b) = sqrt((dx*dx + dy*dy) + dz*dz), d = b - a. */
static float glm_distance(vec3 a, vec3 b) {
    float dx = b.x - a.x, dy = b.y - a.y, dz = b.z - a.z;
    float x2 = dx * dx, y2 = dy * dy, z2 = dz * dz;
    float s = (x2 + y2) + z2;
    return sqrtf(s);
}

/* safeRCP in double (FB/math.cuh:23-29 semantics). */
static double safe_rcp64(double x) {
    if (x > 1e-5 || x < -1e-5) return 1.0 / x;
    return x >= 0 ? (double)FLT_MAX : -(double)FLT_MAX;
}

/* Confocal dist (FB/transient_reconstruction.cuh:108-110). */
static float confocal_dist(const RecInfo* in, vec3 l, vec3 p) {
    float r = glm_distance(l, p);
    float d = r * 2.0f;
    d = d - in->time_offset;
    if (in->discard_first_last_bounces == 0) {
        float a = glm_distance(l, in->laser_position);
        float b = glm_distance(l, in->sensor_position);
        float o = a + b;
        d = d + o;
    }
    return d;
}

/* Exhaustive dist (FB/transient_reconstruction.cuh:155-157). */
static float exhaustive_dist(const RecInfo* in, vec3 l, vec3 s, vec3 p) {
    float a = glm_distance(l, p);
    float b = glm_distance(s, p);
    float d = a + b;
    d = d - in->time_offset;
    if (in->discard_first_last_bounces == 0) {
        float c = glm_distance(l, in->laser_position);
        float e = glm_distance(s, in->sensor_position);
        float o = c + e;
        d = d + o;
    }
    return d;
}

/* (int)(dist / Delta); returns -1 when out of [0, T). */
static int64_t bin_of(float dist, float dt, uint32_t T) {
    float q = dist / dt;
    if (!(q > -1.0f) || !(q < (float)T)) return -1;
    return (int64_t)(int)q;
}

static vec3 voxel_pos(const RecInfo* in, uint32_t vx, uint32_t vy, uint32_t z) {
    vec3 p;
    p.x = vcenter(in->hidden_volume_min.x, in->hidden_volume_voxel_size.x, vx);
    p.y = vcenter(in->hidden_volume_min.y, in->hidden_volume_voxel_size.y, z);
    p.z = vcenter(in->hidden_volume_min.z, in->hidden_volume_voxel_size.z, vy);
    return p;
}

/* Bin of one (voxel, scan point) pair, -1 when invalid (exposed for tests). */
int64_t oracle_confocal_bin(const RecInfo* in, const vec3* laser, uint32_t l, uint32_t vx,
                            uint32_t vy, uint32_t z) {
    vec3 p = voxel_pos(in, vx, vy, z);
    return bin_of(confocal_dist(in, laser[l], p), in->time_step, in->num_time_bins);
}

float oracle_confocal_distf(const RecInfo* in, const vec3* laser, uint32_t l, uint32_t vx,
                            uint32_t vy, uint32_t z) {
    return confocal_dist(in, laser[l], voxel_pos(in, vx, vy, z));
}

/* --------------------------------------------------- confocal backprojection
 * backprojectConfocalVoxel (FB/transient_reconstruction.cuh:93-136).
 * layout 0: I[l][t]; layout 1: I[t][l] (getConfocalTransientIndex).
 * vol: double[Z][Y][X], overwritten. */
void oracle_confocal_voxel(const RecInfo* in, const float* I, int layout, const vec3* laser,
                           double* vol) {
    const uint32_t X = in->voxel_resolution.x, Y = in->voxel_resolution.y,
                   Z = in->voxel_resolution.z;
    const uint32_t L = in->num_laser_targets, T = in->num_time_bins;
    const int64_t n = (int64_t)X * Y * Z;
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t v = 0; v < n; ++v) {
        uint32_t vx = (uint32_t)(v % X), vy = (uint32_t)((v / X) % Y), z = (uint32_t)(v / ((int64_t)X * Y));
        vec3 p = voxel_pos(in, vx, vy, z);
        double acc = 0.0;
        for (uint32_t l = 0; l < L; ++l) {
            float d = confocal_dist(in, laser[l], p);
            int64_t b = bin_of(d, in->time_step, T);
            if (b < 0) continue;
            size_t idx = layout == 0 ? (size_t)l * T + (size_t)b : (size_t)b * L + l;
            float light = I[idx];
            if (light > OR_EPS) {
                float d2 = d * d;
                acc += (double)light * safe_rcp64((double)d2);
            }
        }
        vol[v] = acc;
    }
}

/* ------------------------------------------- voxel-scattering variant
 * The ACCUMULATE_VOXEL_SCATTERING branch of backprojectConfocalVoxel
 * (FB/transient_reconstruction.cuh:112-126): e = length(voxel size) / 2,
 * d = max(0, dist - e); while d < dist + e: bin = (int)(d / Delta), a bin
 * outside [0, T) ends the pair, else add I(l, bin) * safeRCP(dist^2) when
 * I > EPS; d += Delta + EPS.  At most T + 2 steps (a walk whose step falls
 * below the float spacing at d would not advance).  Returns the number of
 * bins written to bins[]. */
static float voxel_extent(const RecInfo* in) {
    vec3 v = in->hidden_volume_voxel_size;
    float x2 = v.x * v.x, y2 = v.y * v.y, z2 = v.z * v.z;
    float s = (x2 + y2) + z2;
    return sqrtf(s) / 2.0f;
}

static int scatter_bins(float dist, float ext, float dt, uint32_t T, int64_t* bins) {
    float lo = dist - ext;
    float d = (0.0f < lo) ? lo : 0.0f;
    float hi = dist + ext;
    float step = dt + OR_EPS;
    int n = 0;
    for (uint32_t k = 0; k < T + 2u && d < hi; ++k) {
        float q = d / dt;
        if (!(q < (float)T)) break;
        bins[n++] = (int64_t)(int)q;
        d = d + step;
    }
    return n;
}

/* Bins of the walk of one pair (exposed for tests); returns the count. */
int oracle_scatter_bins(const RecInfo* in, const vec3* laser, uint32_t l, uint32_t vx,
                        uint32_t vy, uint32_t z, int64_t* bins) {
    float d = confocal_dist(in, laser[l], voxel_pos(in, vx, vy, z));
    return scatter_bins(d, voxel_extent(in), in->time_step, in->num_time_bins, bins);
}

void oracle_confocal_voxel_scatter(const RecInfo* in, const float* I, int layout,
                                   const vec3* laser, double* vol) {
    const uint32_t X = in->voxel_resolution.x, Y = in->voxel_resolution.y,
                   Z = in->voxel_resolution.z;
    const uint32_t L = in->num_laser_targets, T = in->num_time_bins;
    const float ext = voxel_extent(in);
    const int64_t n = (int64_t)X * Y * Z;
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t v = 0; v < n; ++v) {
        uint32_t vx = (uint32_t)(v % X), vy = (uint32_t)((v / X) % Y), z = (uint32_t)(v / ((int64_t)X * Y));
        vec3 p = voxel_pos(in, vx, vy, z);
        int64_t* bins = (int64_t*)malloc(sizeof(int64_t) * (T + 2));
        double acc = 0.0;
        for (uint32_t l = 0; l < L; ++l) {
            float d = confocal_dist(in, laser[l], p);
            int nb = scatter_bins(d, ext, in->time_step, T, bins);
            double w = safe_rcp64((double)(d * d));
            for (int j = 0; j < nb; ++j) {
                size_t idx = layout == 0 ? (size_t)l * T + (size_t)bins[j] : (size_t)bins[j] * L + l;
                float light = I[idx];
                if (light > OR_EPS) acc += (double)light * w;
            }
        }
        free(bins);
        vol[v] = acc;
    }
}

/* ------------------------------------------------- exhaustive backprojection
 * backprojectExhaustiveVoxel (FB/transient_reconstruction.cuh:138-166) with
 * every pair visited, l = ls / S, s = ls % S. */
void oracle_exhaustive_voxel(const RecInfo* in, const float* I, const vec3* laser,
                             const vec3* sensor, double* vol) {
    const uint32_t X = in->voxel_resolution.x, Y = in->voxel_resolution.y,
                   Z = in->voxel_resolution.z;
    const uint32_t L = in->num_laser_targets, S = in->num_sensor_targets, T = in->num_time_bins;
    const int64_t n = (int64_t)X * Y * Z;
#pragma omp parallel for schedule(dynamic, 16)
    for (int64_t v = 0; v < n; ++v) {
        uint32_t vx = (uint32_t)(v % X), vy = (uint32_t)((v / X) % Y), z = (uint32_t)(v / ((int64_t)X * Y));
        vec3 p = voxel_pos(in, vx, vy, z);
        double acc = 0.0;
        for (uint32_t l = 0; l < L; ++l)
            for (uint32_t s = 0; s < S; ++s) {
                float d = exhaustive_dist(in, laser[l], sensor[s], p);
                int64_t b = bin_of(d, in->time_step, T);
                if (b < 0) continue;
                float light = I[((size_t)l * S + s) * T + (size_t)b];
                if (light > OR_EPS) {
                    float d2 = d * d;
                    acc += (double)light * safe_rcp64((double)d2);
                }
            }
        vol[v] = acc;
    }
}

/* Bins (-1 when outside [0, T)) and FP32 dist of every voxel for one scan
 * point (confocal) or one (laser, sensor) pair (exhaustive); [Z][Y][X]. */
void oracle_confocal_bins(const RecInfo* in, const vec3* laser, uint32_t l, int64_t* bins,
                          float* dist) {
    const uint32_t X = in->voxel_resolution.x, Y = in->voxel_resolution.y,
                   Z = in->voxel_resolution.z;
    for (int64_t v = 0; v < (int64_t)X * Y * Z; ++v) {
        uint32_t vx = (uint32_t)(v % X), vy = (uint32_t)((v / X) % Y), z = (uint32_t)(v / ((int64_t)X * Y));
        float d = confocal_dist(in, laser[l], voxel_pos(in, vx, vy, z));
        dist[v] = d;
        bins[v] = bin_of(d, in->time_step, in->num_time_bins);
    }
}

void oracle_exhaustive_bins(const RecInfo* in, const vec3* laser, const vec3* sensor, uint32_t l,
                            uint32_t s, int64_t* bins, float* dist) {
    const uint32_t X = in->voxel_resolution.x, Y = in->voxel_resolution.y,
                   Z = in->voxel_resolution.z;
    for (int64_t v = 0; v < (int64_t)X * Y * Z; ++v) {
        uint32_t vx = (uint32_t)(v % X), vy = (uint32_t)((v / X) % Y), z = (uint32_t)(v / ((int64_t)X * Y));
        float d = exhaustive_dist(in, laser[l], sensor[s], voxel_pos(in, vx, vy, z));
        dist[v] = d;
        bins[v] = bin_of(d, in->time_step, in->num_time_bins);
    }
}

/* ------------------------------------------------------------ depth planes
 * backprojectConfocal / backprojectExhaustive (FB/transient_reconstruction.cuh:41-91).
 * act: double[L][D]. */
static vec3 depth_point(const RecInfo* in, const vec3* laser, uint32_t c, double depth_d) {
    const vec3 nrm = in->relay_wall_normal;
    float depth = (float)depth_d;
    vec3 cp;
    cp.x = laser[c].x + nrm.x * depth;
    cp.y = laser[c].y + nrm.y * depth;
    cp.z = laser[c].z + nrm.z * depth;
    return cp;
}

/* distance(lPos, checkPos) + distance(checkPos, sPos) - t0 (+ offsets),
 * FB/transient_reconstruction.cuh:25-27, 82-84. */
static float depth_exhaustive_dist(const RecInfo* in, vec3 l, vec3 s, vec3 cp) {
    float a = glm_distance(l, cp);
    float b2 = glm_distance(cp, s);
    float t = a + b2;
    t = t - in->time_offset;
    if (in->discard_first_last_bounces == 0) {
        float o = glm_distance(l, in->laser_position) + glm_distance(s, in->sensor_position);
        t = t + o;
    }
    return t;
}

/* Bins (-1 outside [0, T)) of the depth-plane mode for one lit scan point l
 * (confocal) or one lit pair (l, s) (exhaustive): bins[c][d]. */
void oracle_depth_bins(const RecInfo* in, const vec3* laser, const vec3* sensor,
                       const double* depths, uint32_t D, int exhaustive, uint32_t l, uint32_t s,
                       int64_t* bins) {
    const uint32_t L = in->num_laser_targets;
    for (int64_t idx = 0; idx < (int64_t)L * D; ++idx) {
        uint32_t c = (uint32_t)(idx / D), d = (uint32_t)(idx % D);
        vec3 cp = depth_point(in, laser, c, depths[d]);
        float t = exhaustive ? depth_exhaustive_dist(in, laser[l], sensor[s], cp)
                             : confocal_dist(in, laser[l], cp);
        bins[idx] = bin_of(t, in->time_step, in->num_time_bins);
    }
}

void oracle_depth(const RecInfo* in, const float* I, const vec3* laser, const vec3* sensor,
                  const double* depths, uint32_t D, int exhaustive, double* act) {
    const uint32_t L = in->num_laser_targets, S = in->num_sensor_targets, T = in->num_time_bins;
#pragma omp parallel for schedule(dynamic, 4)
    for (int64_t idx = 0; idx < (int64_t)L * D; ++idx) {
        uint32_t c = (uint32_t)(idx / D), d = (uint32_t)(idx % D);
        vec3 cp = depth_point(in, laser, c, depths[d]);
        double acc = 0.0;
        if (!exhaustive) {
            for (uint32_t l = 0; l < L; ++l) {
                float t = confocal_dist(in, laser[l], cp);
                int64_t b = bin_of(t, in->time_step, T);
                if (b < 0) continue;
                float light = I[(size_t)l * T + (size_t)b];
                if (light > OR_EPS) acc += light;
            }
        } else {
            for (uint32_t s = 0; s < S; ++s)
                for (uint32_t l = 0; l < L; ++l) {
                    float t = depth_exhaustive_dist(in, laser[l], sensor[s], cp);
                    int64_t b = bin_of(t, in->time_step, T);
                    if (b < 0) continue;
                    float light = I[((size_t)l * S + s) * T + (size_t)b];
                    if (light > OR_EPS) acc += light;
                }
        }
        act[idx] = acc;
    }
}

/* -------------------------------------------------------------------- MIS
 * backprojectConfocalVoxelMIS (FB/transient_reconstruction.cuh:194-260) with
 * the samples of one voxel summed in order.  Tables (cdf, alias, prob) and
 * the noise buffer are inputs.  mode 0 = reference weights, 1 = unbiased
 * one-sample MIS estimator.  vol: double[Z][Y][X]. */
static float mis_uniform(const float* noise, uint32_t n, uint32_t* state) {
    *state = *state * 747796405u + 2891336453u;
    return noise[*state % n];
}

void oracle_mis(const RecInfo* in, const float* I, const vec3* laser, const float* cdf,
                const uint32_t* alias, const float* prob, const float* noise, uint32_t noise_size,
                uint32_t num_samples, int mode, int scatter, double* vol) {
    const uint32_t X = in->voxel_resolution.x, Y = in->voxel_resolution.y,
                   Z = in->voxel_resolution.z;
    const uint32_t L = in->num_laser_targets, T = in->num_time_bins;
    const uint32_t slice = X * Y;
    const float invL = 1.0f / (float)L;
    const float ext = voxel_extent(in);
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t vid = 0; vid < (int64_t)slice * Z; ++vid) {
        uint32_t v = (uint32_t)(vid % slice), z = (uint32_t)(vid / slice);
        uint32_t vx = v % X, vy = v / X;
        vec3 p = voxel_pos(in, vx, vy, z);
        int64_t* bins = scatter ? (int64_t*)malloc(sizeof(int64_t) * (T + 2)) : NULL;
        double acc = 0.0;
        for (uint32_t s = 0; s < num_samples; ++s) {
            uint32_t state = slice * z + v + s;
            uint32_t l;
            float pdf, gpdf, pa;
            if (mis_uniform(noise, noise_size, &state) < 0.5f) {
                float r1 = mis_uniform(noise, noise_size, &state);
                float r2 = mis_uniform(noise, noise_size, &state);
                uint32_t k = (uint32_t)(r1 * (float)L);
                if (k >= L) k = L - 1;
                l = (r2 < prob[k]) ? k : alias[k];
                pdf = l > 0 ? cdf[l] - cdf[l - 1] : cdf[l];
                gpdf = pdf + 1.0f * invL;
                pa = pdf;
            } else {
                l = (uint32_t)(mis_uniform(noise, noise_size, &state) * (float)L);
                if (l >= L) l = L - 1;
                pdf = 1.0f * invL;
                pa = l > 0 ? cdf[l] - cdf[l - 1] : cdf[l];
                gpdf = pdf + pa;
            }
            float d = confocal_dist(in, laser[l], p);
            double wd = safe_rcp64((double)(d * d));
            double sw;
            if (mode == 0)
                sw = (double)pdf * safe_rcp64((double)gpdf);
            else
                sw = 1.0 / (0.5 * ((double)pa + (double)invL) * (double)num_samples);
            if (scatter) {
                /* ACCUMULATE_VOXEL_SCATTERING branch (:235-250). */
                int nb = scatter_bins(d, ext, in->time_step, T, bins);
                for (int j = 0; j < nb; ++j) {
                    float light = I[(size_t)l * T + (size_t)bins[j]];
                    if (light > OR_EPS) acc += sw * (double)light * wd;
                }
                continue;
            }
            int64_t b = bin_of(d, in->time_step, T);
            if (b < 0) continue;
            float light = I[(size_t)l * T + (size_t)b];
            if (!(light > OR_EPS)) continue;
            acc += sw * (double)light * wd;
        }
        free(bins);
        vol[vid] = acc;
    }
}

/* Vose alias tables, Laser::buildAliasTables (FB/Laser.cu:253-315): float
 * arithmetic, back-of-stack pops.  Returns 0, or 6 on invalid input. */
int oracle_alias_tables(const float* cdf, uint32_t n, uint32_t* alias, float* prob) {
    float* p = (float*)malloc(sizeof(float) * n);
    uint32_t* sm = (uint32_t*)malloc(sizeof(uint32_t) * n);
    uint32_t* lg = (uint32_t*)malloc(sizeof(uint32_t) * n);
    uint32_t ns = 0, nl = 0;
    float sum = 0.0f;
    int rc = 0;
    p[0] = cdf[0];
    for (uint32_t i = 1; i < n; ++i) p[i] = cdf[i] - cdf[i - 1];
    for (uint32_t i = 0; i < n; ++i) {
        alias[i] = 0;
        if (p[i] < 0.0f) { rc = 6; goto done; }
        prob[i] = p[i] * (float)n;
        sum += p[i];
        if (prob[i] < 1.0f) sm[ns++] = i; else lg[nl++] = i;
    }
    if (fabsf(sum - 1.0f) > 1e-5f) { rc = 6; goto done; }
    while (ns > 0 && nl > 0) {
        uint32_t s = sm[--ns];
        uint32_t l = lg[--nl];
        alias[s] = l;
        prob[l] += prob[s] - 1.0f;
        if (prob[l] < 1.0f) sm[ns++] = l; else lg[nl++] = l;
    }
    while (ns > 0) prob[sm[--ns]] = 1.0f;
    while (nl > 0) prob[lg[--nl]] = 1.0f;
done:
    free(p); free(sm); free(lg);
    return rc;
}

/* ----------------------------------------------------------------- filters */

/* laplacianFilter (FB/transient_postprocessing.cuh:21-50), FP64. */
void oracle_laplacian(const float* in, double* out, uint32_t X, uint32_t Y, uint32_t Z, int k) {
    const int h = (int)floor((double)k / 2.0);
    const double k3 = (double)k * k * k;
#pragma omp parallel for schedule(static)
    for (int64_t v = 0; v < (int64_t)X * Y * Z; ++v) {
        int tx = (int)(v % X), ty = (int)((v / X) % Y), tz = (int)(v / ((int64_t)X * Y));
        double s = 0.0;
        for (int x = -h; x <= h; ++x)
            for (int y = -h; y <= h; ++y)
                for (int z = -h; z <= h; ++z) {
                    int nx = tx + x, ny = ty + y, nz = tz + z;
                    if (nx < 0 || nx >= (int)X || ny < 0 || ny >= (int)Y || nz < 0 || nz >= (int)Z)
                        continue;
                    s += in[((int64_t)nz * Y + ny) * X + nx];
                }
        out[v] = k3 * (double)in[v] - s;
    }
}

/* LoG weights from the closed form in FP64 (FB/PostprocessingFilters.cu:102-175):
 * G = exp(-r^2 / 2s^2), taps below FLT_EPSILON * max zeroed, normalised to
 * sum 1; w1 = G (r^2 - 3 s^2) / s^4; w = w1 - mean(w1).  w[kz][ky][kx]. */
void oracle_log_kernel(int k, double sigma, double* w) {
    int lim = (k - 1) / 2, n = k * k * k;
    double s2 = sigma * sigma, mx = 0.0, sum = 0.0, sum1 = 0.0;
    for (int kz = 0; kz < k; ++kz)
        for (int ky = 0; ky < k; ++ky)
            for (int kx = 0; kx < k; ++kx) {
                double x = kx - lim, y = ky - lim, z = kz - lim;
                double g = exp(-(x * x + y * y + z * z) / (2.0 * s2));
                w[(kz * k + ky) * k + kx] = g;
                if (g > mx) mx = g;
            }
    for (int i = 0; i < n; ++i) {
        if (w[i] < (double)FLT_EPSILON * mx) w[i] = 0.0;
        sum += w[i];
    }
    for (int i = 0; i < n; ++i) w[i] = sum != 0.0 ? w[i] / sum : w[i];
    for (int kz = 0; kz < k; ++kz)
        for (int ky = 0; ky < k; ++ky)
            for (int kx = 0; kx < k; ++kx) {
                double x = kx - lim, y = ky - lim, z = kz - lim;
                int i = (kz * k + ky) * k + kx;
                w[i] = w[i] * (x * x + y * y + z * z - 3.0 * s2) / (s2 * s2);
                sum1 += w[i];
            }
    for (int i = 0; i < n; ++i) w[i] -= sum1 / n;
}

/* LoGFilter (FB/transient_postprocessing.cuh:52-83): dense, clamp-to-edge, FP64. */
void oracle_log_filter(const float* in, double* out, uint32_t X, uint32_t Y, uint32_t Z, int k,
                       const double* w) {
    int h = (k - 1) / 2;
#pragma omp parallel for schedule(static)
    for (int64_t v = 0; v < (int64_t)X * Y * Z; ++v) {
        int x = (int)(v % X), y = (int)((v / X) % Y), z = (int)(v / ((int64_t)X * Y));
        double s = 0.0;
        for (int kz = -h; kz <= h; ++kz)
            for (int ky = -h; ky <= h; ++ky)
                for (int kx = -h; kx <= h; ++kx) {
                    int ix = x + kx, iy = y + ky, iz = z + kz;
                    ix = ix < 0 ? 0 : (ix > (int)X - 1 ? (int)X - 1 : ix);
                    iy = iy < 0 ? 0 : (iy > (int)Y - 1 ? (int)Y - 1 : iy);
                    iz = iz < 0 ? 0 : (iz > (int)Z - 1 ? (int)Z - 1 : iz);
                    s += (double)in[((int64_t)iz * Y + iy) * X + ix] *
                         w[((kz + h) * k + (ky + h)) * k + (kx + h)];
                }
        out[v] = s;
    }
}

/* normalizeReconstruction after min/max (FB/Laser.cu:317-339). */
void oracle_normalize(double* v, int64_t n) {
    double mn = v[0], mx = v[0];
    for (int64_t i = 1; i < n; ++i) {
        if (v[i] < mn) mn = v[i];
        if (v[i] > mx) mx = v[i];
    }
    double r = safe_rcp64(mx - mn);
    for (int64_t i = 0; i < n; ++i) v[i] = (v[i] - mn) * r;
}

/* ------------------------------------------------------------- filter_H
 * Laser::filter_H_cuda (FB/Laser.cu:128-240) in FP64: pad every time series
 * by t6sigma - 1 bins (zero or edge), convolve circularly with the ifftshifted
 * complex Gabor kernel K, keep the real part of the UNNORMALISED inverse
 * transform (= nt_pad * circular convolution), unpad.  Direct O(nt_pad^2). */
uint32_t oracle_filter_h_nt_pad(uint32_t nt, float dt, float wl_mean, float wl_sigma) {
    if (fabsf(wl_sigma) <= FLT_EPSILON) wl_sigma = wl_mean / sqrtf(2.0f);
    int t6 = (int)lround(6.0 * (double)wl_sigma / (double)dt);
    if (t6 % 2 == 1) t6 += 1;
    return nt + 2u * (uint32_t)(t6 - 1);
}

void oracle_filter_h(const float* data, uint64_t batch, uint32_t nt, float dt, float wl_mean,
                     float wl_sigma, int border, double* out) {
    if (fabsf(wl_sigma) <= FLT_EPSILON) wl_sigma = wl_mean / sqrtf(2.0f);
    const uint32_t N = oracle_filter_h_nt_pad(nt, dt, wl_mean, wl_sigma);
    const uint32_t pad = (N - nt) / 2;
    const double tmax = (double)dt * (double)(N - 1);
    double* kr = (double*)malloc(sizeof(double) * N);
    double* ki = (double*)malloc(sizeof(double) * N);
    double* er = (double*)malloc(sizeof(double) * N);
    double se = 0.0;
    for (uint32_t i = 0; i < N; ++i) {
        double t = (double)i * dt;
        double u = (t - tmax * 0.5) / (double)wl_sigma;
        er[i] = exp(-(u * u) * 0.5);
        se += er[i];
    }
    for (uint32_t i = 0; i < N; ++i) {
        /* ifftshift: K'[i] = K[(i + N/2) mod N] */
        uint32_t j = (i + N / 2) % N;
        double t = (double)j * dt;
        double ph = 2.0 * M_PI * t / (double)wl_mean;
        kr[i] = er[j] * cos(ph) / se;
        ki[i] = er[j] * sin(ph) / se;
    }
#pragma omp parallel
    {
        double* h = (double*)malloc(sizeof(double) * N);
#pragma omp for schedule(dynamic, 1)
        for (int64_t b = 0; b < (int64_t)batch; ++b) {
            const float* row = data + (size_t)b * nt;
            for (uint32_t t = 0; t < N; ++t) {
                if (t < pad) h[t] = border == 1 ? row[0] : 0.0;
                else if (t >= pad + nt) h[t] = border == 1 ? row[nt - 1] : 0.0;
                else h[t] = row[t - pad];
            }
            for (uint32_t t = 0; t < nt; ++t) {
                uint32_t n = t + pad;
                double acc = 0.0;
                for (uint32_t m = 0; m < N; ++m) {
                    uint32_t q = (n + N - m) % N;
                    acc += h[m] * kr[q];
                }
                out[(size_t)b * nt + t] = acc * (double)N;
            }
        }
        free(h);
    }
    free(kr); free(ki); free(er);
}

/* ---- Depth-plane slice images (TransientImage::save) ---------------------
 * FB/TransientImage.cpp:27-45 with Image<uchar> (FB/Image.h:45-67, 93-137,
 * 161-177), in the reference's order: build the W x H RGBA image of plane d
 * (gray = clamp(uint(a[p * D + d] * 255), 0, 255), alpha 255), flip it
 * vertically, then, when (OW, OH) differs, resize it bilinearly channel by
 * channel with the FP32 expression of Image::interpolate.  Negative and NaN
 * activations (undefined in the reference's cast) give 0.
 * out: uint8[D][OH][OW][4]. */
static unsigned char slice_gray(float v) {
    float s = v * 255.0f;
    if (!(s > 0.0f)) return 0;
    if (s >= 4294967040.0f) return 255;
    uint32_t u = (uint32_t)s;
    return (unsigned char)(u > 255u ? 255u : u);
}

void oracle_depth_slices_rgba(const float* act, uint32_t W, uint32_t H, uint32_t D, uint32_t OW,
                              uint32_t OH, unsigned char* out) {
    const size_t P = (size_t)W * H;
    unsigned char* img = (unsigned char*)malloc(P * 4);
    unsigned char* row = (unsigned char*)malloc((size_t)W * 4);
    for (uint32_t d = 0; d < D; ++d) {
        for (size_t p = 0; p < P; ++p) {
            unsigned char g = slice_gray(act[p * D + d]);
            img[p * 4 + 0] = img[p * 4 + 1] = img[p * 4 + 2] = g;
            img[p * 4 + 3] = 255;
        }
        /* flipImageVertically: swap row i with row H-1-i. */
        for (uint32_t i = 0; i < H / 2; ++i) {
            memcpy(row, img + (size_t)i * W * 4, (size_t)W * 4);
            memcpy(img + (size_t)i * W * 4, img + (size_t)(H - 1 - i) * W * 4, (size_t)W * 4);
            memcpy(img + (size_t)(H - 1 - i) * W * 4, row, (size_t)W * 4);
        }
        unsigned char* o = out + (size_t)d * OW * OH * 4;
        if (OW == W && OH == H) {
            memcpy(o, img, P * 4);
            continue;
        }
        const float xRatio = (float)(W - 1) / (float)(OW - 1);
        const float yRatio = (float)(H - 1) / (float)(OH - 1);
        for (uint32_t y = 0; y < OH; ++y)
            for (uint32_t x = 0; x < OW; ++x) {
                const float xs = (float)(int)x * xRatio, ys = (float)(int)y * yRatio;
                const int x0 = (int)xs, y0 = (int)ys;
                const int x1 = x0 + 1 < (int)W - 1 ? x0 + 1 : (int)W - 1;
                const int y1 = y0 + 1 < (int)H - 1 ? y0 + 1 : (int)H - 1;
                const float dx = xs - (float)x0, dy = ys - (float)y0;
                for (int c = 0; c < 4; ++c) {
                    const float p00 = img[((size_t)y0 * W + x0) * 4 + c];
                    const float p01 = img[((size_t)y0 * W + x1) * 4 + c];
                    const float p10 = img[((size_t)y1 * W + x0) * 4 + c];
                    const float p11 = img[((size_t)y1 * W + x1) * 4 + c];
                    const float v = (1.0f - dx) * (1.0f - dy) * p00 + dx * (1.0f - dy) * p01 +
                                    (1.0f - dx) * dy * p10 + dx * dy * p11;
                    o[((size_t)y * OW + x) * 4 + c] = (unsigned char)v;
                }
            }
    }
    free(row);
    free(img);
}

/* Threads of the OpenMP loops (the CPU baseline is timed at all cores and
 * at one). */
void oracle_set_threads(int n) { omp_set_num_threads(n > 0 ? n : 1); }
int oracle_max_threads(void) { return omp_get_max_threads(); }
