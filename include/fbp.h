/*
 * fbp.h - C ABI of the MI355X (gfx950) NLOS reconstruction core.
 *
 * Plain C: fixed-size integers, float pointers, opaque handles.  Every entry
 * point returns an fbp_status (0 = OK) and takes the HIP stream it runs on as
 * a void* (NULL = the null stream).  All data pointers are DEVICE pointers
 * unless a parameter says "host".  Nothing here allocates the caller's
 * buffers; a context owns scratch memory (partial sums, FFT work areas,
 * tables) and grows it on demand.  One context may be used from one host
 * thread at a time; different contexts are independent (no global state,
 * unlike the reference's __constant__ symbols, GpuStructs.cu:4-16).
 *
 * Each declaration names the reference interface it replaces (paths relative
 * to the reference repository; FB/ = FasterBackprojection/).
 */
#ifndef FBP_H
#define FBP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------ status */

typedef enum fbp_status {
    FBP_OK = 0,
    FBP_ERR_INVALID_ARGUMENT = 1, /* null pointer, zero size, bad enum      */
    FBP_ERR_DEVICE = 2,           /* a HIP runtime call failed              */
    FBP_ERR_OUT_OF_MEMORY = 3,    /* scratch allocation failed              */
    FBP_ERR_FFT = 4,              /* rocFFT plan/execute failed             */
    FBP_ERR_UNSUPPORTED = 5,      /* shape or mode outside what is built    */
    FBP_ERR_PROBABILITIES = 6,    /* alias table input invalid (Laser.cu:270-287) */
    FBP_ERR_IO = 7                /* file could not be read or written      */
} fbp_status;

/* Replaces CudaHelper::checkError's exception text (FB/CudaHelper.cpp:66-69). */
const char* fbp_status_string(int status);

/* ------------------------------------------------------------- data model */

typedef struct fbp_vec3 { float x, y, z; } fbp_vec3;       /* glm::vec3 layout  */
typedef struct fbp_uvec3 { uint32_t x, y, z; } fbp_uvec3;  /* glm::uvec3 layout */

enum { FBP_CONFOCAL = 0, FBP_EXHAUSTIVE = 1 };           /* FB/GpuStructs.cuh:3-6 */

/* Field-for-field mirror of ReconstructionInfo (FB/GpuStructs.cuh:8-38):
 * 148 bytes, no padding.  Volume axis convention of the reference kernels
 * (FB/transient_reconstruction.cuh:102-106): voxel (vx, vy, z) is stored at
 * z*X*Y + vy*X + vx and its centre is
 *   hidden_volume_min + (vx, z, vy) * voxel_size + 0.5 * voxel_size,
 * i.e. the slowest index walks world y (depth), the middle one world z. */
typedef struct fbp_reconstruction_info {
    fbp_vec3 sensor_position;
    uint32_t capture_system;            /* FBP_CONFOCAL / FBP_EXHAUSTIVE     */
    fbp_vec3 laser_position;
    float time_offset;                  /* t0, metres                        */
    uint32_t num_laser_targets;         /* L                                 */
    uint32_t num_sensor_targets;        /* S                                 */
    uint32_t num_time_bins;             /* T                                 */
    float time_step;                    /* bin width in metres (Delta)       */
    fbp_vec3 hidden_volume_min;
    uint32_t discard_first_last_bounces;
    fbp_vec3 hidden_volume_max;
    fbp_vec3 hidden_volume_size;
    fbp_vec3 hidden_volume_voxel_size;
    fbp_vec3 relay_wall_min_position;
    fbp_vec3 relay_wall_size;
    fbp_vec3 relay_wall_normal;
    fbp_uvec3 voxel_resolution;
} fbp_reconstruction_info;

/* Storage of a transient on the device. */
enum {
    FBP_STORAGE_F32 = 0,
    FBP_STORAGE_F16 = 1,  /* IEEE half, value = q * scale[row]   */
    FBP_STORAGE_BF16 = 2  /* bfloat16,  value = q * scale[row]   */
};
/* Element order of a confocal transient.  TIME_FASTEST = [l][t] is what the
 * loader, filter_H and MIS produce (FB/NLosData.cpp:264, FB/Laser.cu:80-126);
 * TIME_MAJOR = [t][l] is what getConfocalTransientIndex reads
 * (FB/transient_utils.cuh:6-9).  Exhaustive transients are always
 * [l][s][t] (FB/transient_utils.cuh:11-14). */
enum { FBP_LAYOUT_TIME_FASTEST = 0, FBP_LAYOUT_TIME_MAJOR = 1 };

typedef struct fbp_transient {
    const void* data;     /* device pointer, element type given by storage    */
    int32_t storage;      /* FBP_STORAGE_*                                    */
    int32_t layout;       /* FBP_LAYOUT_* (confocal only)                      */
    const float* scales;  /* one scale per time series (F16/BF16), else NULL  */
} fbp_transient;

/* --------------------------------------------------------------- context */

typedef struct fbp_context_s* fbp_context;

/* Replaces CudaContext::setDevice (FB/CudaContext.cpp:22-58).  device < 0
 * picks the device with the largest clock x multiprocessor count, as the
 * reference does. */
fbp_status fbp_context_create(int device, fbp_context* out);
fbp_status fbp_context_destroy(fbp_context ctx);
/* Detail text of the last failure on this context (empty if none). */
const char* fbp_context_last_error(fbp_context ctx);
int fbp_context_device(fbp_context ctx);
/* Bytes of scratch the context currently holds on its device. */
size_t fbp_context_scratch_bytes(fbp_context ctx);

/* Options of a context (no reference counterpart). */
typedef enum fbp_option {
    /* 1 (default): fbp_backproject_confocal_voxel may take the aligned-grid
     * path (same results; see that function); 0: always the general kernel. */
    FBP_OPTION_CONFOCAL_ALIGNED = 0,
    /* 0 (default): one bin per (voxel, scan point).  1: the reference's
     * ACCUMULATE_VOXEL_SCATTERING variant (FB/transient_reconstruction.cuh:
     * 112-126, 235-250): fbp_backproject_confocal_voxel and
     * fbp_backproject_confocal_mis add, for each pair, the samples of every
     * bin visited by d = max(0, dist - e), d += (Delta + 1e-5) while
     * d < dist + e (e = half the voxel diagonal), each weighted by
     * safeRCP(dist^2); the walk ends at the first bin >= T.  Runs a
     * dedicated kernel (not the aligned-grid path). */
    FBP_OPTION_VOXEL_SCATTERING = 1,
    /* 1 (default): fbp_backproject_exhaustive_voxel and
     * fbp_backproject_depth_exhaustive may run their laser-group kernels
     * (sensor distances shared by groups of eight lasers) when the capture
     * qualifies; 0: always the per-pair / staged-window kernels.  Same bins;
     * results agree to FP32 rounding. */
    FBP_OPTION_EXHAUSTIVE_LASER_GROUPS = 2
} fbp_option;
fbp_status fbp_context_set_option(fbp_context ctx, int32_t option, int32_t value);
fbp_status fbp_context_get_option(fbp_context ctx, int32_t option, int32_t* value);

/* Which path the last fbp_backproject_confocal_voxel or
 * fbp_backproject_depth_confocal call on this context took: *aligned = 1 for
 * the aligned-grid path, 0 for the general (gather) kernel.  Synchronises
 * the device. */
fbp_status fbp_context_last_confocal_path(fbp_context ctx, int32_t* aligned);

/* Which kernel the last fbp_backproject_exhaustive_voxel[_noalloc] or
 * fbp_backproject_depth_exhaustive call on this context launched:
 * *laser_groups = 1 for a laser-group kernel, 0 for the per-pair (voxel) or
 * staged-window (depth) kernel (host-side record, no synchronisation; no
 * reference counterpart). */
fbp_status fbp_context_last_exhaustive_kernel(fbp_context ctx, int32_t* laser_groups);

/* --------------------------------------------------------- backprojection */

/* Replaces backprojectConfocalVoxel + Laser::reconstructAABBConfocal
 * (FB/transient_reconstruction.cuh:93-136, FB/Laser.cu:482-498).
 * volume: float[Z][Y][X] (X,Y,Z = voxel_resolution), OVERWRITTEN with
 *   sum_l [I > 1e-5] * I(l, bin) * safeRCP(dist^2).
 * laser_targets: fbp_vec3[L].  Deterministic (no atomics): per-chunk partial
 * sums are reduced in a fixed order in FP64.
 * Two implementations, chosen on the device per call: the aligned-grid path
 * (a ring convolution over transient planes) runs when t0 = 0 with bounces
 * discarded, the scan points form an n x n grid on a plane of constant y,
 * X = 64 or 128 voxels along x with n = X or 2X, Y a multiple of 16 (at most
 * X), storage F32 (either layout) or F16/BF16 time-fastest, and a device
 * check proves in FP32 that the per-axis squared distances depend only on
 * the voxel-to-scan-point offset (mirrored for odd scan columns when
 * n = 2X); otherwise the general kernel runs.  Both give the bins of the
 * IEEE chain; results agree within the FP32 tolerance. */
fbp_status fbp_backproject_confocal_voxel(fbp_context ctx, const fbp_reconstruction_info* info,
                                          const fbp_transient* transient,
                                          const fbp_vec3* laser_targets, float* volume,
                                          void* stream);

/* Scatter form of the confocal backprojection; stands in for
 * DrawEllipsoids::solveBackprojection (FB/DrawEllipsoids.cpp:332-440), which
 * rasterises one half-ellipsoid instance per lit sample (l, t) (I > epsilon,
 * FB/DrawEllipsoids.cpp:29-56) and adds its intensity into the voxel under
 * each fragment (FB/assets/shaders/shading/backprojection-fs.glsl:24).  Here
 * each lit sample (I > 1e-5) adds I * safeRCP(dist^2) into exactly the voxels
 * whose bin for scan point l is t under backprojectConfocalVoxel's arithmetic,
 * so the volume equals fbp_backproject_confocal_voxel's within FP32 summation
 * order.  Work scales with the voxels that lie in lit shells, not with the
 * full (voxel, scan point) product.  Deterministic (FP64 chunk reduction).
 * Storage F32 (either layout); FBP_ERR_UNSUPPORTED for F16/BF16, with the
 * voxel-scattering option set, when a 16 x 16 x 32 voxel tile spans more than
 * 4095 bins, or for a negative depth voxel size. */
fbp_status fbp_backproject_confocal_splat(fbp_context ctx, const fbp_reconstruction_info* info,
                                          const fbp_transient* transient,
                                          const fbp_vec3* laser_targets, float* volume,
                                          void* stream);

/* Replaces backprojectExhaustiveVoxel + Laser::reconstructAABBExhaustive
 * (FB/transient_reconstruction.cuh:138-166, FB/Laser.cu:500-516).  Visits every
 * (laser, sensor) pair with l = ls / S, s = ls % S (the reference splits by L
 * and swaps grid axes; identical when L == S and every pair is visited). */
fbp_status fbp_backproject_exhaustive_voxel(fbp_context ctx, const fbp_reconstruction_info* info,
                                            const fbp_transient* transient,
                                            const fbp_vec3* laser_targets,
                                            const fbp_vec3* sensor_targets, float* volume,
                                            void* stream);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): fbp_backproject_exhaustive_voxel with the same
 * arguments and results, whose per-chunk partial sums live in a fixed
 * module-scope device array of FBP_NOALLOC_SCRATCH_BYTES per GPU (held from
 * fbp_context_create until the process ends) instead of the context's
 * growing arena.  Calls no hipMalloc/hipFree or allocating library routine.
 * The pair chunks are capped so their partial sums fit: volumes whose
 * 16 x 16 x 32-tile-padded voxel count times 4 bytes exceeds the array
 * (more than 8 Mi padded voxels, e.g. above 128 x 128 x 512) are refused
 * with FBP_ERR_UNSUPPORTED.  Every call on one GPU shares that GPU's array,
 * whatever the context: calls that overlap on one GPU (other streams,
 * other threads) race and must be serialised by the caller. */
#define FBP_NOALLOC_SCRATCH_BYTES (32u << 20)
fbp_status fbp_backproject_exhaustive_voxel_noalloc(fbp_context ctx,
                                                    const fbp_reconstruction_info* info,
                                                    const fbp_transient* transient,
                                                    const fbp_vec3* laser_targets,
                                                    const fbp_vec3* sensor_targets, float* volume,
                                                    void* stream);

/* Depth-plane mode: replaces backprojectConfocal / backprojectExhaustive and
 * Laser::reconstructDepthConfocal / Exhaustive
 * (FB/transient_reconstruction.cuh:41-91, FB/Laser.cu:379-480).
 * depths: double[num_depths] (device).  activation: float[L][num_depths],
 * OVERWRITTEN with the unweighted sum of I > 1e-5 samples.  Confocal calls on
 * a 64^2 or 128^2 scan grid with the wall normal along y, t0 = 0 and the
 * first/last bounces discarded run the ring-convolution path when the
 * device-side grid check passes (same sums, FP32 partials reduced in FP64). */
fbp_status fbp_backproject_depth_confocal(fbp_context ctx, const fbp_reconstruction_info* info,
                                          const fbp_transient* transient,
                                          const fbp_vec3* laser_targets, const double* depths,
                                          uint32_t num_depths, float* activation, void* stream);
fbp_status fbp_backproject_depth_exhaustive(fbp_context ctx, const fbp_reconstruction_info* info,
                                            const fbp_transient* transient,
                                            const fbp_vec3* laser_targets,
                                            const fbp_vec3* sensor_targets, const double* depths,
                                            uint32_t num_depths, float* activation, void* stream);

/* Depth-plane slice images: replaces the per-depth image output of
 * Laser::reconstructDepthConfocal / Exhaustive (FB/Laser.cu:410-424, 463-476)
 * made by TransientImage::save (FB/TransientImage.cpp:27-45) with
 * Image<uchar>::flipImageVertically / resize / interpolate
 * (FB/Image.h:45-67, 93-137, 161-177).
 * activation: float[width * height][num_depths] (device, depth fastest, as
 * written by fbp_backproject_depth_*; normally min-max normalised).
 * rgba: uint8[num_depths][out_height][out_width][4] (device, 4-byte aligned):
 * gray = clamp(uint(a * 255), 0, 255) (negative and NaN -> 0), alpha 255,
 * rows flipped, then, when the output size differs, the reference's FP32
 * bilinear resampling of all four channels with ratio (W-1)/(W'-1).  The
 * reference's sizes: confocal W' = W * (W < 256 ? 256 / W : 1), exhaustive
 * W' = 4 W.  Edges at most 65535 (16-bit in the reference). */
fbp_status fbp_depth_slices_rgba(fbp_context ctx, const float* activation, uint32_t width,
                                 uint32_t height, uint32_t num_depths, uint32_t out_width,
                                 uint32_t out_height, uint8_t* rgba, void* stream);

/* PNG file (8-bit RGBA, zlib-compressed) of one host image, rows top to
 * bottom: the SOIL_save_image(..., SOIL_SAVE_TYPE_PNG, w, h, 4, ...) call of
 * TransientImage::save (FB/TransientImage.cpp:44).  The encoded bytes differ
 * from SOIL's encoder; the decoded pixels are the input bytes. */
fbp_status fbp_write_png_rgba(const char* path, const uint8_t* rgba, uint32_t width,
                              uint32_t height);

/* MIS-sampled confocal backprojection: replaces Laser::reconstructAABBConfocalMIS,
 * buildAliasTables, spatioTemporalPrefixSum and backprojectConfocalVoxelMIS
 * (FB/Laser.cu:253-315, 526-611; FB/transient_reconstruction.cuh:170-260;
 * FB/transient_utils.cuh:18-53). */
enum {
    FBP_MIS_REFERENCE = 0, /* reference weights: w = pdf/(pdf + pdf_other), no 1/pdf */
    FBP_MIS_UNBIASED = 1,  /* one-sample MIS estimator f/(N_s (p_a + p_u)/2)         */
    /* Tile-shared samples (no reference counterpart): the voxels of each
     * 16x16x16 tile (tiles in the order of the voxel kernels) share
     * num_samples scan points, drawn with probability 1/2 from
     * p_a(l) = R(l) / sum R (R = FP64 sum of the samples of row l above 1e-5)
     * and 1/2 uniformly, each weighted 1 / (num_samples * (p_a(l) + 1/L) / 2):
     * unbiased.  Sample i of tile t: LCG state = t * num_samples + i, two
     * draws (branch, position) from noise; the importance position is the
     * first l whose inclusive prefix of R exceeds u * sum R.  The tables stay
     * on the device (no host round trip); contributions use the exact bins
     * and weights of fbp_backproject_confocal_voxel. */
    FBP_MIS_TILED = 2
};
typedef struct fbp_mis_params {
    uint32_t num_samples;   /* samples per voxel (reference: L / 4)            */
    uint32_t mode;          /* FBP_MIS_*                                       */
    const float* noise;     /* device float[noise_size] in [0,1)               */
    uint32_t noise_size;    /* reference: number of voxels                     */
    /* FBP_MIS_TILED with noise == NULL: entry p of the noise array is
     * (mix(p ^ key(seed)) >> 8) / 2^24 (no reference counterpart; mix is the
     * "lowbias32" integer hash x ^= x >> 16, x *= 0x7feb352d, x ^= x >> 15,
     * x *= 0x846ca68b, x ^= x >> 16, and key(seed) = mix(lo ^ mix(hi +
     * 0x9e3779b9)) of the seed's 32-bit halves): the draw needs no noise
     * buffer.                                                                 */
    uint64_t seed;
} fbp_mis_params;
fbp_status fbp_backproject_confocal_mis(fbp_context ctx, const fbp_reconstruction_info* info,
                                        const fbp_transient* transient,
                                        const fbp_vec3* laser_targets,
                                        const fbp_mis_params* params, float* volume,
                                        void* stream);
/* Vose alias tables from a CDF (host arrays): Laser::buildAliasTables
 * (FB/Laser.cu:253-315), same pop order and error checks. */
fbp_status fbp_build_alias_tables_host(const float* cdf, uint32_t n, uint32_t* alias_table,
                                       float* prob_table);

/* Precision form: quantise a float32 transient to F16/BF16 with one scale
 * per time series (scale = max_t I, values <= 1e-5 zeroed first).  No
 * reference counterpart.  rows = number of time series, cols = T. */
fbp_status fbp_quantize_transient(fbp_context ctx, const float* src, uint64_t rows, uint32_t cols,
                                  int32_t storage, void* dst, float* scales, void* stream);

/* ---------------------------------------------------------------- filters */

/* Laplacian::compute + laplacianFilter (FB/PostprocessingFilters.cu:81-98,
 * FB/transient_postprocessing.cuh:21-50): out = k^3 v - sum of in-bounds box. */
fbp_status fbp_filter_laplacian(fbp_context ctx, const float* in, float* out, fbp_uvec3 size,
                                int32_t kernel_size, void* stream);
/* LoG::compute + calculateLaplacianKernel + LoGFilter
 * (FB/PostprocessingFilters.cu:12-32, 102-175; FB/transient_postprocessing.cuh:52-83),
 * clamp-to-edge.  Evaluated in separable form (1D passes plus a box term,
 * fused into one pass over the volume) when the kernel is exactly separable
 * (no Gaussian tap thresholded away, odd size), else as the dense k^3 sum. */
fbp_status fbp_filter_log(fbp_context ctx, const float* in, float* out, fbp_uvec3 size,
                          int32_t kernel_size, float sigma, void* stream);
/* The k^3 LoG weights exactly as the reference builds them (host array). */
fbp_status fbp_log_kernel_host(int32_t kernel_size, float sigma, float* weights);
/* LoGFFT::compute (FB/PostprocessingFilters.cu:36-77): in-place, unnormalised
 * (result scaled by X*Y*Z like the reference's C2R). */
fbp_status fbp_filter_log_fft(fbp_context ctx, float* volume, fbp_uvec3 size, float sigma,
                              void* stream);
/* Laser::normalizeMatrix + normalizeReconstruction (FB/Laser.cu:317-339,
 * FB/transient_postprocessing.cuh:10-17): v = (v - min) * safeRCP(max - min). */
fbp_status fbp_normalize_minmax(fbp_context ctx, float* v, uint64_t n, void* stream);
/* fbp_filter_log followed by fbp_normalize_minmax on out, with the same
 * results, in two launches: the LoG kernel reduces the min and max of its
 * output itself (the reference runs LoGFilter, then normalizeMatrix's
 * reductions and normalizeReconstruction: FB/Laser.cu:317-339). */
fbp_status fbp_filter_log_normalize(fbp_context ctx, const float* in, float* out,
                                    fbp_uvec3 size, int32_t kernel_size, float sigma,
                                    void* stream);
/* min and max of v written to host pointers (synchronises the stream). */
fbp_status fbp_minmax(fbp_context ctx, const float* v, uint64_t n, float* host_min,
                      float* host_max, void* stream);

/* Laser::filter_H_cuda (FB/Laser.cu:128-240): temporal band-pass of every
 * time series of data[batch][nt], in place on the device.  wl_sigma <= 0
 * means wl_mean / sqrt(2).  border: 0 = zero, 1 = edge.  Like the reference
 * the inverse transform is unnormalised (output scaled by nt_pad). */
fbp_status fbp_filter_h(fbp_context ctx, float* data, uint64_t batch, uint32_t nt, float delta_t,
                        float wl_mean, float wl_sigma, int32_t border, void* stream);
/* nt_pad the reference uses for these parameters (t6 = round(6 sigma /
 * delta_t) made even, nt_pad = nt + 2 (t6 - 1)); 0 when the rule gives no
 * padding (t6 = 0, i.e. six sigmas below half a bin: fbp_filter_h rejects
 * it) or more than 2^24 samples. */
uint32_t fbp_filter_h_padded_length(uint32_t nt, float delta_t, float wl_mean, float wl_sigma);

/* ------------------------------------------------------------------- LCT */

/* Confocal LCT after python/lct.py:cnlos_reconstruction (:95-211), the
 * reconstruction the reference's LCT class (FB/LCT.h:6-12) leaves empty.
 * data: float[N][N][m] (rect_data order: x, y, t), m and N powers of two.
 * Downsampling (:98-105): K = downsample rounds, each summing adjacent time
 * bins pairwise (rect[:, :, 0::2] + rect[:, :, 1::2], so the 2^K bins of one
 * output bin are added as a balanced tree in that order) and halving z_trim
 * with Python's round (half to even); M = m >> K.  Then the first z_trim
 * bins are zeroed (:109) and
 * out:  float[M][N][N] = max(Re(mtxi @ crop(ifftn(fftn(pad(mtx @ (data'*z^p))) * invpsf))), 0)
 * with p = 4 if is_diffuse else 2.  width / range_max is the psf slope, as
 * in define_psf(N, M, width / range_max); range_max is the pre-downsampling
 * m * c * bin_resolution, as lct.py computes it before the loop (:95). */
typedef struct fbp_lct_params {
    uint32_t n;          /* spatial resolution N                         */
    uint32_t m;          /* time bins of data (before downsampling)      */
    float width;         /* half-width of the scanned wall, metres       */
    float range_max;     /* m * c * bin_resolution, metres               */
    float snr;           /* Wiener SNR (invpsf = conj(F)/(|F|^2 + 1/snr)) */
    int32_t is_diffuse;  /* z^4 instead of z^2 radiometric term          */
    int32_t is_backprop; /* invpsf = conj(F) only                        */
    uint32_t z_trim;     /* bins zeroed, counted before downsampling     */
    uint32_t downsample; /* K (lct.py:31), 0..16; m >> K >= 2            */
} fbp_lct_params;
fbp_status fbp_lct(fbp_context ctx, const float* data, const fbp_lct_params* params, float* out,
                   void* stream);

/* Confocal voxel backprojection evaluated only at the listed voxels (linear
 * indices z*X*Y + vy*X + vx): values[i] = sum_l [I > 1e-5] I(l, bin) *
 * safeRCP(dist^2) for voxel voxels[i], with the reference's IEEE chain per
 * pair (FB/transient_reconstruction.cuh:93-136) and an FP64 sum.  F32
 * time-fastest transients.  values: double[count] (device). */
fbp_status fbp_backproject_confocal_points(fbp_context ctx, const fbp_reconstruction_info* info,
                                           const fbp_transient* transient,
                                           const fbp_vec3* laser_targets, const uint32_t* voxels,
                                           uint32_t count, double* values, void* stream);

/* Learned form (no reference counterpart): maps a refined, min-max
 * normalised LoG volume (volume, float[Z][Y][X], IN/OUT) onto the exactly
 * normalised LoG of the plain backprojection.  reference (float[Z][Y][X]):
 * an unbiased estimate of the LoG volume in physical units whose noise is
 * independent of volume's (e.g. the LoG of a second MIS draw); its
 * least-squares fit on volume over all voxels gives the LoG as a v + b.
 * The true range is the exact LoG (kernel_size 5, sigma) at the `extremes`
 * (1..32) lowest and highest voxels of volume, from exact backprojections of
 * their 5^3 clamp-to-edge neighbourhoods.  Output (a v + b - min) *
 * safeRCP(max - min).  F32 time-fastest transients. */
fbp_status fbp_learned_calibrate(fbp_context ctx, const fbp_reconstruction_info* info,
                                 const fbp_transient* transient, const fbp_vec3* laser_targets,
                                 float* volume, const float* reference, int32_t kernel_size,
                                 float sigma, uint32_t extremes, void* stream);

/* ------------------------------------------------------ scene forward model */

/* Confocal transient of a set of Lambertian points: for every scan point l and
 * point p, I[l][round(2 r / delta)] += albedo / r^4 (r = |l - p|).  Used to
 * render training scenes and for the refiner's residual check.  No reference
 * counterpart.  transient: float[L][T] time-fastest, OVERWRITTEN. */
fbp_status fbp_render_confocal(fbp_context ctx, const fbp_vec3* scan_points, uint32_t num_scan,
                               const fbp_vec3* points, const float* albedo, uint32_t num_points,
                               uint32_t num_time_bins, float delta, float* transient,
                               void* stream);

/* Exhaustive transient of Lambertian points: I[l][s][round((|l-p| + |p-s|) /
 * delta)] += albedo / (|l-p|^2 |p-s|^2).  transient: float[L][S][T],
 * OVERWRITTEN.  No reference counterpart. */
fbp_status fbp_render_exhaustive(fbp_context ctx, const fbp_vec3* laser_points, uint32_t num_lasers,
                                 const fbp_vec3* sensor_points, uint32_t num_sensors,
                                 const fbp_vec3* points, const float* albedo, uint32_t num_points,
                                 uint32_t num_time_bins, float delta, float* transient,
                                 void* stream);

/* Forward-model residual of a reconstructed volume (learned form's failure
 * check, no reference counterpart): with polarity +1 the voxels above
 * thr = min + threshold_fraction * (max - min) are rendered as points
 * (albedo = excess over thr), with polarity -1 the voxels below it (albedo =
 * thr - v: surfaces are minima of a LoG-filtered volume); 1/r^4, bin
 * round(2r/Delta), at the scan points
 * rows[0..num_rows); both that and the measured rows are box-blurred over
 * +-blur bins; host_residual = 1 - Pearson correlation of the rendered and
 * measured rows (each row centred on its own mean, second moments pooled
 * over the rows).  1 when either side is constant.  Synchronises. */
fbp_status fbp_volume_residual(fbp_context ctx, const fbp_reconstruction_info* info,
                               const float* volume, const float* transient,
                               const fbp_vec3* laser_targets, const uint32_t* rows,
                               uint32_t num_rows, float threshold_fraction, int32_t polarity,
                               int32_t blur, double* host_residual, void* stream);

/* ---------------------------------------------------------- learned form */

/* Learned refiner (no reference counterpart): a 3-level 3D U-Net
 * (16/32/64 channels, 3x3x3 kernels) mapping a max-normalised MIS-sampled
 * backprojection to the backprojection + LoG volume.  Forward convolutions
 * with 16+ channels are BF16 MFMA implicit GEMMs with FP32 accumulation;
 * training (L1 + SSIM loss, Adam) keeps FP32 master weights.  Volumes are
 * float[S][S][S] on the device, S a multiple of 16. */
typedef struct fbp_refiner_s* fbp_refiner;
size_t fbp_refiner_param_count(void);
fbp_status fbp_refiner_create(fbp_context ctx, uint32_t size, uint64_t seed, fbp_refiner* out);
fbp_status fbp_refiner_destroy(fbp_refiner r);
/* Parameters as one FP32 host array (fbp_refiner_param_count() values). */
fbp_status fbp_refiner_load(fbp_refiner r, const float* host_params, uint64_t count);
fbp_status fbp_refiner_save(fbp_refiner r, float* host_params, uint64_t count);
fbp_status fbp_refiner_forward(fbp_refiner r, const float* input, float* output, void* stream);
/* Learned form in one call (no reference counterpart): the tile-shared MIS
 * backprojection (FBP_MIS_TILED, num_samples scan points per tile, uniforms
 * from `seed` over a noise array of S^3 entries), its LoG (kernel_size 5,
 * sigma), the refiner on the LoG's min-max normalisation (the min and max
 * come from the LoG kernel itself), then fbp_learned_calibrate of the
 * refined volume against that LoG with `extremes` voxels per side.  The
 * voxel grid must be the refiner's S^3; F32 time-fastest transients.
 * volume: the calibrated [Z][Y][X] result. */
fbp_status fbp_learned_reconstruct(fbp_refiner r, const fbp_reconstruction_info* info,
                                   const fbp_transient* transient,
                                   const fbp_vec3* laser_targets, uint32_t num_samples,
                                   uint64_t seed, int32_t kernel_size, float sigma,
                                   uint32_t extremes, float* volume, void* stream);
/* One Adam step over `batch` (input, target) volume pairs laid out back to
 * back; loss = mean|y - t| + ssim_weight * (1 - mean SSIM(y, t)).
 * host_loss (may be NULL, else synchronises) receives the batch mean. */
fbp_status fbp_refiner_train_step(fbp_refiner r, const float* inputs, const float* targets,
                                  uint32_t batch, float lr, float ssim_weight, float* host_loss,
                                  void* stream);
/* Gradient of the loss w.r.t. all parameters for one pair (host output). */
fbp_status fbp_refiner_gradient(fbp_refiner r, const float* input, const float* target,
                                float ssim_weight, float* host_grad, float* host_loss);

#ifdef __cplusplus
}
#endif

#endif /* FBP_H */
