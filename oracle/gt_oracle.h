/*
 * gt_oracle.h -- CPU restatement of gpu_tracking's per-frame detection
 * pipeline.  TEST INFRASTRUCTURE ONLY: the product (libgtrack.so and the
 * gpu_tracking_amd package) never links, loads or calls this code.  Only
 * tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg use it, as
 * the result the HIP path is checked against.
 *
 * Every function names the reference file:line it restates.  The reference's
 * arithmetic is WGSL compiled by naga/wgpu; the few places where WGSL leaves
 * the result to the driver are fixed here as follows (and the HIP kernels
 * follow the same choices, so the two agree bit for bit):
 *   - per-tap multiply-accumulates `acc += w * x` are contracted (fmaf);
 *     SPIR-V without NoContraction lets a driver do this;
 *   - the Gaussian tap weights are one f32 table computed with expf;
 *   - sin(2*theta), cos(2*theta) of atan2(j, i) use the exact rational forms
 *     2ij/(i^2+j^2) and (i^2-j^2)/(i^2+j^2);
 *   - NaN converts to integer 0 (what AMD v_cvt_i32_f32 does).
 */
#ifndef GT_ORACLE_H
#define GT_ORACLE_H
#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Derived per-frame parameters: the 16-word WGSL `Params` uniform
 * (src/shaders/params.wgsl:1-20) plus the frame std and the truncation flag. */
typedef struct {
    int32_t nrows, ncols;
    int32_t smoothing;        /* preprocess kernel side s (composite_dims) */
    float sigma2;             /* noise_size, used as a variance */
    int32_t circle;           /* diameter (circle_dims) */
    int32_t dilation;         /* dilation window side k */
    uint32_t max_iterations;
    float threshold;          /* shift threshold */
    float minmass;
    int32_t margin;
    float snr;
    float minmass_snr;
    float rough_snr_factor;
    float image_std;          /* 0 unless an snr criterion is active */
    int32_t truncate;         /* truncate_preprocessed */
    /* LoG only, no reference counterpart (0 = off): a scale-space maximum is
     * kept when its LoG value exceeds threshold_rel times the largest LoG
     * value of the frame over every scale (skimage blob_log's rule). */
    float threshold_rel;
} or_params;

/* ResultRow (src/execute_gpu.rs:48-61) plus the raster index of the
 * candidate pixel the walk started from (the deterministic order key). */
typedef struct {
    float x, y, mass, r, max_intensity, Rg, raw_mass, signal, ecc, count;
    uint32_t seed;
    uint32_t pad;
} or_row;

/* gpu_setup.rs:528-580 (Trackpy branch). smoothing_size <= 0 means "diameter". */
void or_derive_trackpy(uint32_t diameter, uint32_t separation, int32_t smoothing_size,
                       float noise_size, float minmass, uint32_t max_iterations,
                       float shift_threshold, float snr, float minmass_snr,
                       int adaptive_background, int truncate, int32_t nrows,
                       int32_t ncols, or_params *out);

/* Tap weights of the preprocess Gaussian, i = -s/2..s/2 (preprocess_rows.wgsl:30-47).
 * w must hold 2*(s/2)+1 floats; *norm receives their sequential f32 sum. */
void or_gauss_weights(int32_t s, float sigma2, float *w, float *norm);

/* preprocess_rows.wgsl:14-55 then preprocess_cols.wgsl:11-60.
 * frame, processed: nrows*ncols f32.  scratch: 2*nrows*ncols f32. */
void or_bandpass(const float *frame, const or_params *p, float *processed, float *scratch);

/* max_rows.wgsl:10-32 + extract_max.wgsl:35-78.  Writes candidate raster
 * indices (u*ncols+v) in raster order; returns the count. */
size_t or_local_max(const float *processed, const or_params *p, uint32_t *cand,
                    size_t cap, float *scratch_rows);

/* walk.wgsl:67-229 for every candidate (r = -1, Trackpy).  Returns the
 * number of rows kept (mass >= minmass) in candidate order. */
size_t or_walk(const float *processed, const or_params *p, const uint32_t *cand,
               size_t ncand, or_row *rows);

/* characterize.wgsl:30-155 on rows[0..n).  use_f64 = 1 accumulates the
 * Rg and eccentricity moment sums in f64; 0 keeps the reference's f32 sums.
 * recompute_mass = 1 is the characterize_points feature (characterize.wgsl:75). */
void or_characterize(const float *processed, const float *raw, const or_params *p,
                     or_row *rows, size_t n, int use_f64, int recompute_mass);

/* filter_close_trackpy (execute_gpu.rs:556-590) with within_radius_rd2
 * (linking.rs:208-247).  keep[i] = 1 if row i survives. O(n^2). */
void or_filter_close(const or_row *rows, size_t n, float separation, uint8_t *keep);

/* Doughnut photometry (execute_gpu.rs:717-790; kernels.rs:138-171).
 * out3[3*i..] = raw_mass, raw_bg_median, raw_mass_corrected. */
void or_doughnut(const float *raw, int32_t nrows, int32_t ncols, const or_row *rows,
                 size_t n, float radius, float bg_radius, float gap_radius, float *out3);

/* Sort rows by seed (the deterministic order used by both implementations). */
void or_sort_rows(or_row *rows, size_t n);

/* Whole frame: bandpass -> local max -> walk -> (characterize, f64 moments).
 * Returns the number of rows (sorted by seed), or (size_t)-1 on overflow. */
size_t or_detect_frame(const float *frame, const or_params *p, int characterize,
                       or_row *rows, size_t cap, size_t *ncand_out);

/* Frame-parallel driver for the CPU baseline: nthreads POSIX threads, frame
 * f handled by thread f % nthreads.  dtype 0 = uint16, 1 = float32.
 * rows: nframes*cap_per_frame; counts: nframes.  Returns 0 or -1. */
int or_detect_batch(const void *frames, int dtype, size_t nframes, const or_params *p,
                    int characterize, int nthreads, or_row *rows, size_t cap_per_frame,
                    uint64_t *counts, uint64_t *ncands);

/* walk.wgsl:104-229 with a radius per candidate (cand_r[c] > 0: LoG,
 * rint = ceil(r), r2 = r*r; cand_r == NULL: Trackpy, r = -1). */
size_t or_walk_r(const float *processed, const or_params *p, const uint32_t *cand,
                 const float *cand_r, size_t ncand, or_row *rows);

/* ---- LoG style ---- */
/* The radius scan (gpu_setup.rs:958-984). */
void or_log_radii(float min_r, float max_r, uint32_t n, int log_spacing, float *out);
/* Tap weights of one scale (separable_log.wgsl:22-89): returns R; g, d hold
 * 2R+1 floats (Gaussian and second-derivative taps) when not NULL. */
int32_t or_log_taps(float sigma, float *g, float *d, float *norm1, float *norm3);
/* Scale-normalised LoG plane L = -(sigma^2)(Gxx + Gyy) * raw, by the four 1-D
 * passes of Laplace<2> (utils.rs:114-296).  tmp, tmp2: H*W floats each. */
void or_log_plane(const float *raw, int32_t H, int32_t W, float sigma, float *L, float *tmp,
                  float *tmp2);
/* Scale-space maxima of `middle` (logspace_max.wgsl:62-154); bottom / top
 * NULL at the ends of the scan.  Raster order; returns the count (may exceed cap). */
size_t or_log_max(const float *bottom, const float *middle, const float *top,
                  const float *processed, int32_t H, int32_t W, float snr_threshold,
                  uint32_t *cand, size_t cap);
/* blob_overlap (execute_gpu.rs:597-630). */
float or_blob_overlap(float r1, float r2, float distance);
/* Sort (r desc, mass desc, stable) + prune_blobs_log (execute_gpu.rs:592-712);
 * compacts rows in place and returns the survivor count. */
size_t or_prune_blobs(or_row *rows, size_t n, float overlap_threshold);
/* Whole LoG frame (execute_gpu.rs:237-286 + the LoG branch of post_process);
 * rows come out in the pruned, sorted order; returns their count or (size_t)-1. */
size_t or_log_detect_frame(const float *frame, const or_params *p, const float *radii,
                           uint32_t n_radii, int characterize, float overlap_threshold,
                           or_row *rows, size_t cap, size_t *ncand_out);

/* FP64 twin for error reporting: f64 bandpass (exp in f64, exact divisions). */
/* Noise estimate of a raw frame for the snr / minmass_snr rules
 * (src/utils.rs:298-670 StdArray: MeanArrayOutplace, std_image, MeanArray;
 * shaders sum_image_outplace.wgsl:14-67, sum_image_inplace.wgsl:10-42,
 * std_image.wgsl:16-30, last_reduction.wgsl:18-34).  Eight-way tree: a level
 * of length len has stride s = ceil(len / 8) and out[i] = sum_{j<8} in[i +
 * s j] in j order from 0.0; levels repeat while len > 24; the last values are
 * summed in order.  The mean's first level skips NaN-marker pixels (bits
 * 0xFFFFFFFF), the squared deviations are 0 for them; the divisors are f32(N)
 * and f32(N - 1), the pixel count the runtime writes into the divisor buffer
 * each frame (src/execute_gpu.rs:113-135).  Indices at or past a level's
 * length are skipped (the reference's guard is vacuous there and reads
 * stale or out-of-range values; identical whenever every level length is a
 * multiple of 8).  N must exceed 24. */
void or_frame_std(const float *frame, size_t n, float *mean, float *std);
/* The same with explicit divisors (mean: div_mean, variance: div_var); the
 * adaptive filter's re-estimates divide by the pixel count less the
 * accepted particles' disk counts (adaptive_filter.wgsl:116). */
void or_frame_std_div(const float *frame, size_t n, uint32_t div_mean, uint32_t div_var,
                      float *mean, float *std);
/* Trackpy detection with adaptive_background = n_iter (src/execute_gpu.rs:
 * 290-330, adaptive_filter.wgsl:89-125): the noise estimate of the raw frame;
 * local max and walk with rough_snr_factor 0.5 (p from or_derive_trackpy
 * with adaptive_background); then n_iter passes over the rows not yet
 * accepted: accepted when mass >= (minmass, or minmass_snr * count * std)
 * and max_intensity >= snr * std (no rough factor), which marks the row's
 * walk disk with the NaN marker in a copy of the frame and takes its count
 * off the divisor; from the second pass on the estimate is recomputed on
 * the marked copy.  Accepted rows (in acceptance order), characterized on
 * the unmarked frame. */
size_t or_detect_frame_adaptive(const float *frame, const or_params *p, uint32_t n_iter,
                                int characterize, or_row *rows, size_t cap);
/* Illumination correction (the shaders in src/shaders/correct_illumination;
 * setup_illumination_profile, src/execute_gpu.rs:849-884; submit_work,
 * :138-166).  Taps: rint = i32(4 sigma + 0.5), w[i + rint] =
 * exp(-(i*i) / (2 (sigma*sigma))) in f32; returns rint. */
int32_t or_illum_taps(float sigma, float *w);
/* smooth_mean_frame.wgsl run as SeparableConvolution<2> with n_frames = 1:
 * pass 0 along columns (rows index u) into tmp, pass 1 along rows into out;
 * per output, taps in order, out-of-frame taps skipped and left out of the
 * norm: fmaf(w, x, sum), norm += w, result sum / norm. */
void or_illum_smooth(const float *in, int32_t H, int32_t W, float sigma, float *out, float *tmp);
/* mean_from_iter (src/execute_gpu.rs:819-847): per pixel, the f32 sum over
 * frames in order, then / f32(T). */
void or_frames_mean(const float *frames, size_t T, size_t n, float *out);
/* correct_illumination.wgsl: out[i] = frame[i] / profile[i]. */
void or_illum_divide(const float *frame, const float *profile, size_t n, float *out);
/* LoG detection with adaptive_background = n_iter: the noise estimate of the
 * raw frame sets the scale-space threshold (rough_snr_factor 0.5 in p); the
 * walked rows of every scale go through the adaptive passes (a row with
 * r > 0 marks its disk of radius ceil(r)), the accepted ones are
 * characterized and pruned. */
size_t or_log_detect_frame_adaptive(const float *frame, const or_params *p, const float *radii,
                                    uint32_t n_radii, uint32_t n_iter, int characterize,
                                    float overlap_threshold, or_row *rows, size_t cap);
void or_bandpass_f64(const float *frame, const or_params *p, double *processed, double *scratch);
/* FP64 centroid of the disk at integer position (pu, pv) of a f64 frame. */
void or_centroid_f64(const double *processed, const or_params *p, int32_t pu, int32_t pv,
                     double *dx, double *dy, double *mass);

#ifdef __cplusplus
}
#endif
#endif
