/*
 * gtrack.h -- C ABI of the MI355X (gfx950) detection pipeline.
 *
 * This is the drop-in boundary for gpu_tracking's per-frame detection path
 * (bandpass -> grey-dilation local max -> centroid walk -> characterize, and
 * the host post-processing that follows it).  Plain C: pointers, sizes and
 * POD structs only.  Each declaration names the reference interface it
 * replaces (paths relative to the gpu_tracking repository).
 *
 * Threading: a gt_context is used by one host thread at a time; only
 * gt_set_interrupt() and gt_progress() may be called from another thread
 * (reference: GpuState is !Send, interruption is an Arc<AtomicBool>,
 * src/gpu_setup.rs:502-514, src/execute_gpu.rs:886-895).
 * Ownership: inputs and outputs belong to the caller; the library owns its
 * device buffers and pinned staging and keeps no caller pointer after a call
 * returns (gt_output memory is released with gt_output_free).
 */
#ifndef GTRACK_H
#define GTRACK_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GT_ABI_VERSION 2

/* One code per variant of the reference's error enum (src/error.rs:6-94),
 * in declaration order, plus two codes that have no reference counterpart. */
typedef enum gt_status {
    GT_OK = 0,
    GT_ERR_GPU_ADAPTER = 1,              /* Error::GpuAdapterError: no HIP device */
    GT_ERR_GPU_DEVICE = 2,               /* Error::GpuDeviceError: HIP runtime failure */
    GT_ERR_FILE_NOT_FOUND = 3,           /* Error::FileNotFound */
    GT_ERR_DIMENSION_MISMATCH = 4,       /* Error::DimensionMismatch */
    GT_ERR_EMPTY_ITERATOR = 5,           /* Error::EmptyIterator */
    GT_ERR_NON_SORTED_CHARACTERIZATION = 6, /* Error::NonSortedCharacterization */
    GT_ERR_FRAME_OUT_OF_BOUNDS = 7,      /* Error::FrameOutOfBounds */
    GT_ERR_FRAME_OOB = 8,                /* Error::FrameOOB */
    GT_ERR_THREAD = 9,                   /* Error::ThreadError */
    GT_ERR_NON_STANDARD_ARRAY_LAYOUT = 10, /* Error::NonStandardArrayLayout */
    GT_ERR_NO_EXTENSION = 11,            /* Error::NoExtensionError */
    GT_ERR_UNSUPPORTED_FILEFORMAT = 12,  /* Error::UnsupportedFileformat */
    GT_ERR_ARRAY_DIMENSIONS = 13,        /* Error::ArrayDimensionsError */
    GT_ERR_CAST = 14,                    /* Error::CastError */
    GT_ERR_READ = 15,                    /* Error::ReadError */
    GT_ERR_INVALID_FILENAME = 16,        /* Error::InvalidFileName */
    GT_ERR_CHANNEL_NOT_FOUND = 17,       /* Error::ChannelNotFound */
    GT_ERR_TOO_DENSE_TO_LINK = 18,       /* Error::TooDenseToLink */
    GT_ERR_INTERRUPTED = 19,             /* Error::Interrupted */
    GT_ERR_POLLED_AFTER_TERMINATION = 20, /* Error::PolledAfterTermination */
    GT_ERR_TIFF_WRITE = 21,              /* Error::TiffWrite */
    GT_ERR_WRONG_BUILDER = 22,           /* Error::WrongBuilder */
    GT_ERR_INVALID_ARGUMENT = 100,       /* no counterpart: rejected argument */
    GT_ERR_CAPACITY = 101                /* no counterpart: a frame exceeded the
                                            candidate/detection capacity (the
                                            reference writes past its buffers) */
} gt_status;

/* ParamStyle (src/gpu_setup.rs:10-33) */
typedef enum gt_style { GT_STYLE_TRACKPY = 0, GT_STYLE_LOG = 1 } gt_style;

/* Pixel type of the frames handed in.  The reference casts every frame to f32
 * on the host (src/decoderiter.rs:16-19); here u16/u8 frames go to the GPU as
 * they are and are widened in the kernel (exact). */
typedef enum gt_dtype { GT_U16 = 0, GT_F32 = 1, GT_U8 = 2 } gt_dtype;

/* TrackingParams + ParamStyle, field for field (src/gpu_setup.rs:10-59).
 * Option<T> fields are a has_* flag plus the value.  keys/frames and
 * linker_reset_points are arrays passed to gt_execute instead. */
typedef struct gt_params {
    int32_t style;                    /* gt_style */
    /* ParamStyle::Trackpy */
    uint32_t separation;
    uint32_t diameter;
    float maxsize;                    /* accepted, unused (as in the reference) */
    float threshold;                  /* accepted, unused */
    int32_t invert;                   /* accepted, unused */
    float percentile;                 /* accepted, unused */
    uint32_t topn;                    /* accepted, unused */
    int32_t preprocess;               /* accepted, unused */
    int32_t filter_close;
    /* ParamStyle::Log */
    float min_radius;
    float max_radius;
    uint32_t n_radii;
    int32_t log_spacing;
    float overlap_threshold;
    /* common */
    float minmass;
    uint32_t max_iterations;
    int32_t characterize;
    int32_t has_search_range;  float search_range;
    int32_t has_memory;        uint32_t memory;
    int32_t doughnut_correction;
    int32_t has_bg_radius;     float bg_radius;
    int32_t has_gap_radius;    float gap_radius;
    int32_t has_snr;           float snr;
    int32_t has_minmass_snr;   float minmass_snr;
    int32_t truncate_preprocessed;
    int32_t has_illumination_sigma; float illumination_sigma;
    int32_t has_adaptive_background; uint32_t adaptive_background;
    int32_t include_r_in_output;
    float shift_threshold;
    float noise_size;
    int32_t has_smoothing_size; uint32_t smoothing_size;
    int32_t illumination_correction_per_frame;
    /* No reference counterpart (0 = off, the default).  LoG style only: a
     * scale-space maximum is kept when its LoG value exceeds threshold_rel
     * times the largest LoG value of its frame over every scale and pixel
     * (skimage blob_log's threshold_rel; the reference allocates a global
     * maximum buffer for this but never writes it, src/gpu_setup.rs:481-488,
     * src/shaders/log_style/logspace_max.wgsl:48-49). */
    float threshold_rel;
} gt_params;

/* ResultRow (src/execute_gpu.rs:48-61): 10 x f32, x = row, y = column. */
typedef struct gt_result_row {
    float x, y, mass, r, max_intensity, Rg, raw_mass, signal, ecc, count;
} gt_result_row;

/* A detection as the device path emits it: the reference's ResultRow plus
 * the raster index (row * ncols + col) of the local maximum the walk started
 * from, which orders detections deterministically, and the frame index. */
typedef struct gt_detection {
    gt_result_row row;
    uint32_t seed;
    uint32_t frame;
} gt_detection;

typedef struct gt_context gt_context;

/* Whole-video result (execute_gpu's (Vec<f32>, Vec<(&str,&str)>) return,
 * src/execute_gpu.rs:886-895, with column_names at :482-519). */
typedef struct gt_output {
    float *data;               /* nrows x ncols, row-major */
    uint64_t nrows;
    uint32_t ncols;
    const char *const *col_names;  /* static strings */
    const char *const *col_types;  /* "int" or "float" */
} gt_output;

/* TrackingParams::default() (src/gpu_setup.rs:61-97). */
void gt_params_default(gt_params *p);

/* ABI version of the loaded library (GT_ABI_VERSION). */
int32_t gt_abi_version(void);

/* Message for the last error raised on the calling thread (never NULL). */
const char *gt_last_error(void);

/* setup_state(&TrackingParams, &[u32;2] dims, characterize_new_points, ..)
 * (src/gpu_setup.rs:611-616): validates and derives the parameters
 * (gpuparams_from_tracking_params, :528-580), selects the HIP device that is
 * current on the calling thread and creates the context's stream. */
gt_status gt_context_create(const gt_params *params, uint32_t nrows, uint32_t ncols,
                            int32_t characterize_points, gt_context **out);
void gt_context_destroy(gt_context *ctx);

/* The device-resident hot path: the reference's per-frame GPU pass sequence
 * (submit_work's order preprocess_rows, preprocess_cols, max_rows,
 * extract_max, walk, characterize -- src/execute_gpu.rs:231-235,368-375)
 * over nframes frames already in device memory.
 *   d_frames: nframes * nrows * ncols pixels of `dtype`, contiguous.
 *   d_out:    nframes * cap_per_frame gt_detection slots; frame f's rows go
 *             to d_out[f * cap_per_frame ...] in unspecified order.
 *   d_counts: nframes uint32; d_counts[f] = number of detections of frame f
 *             (it may exceed cap_per_frame, in which case only cap_per_frame
 *             rows were written and the caller must treat it as an error).
 *   stream:   hipStream_t to enqueue on (NULL = the default stream).
 * Asynchronous: returns after enqueueing.  No per-frame D2H copy. */
gt_status gt_detect_device(gt_context *ctx, const void *d_frames, int32_t dtype,
                           uint32_t nframes, gt_detection *d_out, uint32_t cap_per_frame,
                           uint32_t *d_counts, void *stream);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): gt_detect_device's plain Trackpy-style path, without a
 * context, with its scratch (processed planes, counters, slow-walk lists) in
 * a static 64 MiB device array that each GPU holds until the process ends.
 * It calls no allocating routine; the caller passes the parameters, the
 * frame size, the frames, the outputs and the stream, laid out as for
 * gt_detect_device, and rows and counts equal gt_detect_device's bit for bit.
 * A frame needs 4 B/px of processed plane plus 1 B/px of slow-walk list (and
 * 512 bytes of counters): frames above about 13.4 Mpx (GT_ERR_CAPACITY), the
 * LoG style, snr / minmass_snr / adaptive_background and illumination
 * correction (GT_ERR_INVALID_ARGUMENT) are rejected.  All calls on one GPU
 * share its array: calls that overlap on one GPU (two streams, two host
 * threads) race and must be serialised by the caller. */
gt_status gt_detect_device_static(const gt_params *params, uint32_t nrows, uint32_t ncols,
                                  const void *d_frames, int32_t dtype, uint32_t nframes,
                                  gt_detection *d_out, uint32_t cap_per_frame, uint32_t *d_counts,
                                  void *stream);

/* ---- One frame at a time (the reference's per-frame pair) ---------------
 * gt_submit_frame mirrors submit_work (src/execute_gpu.rs:95-105; its body
 * :106-422): the host frame (n_pixels = nrows * ncols samples of `dtype`) is
 * copied into a pinned staging slot before the call returns, then its copy
 * to the GPU, the pass sequence and the read-back of its count are queued on
 * the context's stream; the call does not wait for them.  frame_idx is
 * handed back by gt_fetch_frame.  Contexts created with characterize_points
 * take the frame's positions instead (pos[0..npos): x = row, y = column,
 * r > 0 a radius), as submit_work's characterization branch (:190-229) with
 * FrameSubsetter's per-frame points (src/linking.rs:94-196); other contexts
 * take pos = NULL, npos = 0.  At most gt_context_set_inflight frames (default
 * 2, the reference's staging ring depth, src/gpu_setup.rs:173-183) are in
 * flight: submitting one more fails until a frame is fetched.
 * Errors: GT_ERR_DIMENSION_MISMATCH when n_pixels differs from the context's
 * frame size (:1074-1080); GT_ERR_INVALID_ARGUMENT when the ring is full, or
 * for whole-video illumination correction (its profile needs every frame
 * first: gt_execute).
 * gt_fetch_frame mirrors get_work (:424-467): it blocks until the oldest
 * frame in flight is done and writes its detections (the reference's
 * ResultRow before the host post-processing: no filter_close, doughnut or
 * linking) to out in ascending order of the raster index of the local
 * maximum each walk started from (position order with characterize_points),
 * with *n = their number and *frame_idx = the frame's submit index.  When
 * *n > cap it returns GT_ERR_CAPACITY and the frame stays in flight, so a
 * second call with a larger buffer gets it; a frame whose candidates or
 * detections overflowed the context's capacity is dropped with
 * GT_ERR_CAPACITY.  gt_detect_device, gt_execute and gt_execute_file fail
 * while frames are in flight. */
gt_status gt_submit_frame(gt_context *ctx, const void *px, int32_t dtype, uint64_t n_pixels,
                          uint64_t frame_idx, const gt_result_row *pos, uint32_t npos);
gt_status gt_fetch_frame(gt_context *ctx, uint64_t *frame_idx, gt_result_row *out, uint32_t cap,
                         uint32_t *n);
/* Frames in flight at most (1..16, default 2); only with none in flight. */
gt_status gt_context_set_inflight(gt_context *ctx, uint32_t depth);

/* Lanes of gt_detect_device (1 or 2, default 1): with 2, the chunks of a call
 * alternate between two streams with their own scratch (plain Trackpy-style
 * detection without the noise estimate, adaptive passes, illumination
 * correction or learned proposal), so one chunk's slow walks and
 * characterize overlap the next chunk's bandpass; the work still starts
 * after, and ends before, the caller's stream.  Results do not depend on it.
 * Throughput rises a few percent; each stage's own time (gt_stage_times)
 * then includes the other lane's concurrent work.  No reference
 * counterpart. */
gt_status gt_context_set_lanes(gt_context *ctx, int32_t lanes);
/* Tail overlap of gt_detect_device (0 or 1, default 0; applies where two
 * lanes would, and only with one lane): every chunk's bandpass and local
 * maxima run on the caller's stream and its slow walks, characterize and
 * count copy on a second stream, so that they overlap the next chunk's
 * bandpass and local maxima; chunks alternate between two sets of scratch.
 * The work still starts after, and ends before, the caller's stream.
 * Results do not depend on it.  No reference counterpart. */
gt_status gt_context_set_overlap(gt_context *ctx, int32_t enable);

/* execute_gpu over an in-memory video (execute_ndarray -> execute_provider ->
 * execute_gpu, src/execute_gpu.rs:1378,1272,886): frames are staged through
 * pinned host memory to the GPU in chunks, detected, read back (counts and
 * rows only) and post-processed on a host thread (filter_close, doughnut
 * photometry, column assembly; src/execute_gpu.rs:671-817).
 *   frames:  nframes * nrows * ncols pixels of `dtype`, contiguous (host).
 *   keys:    optional list of frame indices to process, in order (the
 *            reference's `keys`/`frames`, :1020-1056); NULL = all frames.
 *   points:  characterize_points input, npoints rows of point_cols floats
 *            laid out [frame?, y, x, r?] (src/linking.rs:94-196); NULL
 *            otherwise.  Requires a context created with
 *            characterize_points = 1.
 * On success *out holds the output matrix; release it with gt_output_free. */
gt_status gt_execute(gt_context *ctx, const void *frames, int32_t dtype, uint64_t nframes,
                     const uint64_t *keys, uint64_t nkeys, const float *points,
                     uint64_t npoints, uint32_t point_cols, int32_t points_has_frames,
                     int32_t points_has_r, const uint64_t *linker_reset_points,
                     uint64_t nreset, gt_output *out);
void gt_output_free(gt_output *out);

/* ---- Video files -------------------------------------------------------
 * path_to_iter (src/execute_gpu.rs:1186-1270): tif / tiff (multi-page; the
 * reference decodes them with the tiff crate, src/decoderiter.rs:64-127),
 * ets (MinimalETSParser / ETSIterator, src/decoderiter.rs:351-481) and vsi
 * (the frames of x.vsi live in _x_/stack1/frame_t_0.ets beside it).  Frames
 * decode into the caller's memory in their own sample type: u8 -> GT_U8,
 * u16 -> GT_U16, f32 -> GT_F32, i8 / i16 widened to GT_F32 (the reference
 * widens every type to f32 on the host, src/decoderiter.rs:16-19).
 * TIFF: classic or BigTIFF, either byte order, strips or tiles, one sample
 * per pixel, compression none / PackBits / LZW / Deflate, horizontal
 * predictor; every page has page 0's size and type.  ETS: u16 data.
 * Errors: GT_ERR_NO_EXTENSION, GT_ERR_UNSUPPORTED_FILEFORMAT,
 * GT_ERR_FILE_NOT_FOUND, GT_ERR_CHANNEL_NOT_FOUND, GT_ERR_CAST (u32 / u64 /
 * i32 / i64 / f64 samples), GT_ERR_READ, GT_ERR_DIMENSION_MISMATCH (pages of
 * different sizes), GT_ERR_FRAME_OUT_OF_BOUNDS (gt_file_read past the end). */
typedef struct gt_file gt_file;
/* channel: the ETS channel when has_channel (default 0); ignored for TIFF. */
gt_status gt_file_open(const char *path, int32_t has_channel, uint32_t channel, gt_file **out);
void gt_file_close(gt_file *f);
gt_status gt_file_info(const gt_file *f, uint32_t *nrows, uint32_t *ncols, uint64_t *nframes,
                       int32_t *dtype);
/* Frames first .. first+n-1 into dst: n * nrows * ncols samples of the
 * file's dtype.  One file may be read from several threads at once. */
gt_status gt_file_read(const gt_file *f, uint64_t first, uint64_t n, void *dst);
/* The ETS channels of the file in ascending order (parse_ets,
 * gpu_tracking_python/src/expansion.rs:849-867): returns their number and
 * writes at most cap of them; 0 for TIFF. */
uint32_t gt_file_channels(const gt_file *f, uint32_t *channels, uint32_t cap);
/* mean_from_iter (src/execute_gpu.rs:819-847): per pixel, the f32 sum over
 * the frames in order, divided by f32(nframes); out holds nrows * ncols. */
gt_status gt_file_mean(const gt_file *f, float *out);
/* execute_file (src/execute_gpu.rs:1354-1376): gt_execute over the file's
 * frames, decoded by reader threads straight into the context's pinned
 * staging while the previous chunk runs on the GPU (no f32 widening of
 * integer frames).  The context must have the file's frame size. */
gt_status gt_execute_file(gt_context *ctx, const gt_file *f, const uint64_t *keys, uint64_t nkeys,
                          const float *points, uint64_t npoints, uint32_t point_cols,
                          int32_t points_has_frames, int32_t points_has_r,
                          const uint64_t *linker_reset_points, uint64_t nreset, gt_output *out);
/* to_tiff (gpu_tracking_python/src/expansion.rs:925-950): one uncompressed
 * 16-bit grey page per frame, values cast as Rust's `f32 as u16` (toward
 * zero, saturating at 0 and 65535, NaN -> 0); width = ncols (the reference
 * passes the row count as the width). */
gt_status gt_write_tiff(const char *path, const float *frames, uint64_t nframes, uint32_t nrows,
                        uint32_t ncols);

/* Column names and types for a parameter set (column_names,
 * src/execute_gpu.rs:482-519).  Returns the column count; the arrays point
 * to static strings and hold at most 16 entries. */
uint32_t gt_column_names(const gt_params *params, const char **names, const char **types);

/* filter_close_trackpy (src/execute_gpu.rs:556-590) over one frame's rows:
 * keep[i] = 0 when another row j lies strictly closer than `separation`
 * (squared distance < separation^2, src/linking.rs:238-243) and has a larger
 * mass, or the same mass and j < i.  Host-only, uniform-grid neighbour search. */
gt_status gt_filter_close(const gt_result_row *rows, uint64_t n, float separation,
                          uint8_t *keep);

/* Doughnut photometry of post_process (src/execute_gpu.rs:717-790) with the
 * stencils circle_inds/annulus_inds (src/kernels.rs:138-171): for each row,
 * out3[3i..3i+2] = raw_mass, raw_bg_median, raw_mass_corrected over the raw
 * f32 frame.  Host-only. */
gt_status gt_doughnut(const float *raw, uint32_t nrows, uint32_t ncols, const gt_result_row *rows,
                      uint64_t n, float radius, float bg_radius, float gap_radius, float *out3);

/* LoG post-processing of one frame (src/execute_gpu.rs:699-712 with
 * prune_blobs_log, :592-669): when overlap_threshold < 1, rows are stably
 * sorted by r descending then mass descending, and each surviving row in that
 * order drops every surviving other row with squared distance < (2 r)^2 whose
 * disk overlap (:597-630) exceeds the threshold.  Survivors are compacted in
 * sorted order into rows[0..*n_out).  Host-only. */
gt_status gt_prune_blobs(gt_result_row *rows, uint64_t n, float overlap_threshold,
                         uint64_t *n_out);

/* Frame-to-frame linking of detections given as rows [frame, y, x] sorted by
 * frame (linker_all over FrameSubsetter groups, src/linking.rs:94-196 and
 * :818-831, with Linker at :346-521 and link at :564-816): ids[i] receives
 * row i's particle id.  search_range > 0; memory = frames a particle may be
 * missing.  GT_ERR_TOO_DENSE_TO_LINK when a subnetwork is too large,
 * GT_ERR_NON_SORTED_CHARACTERIZATION when frames decrease.  Host-only. */
gt_status gt_link(const float *points, uint64_t n, float search_range, uint32_t memory,
                  uint64_t *ids);

/* Links two sets of detections frame by frame (connect_all,
 * src/linking.rs:960-1012, bound as connect_rust at
 * gpu_tracking_python/src/python_bindings.rs:293-311): points1 (n1 rows)
 * and points2 (n2 rows) are [frame, y, x] rows, each sorted by frame and
 * grouped as gt_link groups them.  In ascending frame order, a frame present
 * in both sets is linked set 1 -> set 2 with memory 0 (Linker::connect,
 * :405-432): set 1's points take fresh ids in row order, set 2's points the
 * id of the set-1 point they link to, else a fresh id; a frame present in
 * one set only takes fresh ids (fill_base, :914-919).  ids1[i] / ids2[i]
 * receive the ids; one id counter serves both sets.  Where the reference's
 * sync_iterators (:945-955) gives ids twice to a frame of the lagging set
 * that skips past the frame it was catching up to (its id list then comes
 * out longer than its rows), that frame takes its ids once, when reached.
 * Errors as gt_link.  Host-only. */
gt_status gt_connect(const float *points1, uint64_t n1, const float *points2, uint64_t n2,
                     float search_range, uint64_t *ids1, uint64_t *ids2);

/* ---- Batched detection over frames of different sizes ------------------
 * No counterpart in the reference: one GpuState serves one frame size
 * (src/execute_gpu.rs:1074-1080).  A batch context holds the parameters
 * (Trackpy style, u8/u16 frames) and scratch for launches of up to
 * max_launch_pixels pixels in up to max_slots frames; gt_batch_detect runs
 * any mix of frame sizes through one kernel per stage per launch (grid y =
 * frame of the launch's table, grid x = tile), uploading every launch's table
 * at once and replaying a HIP graph of the stage sequence once per launch.
 * Rows of frames[i] land at d_out[i * cap_per_frame], their count at
 * d_counts[i]; each row's `frame`
 * field is frames[i].frame.  Results equal gt_detect_device's on the same
 * frames bit for bit. */
typedef struct gt_frame_desc {
    const void *data;      /* device address of the frame (row-major) */
    uint32_t nrows, ncols;
    uint32_t frame;        /* frame field of its detections */
    uint32_t pad;
} gt_frame_desc;

typedef struct gt_batch gt_batch;

gt_status gt_batch_create(const gt_params *params, uint64_t max_frame_pixels,
                          uint64_t max_launch_pixels, uint32_t max_slots, gt_batch **out);
void gt_batch_destroy(gt_batch *b);
gt_status gt_batch_detect(gt_batch *b, const gt_frame_desc *frames, uint32_t n, int32_t dtype,
                          gt_detection *d_out, uint32_t cap_per_frame, uint32_t *d_counts,
                          void *stream);
/* Autotuner: times candidate schedules on the sample frames (pixels per
 * launch, then the slow-walk / characterize grid, then the lanes; coordinate
 * search, warm-up + the faster of two timed runs each) and keeps the fastest for later
 * calls; *best_ms receives its time. */
gt_status gt_batch_tune(gt_batch *b, const gt_frame_desc *frames, uint32_t n, int32_t dtype,
                        gt_detection *d_out, uint32_t cap_per_frame, uint32_t *d_counts,
                        float *best_ms);
/* The schedule in use (pixels per launch, slow-walk blocks per frame, frames
 * per launch at most, the context's pixel budget per launch) and the number
 * of launches of the last call. */
gt_status gt_batch_get_schedule(gt_batch *b, uint64_t *sched4, uint64_t *launches);
/* The whole schedule: the knobs above, the lanes, the tile shape of the
 * bandpass and detect kernels (fixed: 64 columns, one per lane of a
 * wavefront, by 32 rows; 16 and 64 rows were measured slower at every batch
 * frame size) and the last call's grid (tile workgroups per launch, frames
 * per launch, most runs of one frame size in a launch).  No reference
 * counterpart: the reference has no batched path. */
typedef struct gt_batch_schedule_info {
    uint64_t pixels_per_launch;     /* pixel budget per launch in use */
    uint64_t max_pixels_per_launch; /* the context's limit */
    uint32_t walk_blocks;           /* slow-walk / characterize blocks per frame */
    uint32_t max_frames_per_launch;
    uint32_t lanes;
    uint32_t tile_rows, tile_cols;
    uint32_t launches;              /* launches of the last call */
    uint32_t tile_blocks;           /* tile workgroups per launch, last call */
    uint32_t launch_slots;          /* frames per launch the grids cover, last call */
    uint32_t max_segments;          /* most size runs in one launch, last call */
    int32_t graph;                  /* graph replay on */
} gt_batch_schedule_info;
gt_status gt_batch_get_schedule_info(gt_batch *b, gt_batch_schedule_info *out);
/* Graph replay on (default) or direct launches. */
gt_status gt_batch_set_graph(gt_batch *b, int32_t enable);
/* Lanes (1..4): launch k of a call runs on lane k % lanes, each lane a stream
 * with its own scratch and graph, so consecutive launches overlap; results do
 * not depend on the lane count.  Default 1; gt_batch_tune also picks it. */
gt_status gt_batch_set_lanes(gt_batch *b, uint32_t lanes);
uint32_t gt_batch_get_lanes(const gt_batch *b);

/* Interruption and progress (the Arc<AtomicBool> / Arc<Mutex<(usize,
 * Option<usize>)>> pair of execute_gpu, src/execute_gpu.rs:1094-1099,
 * :974-977).  Safe to call from any thread while gt_execute runs. */
gt_status gt_set_interrupt(gt_context *ctx, int32_t value);
gt_status gt_progress(gt_context *ctx, uint64_t *cur, uint64_t *total);

/* Per-stage device time of the most recent gt_detect_device call when timing
 * is enabled (hipEvents on the launch stream around each stage; adds a few
 * microseconds per chunk).  stage_ms[0..n) receives the summed time of
 * bandpass, local max, walk+characterize, in that order; *launches the
 * number of launches per stage and *frames_per_launch the chunk size.
 * Returns the number of stages written (0 when nothing was timed).  There is
 * no reference counterpart (the reference has no GPU timestamps). */
gt_status gt_set_timing(gt_context *ctx, int32_t enable);
int32_t gt_stage_times(gt_context *ctx, double *stage_ms, int32_t n, uint32_t *launches,
                       uint32_t *frames_per_launch);

/* The noise estimate the snr and minmass_snr rules use: the sample standard
 * deviation of each raw frame (StdArray, src/utils.rs:298-670: an eight-way
 * summation tree; NaN-marker pixels, bits 0xFFFFFFFF, are left out of the
 * sums but not the pixel count), for nframes device frames of the context's
 * size; d_std receives nframes floats.  Asynchronous on `stream` (NULL = the
 * default stream). */
gt_status gt_frame_std(gt_context *ctx, const void *d_frames, int32_t dtype, uint32_t nframes,
                       float *d_std, void *stream);

/* ---- Learned centroid proposal ---------------------------------------------
 * No counterpart in the reference.  The reference's walk is the iterative
 * centroid loop of src/shaders/walk.wgsl:104-198.  Here a small CNN predicts,
 * for every candidate pixel, where that walk ends.  One walk evaluation at the
 * predicted centre verifies the prediction; the full walk runs when the
 * verification fails or the confidence is low.
 * The network: the 11x11 processed patch around the candidate (divided by
 * its largest |value|, bf16) -> conv3x3 1->16 -> 16->32 -> 32->32 (valid,
 * ReLU, bf16 activations, fp32 accumulation on the matrix cores) -> mean
 * pooling -> linear 32->3: (dy, dx, confidence logit).
 * Parameters: a float32 vector of GT_CNN_NPARAM entries, torch layouts in the
 * order w1 (16,1,3,3) b1 (16) w2 (32,16,3,3) b2 (32) w3 (32,32,3,3) b3 (32)
 * wfc (3,32) bfc (3).
 * Training samples: patches bf16 [n][GT_CNN_PATCH_PITCH] (entries 121..127
 * zero) and targets float32 [n][4] = (end row - candidate row, end column -
 * candidate column, 1 if the walk converges within max_iterations, the
 * candidate's raster index as uint32 bits), all in device memory.  Loss: label * |d - t|^2 + BCE(logit, label), batch mean;
 * Adam with torch.optim.Adam's arithmetic (betas 0.9, 0.999; eps 1e-8). */
#define GT_CNN_NPARAM 14147
#define GT_CNN_PATCH_PITCH 128
typedef struct gt_cnn gt_cnn;
/* A network with He-uniform weights drawn from `seed` and zero biases. */
gt_status gt_cnn_create(uint64_t seed, gt_cnn **out);
void gt_cnn_destroy(gt_cnn *net);
/* Host float32[GT_CNN_NPARAM]; setting resets the Adam state. */
gt_status gt_cnn_set_params(gt_cnn *net, const float *params);
gt_status gt_cnn_get_params(gt_cnn *net, float *params);
gt_status gt_cnn_adam_steps(gt_cnn *net, uint64_t *steps);
/* d_out: float32 [n][4] = (dy, dx, logit, 0) for n stored patches;
 * asynchronous on `stream` (NULL = the default stream). */
gt_status gt_cnn_forward(gt_cnn *net, const uint16_t *d_patches, uint64_t n, float *d_out,
                         void *stream);
/* Gradient, step and train run on the network's own stream and return when
 * done; the samples and indices must already be complete in device memory.
 * Gradient (host float32[GT_CNN_NPARAM]) and loss of the batch of samples
 * d_idx[0..nb) (device uint32 indices, each < the number of samples). */
gt_status gt_cnn_gradient(gt_cnn *net, const uint16_t *d_patches, const float *d_targets,
                          const uint32_t *d_idx, uint32_t nb, float *grad, float *loss);
/* One Adam step on the batch d_idx[0..nb); *loss = the batch loss before it. */
gt_status gt_cnn_step(gt_cnn *net, const uint16_t *d_patches, const float *d_targets,
                      const uint32_t *d_idx, uint32_t nb, float lr, float *loss);
/* `epochs` passes over the n samples in batches of `batch` (a fresh seeded
 * shuffle per epoch); loss_per_epoch[e] = mean batch loss of epoch e. */
gt_status gt_cnn_train(gt_cnn *net, const uint16_t *d_patches, const float *d_targets,
                       uint64_t n, uint32_t epochs, uint32_t batch, float lr, uint64_t seed,
                       float *loss_per_epoch);
/* Training samples of every candidate of nframes device frames, using the
 * context's detection parameters (tiled Trackpy style): the plain method's
 * full walk from each candidate gives the target.  Writes at most cap
 * samples; *n_written = samples written, *n_total = candidates seen.  Runs
 * on `stream` (NULL = the default stream) and returns when done. */
gt_status gt_cnn_make_samples(gt_context *ctx, const void *d_frames, int32_t dtype,
                              uint32_t nframes, uint16_t *d_patches, float *d_targets,
                              uint64_t cap, uint64_t *n_written, uint64_t *n_total, void *stream);
/* Detect with the network's proposals (NULL: the plain walk again).
 * scope GT_CNN_MOVING: every candidate takes one walk step first; a walk
 * that stops there is finished (the plain result), and the network proposes
 * for the walks still moving.  GT_CNN_ALL: the network proposes for every
 * candidate.  Proposals with logit < conf_logit, and proposals whose
 * verifying evaluation moves or lands more than tol px from the prediction,
 * take the rest of the full walk.  The context keeps the pointer: the
 * network must outlive its use, and is not trained while the context
 * detects. */
#define GT_CNN_MOVING 0
#define GT_CNN_ALL 1
gt_status gt_context_set_cnn(gt_context *ctx, const gt_cnn *net, float conf_logit, float tol,
                             int32_t scope);
/* Accepted proposals and candidates since the last reset. */
gt_status gt_context_cnn_stats(gt_context *ctx, uint64_t *accepted, uint64_t *candidates,
                               int32_t reset);

#ifdef __cplusplus
}
#endif
#endif /* GTRACK_H */
