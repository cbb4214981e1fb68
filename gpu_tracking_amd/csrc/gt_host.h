// gt_host.h -- host-side logic of the runtime that needs no GPU: parameter
// defaults and derivation, column names, frame planning, and the per-frame
// post-processing of execute_gpu (filter_close, doughnut, output assembly).
#pragma once

#include <algorithm>
#include <atomic>
#include <cstdlib>
#include <condition_variable>
#include <deque>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <utility>
#include <vector>

#include "gt_kparams.h"
#include "gt_linker.h"
#include "gtrack.h"

// Parameters after validation and derivation (gpuparams_from_tracking_params,
// src/gpu_setup.rs:528-580).
struct gt_derived {
    int32_t style = GT_STYLE_TRACKPY;
    uint32_t smoothing = 9;     // preprocess kernel side s
    float sigma2 = 1.0f;        // noise_size (used as a variance)
    uint32_t circle = 9;        // diameter
    uint32_t dilation = 12;     // k
    int32_t margin = 4;
    uint32_t separation = 10;
    uint32_t max_iterations = 10;
    float shift_threshold = 0.6f;
    float minmass = 0.0f;
    float snr = 0.0f, minmass_snr = 0.0f, rough = 1.0f;
    int32_t need_std = 0;       // snr, minmass_snr or adaptive_background: noise estimate
    uint32_t adaptive = 0;      // adaptive_background passes (0: off)
    float illum_sigma = 0.0f;   // illumination correction (0: off)
    int32_t illum_per_frame = 0;
    int32_t truncate = 1;
    int32_t characterize = 0;
    int32_t filter_close = 1;
    int32_t doughnut = 0;
    float bg_radius = 0.0f;     // resolved doughnut outer radius
    float gap_radius = 0.0f;
    int32_t include_r = 0;
    // LoG style
    std::vector<float> radii;   // the radius scan (gpu_setup.rs:958-984)
    float overlap_threshold = 0.0f;
    float threshold_rel = 0.0f; // 0: off
    int32_t bg_radius_given = 0;  // LoG: else the ring's outer radius is 2 r per row
};

void gt_host_params_default(gt_params *p);
bool gt_host_derive(const gt_params *p, uint32_t H, uint32_t W, gt_derived *d, std::string *why);
void gt_host_fill_kparams(const gt_derived *d, uint32_t H, uint32_t W, GtKParams *kp);
// Whether q = RN(a * RN(1/b)), then RN(q + fma(-q, b, a) * RN(1/b)) equals
// IEEE a / b for every a in [1, 2) (all 2^23 mantissas; cached per divisor).
bool gt_host_one_correction_exact(float b);
// Whether RN(a*rh + RN(a*rl)) (rh = RN(1/b), rl = RN(1/b - rh)) equals a / b
// for every mantissa of a binade (then for every dividend whose a*rl is
// normal); rh, rl receive the split reciprocal.
bool gt_host_split_exact(float b, float *rh, float *rl);
// The LoG scale table (radii, tap radii, tap weights and norms).
void gt_host_fill_logtable(const gt_derived *d, GtLogTable *t);
// LoG post-processing of one frame (execute_gpu.rs:699-712): stable sort by
// r descending then mass descending, then prune_blobs_log (:592-669) when
// overlap_threshold < 1.  Compacts rows in place; returns the survivor count.
size_t gt_host_prune_blobs(gt_result_row *rows, size_t n, float overlap_threshold);
uint32_t gt_host_column_names(const gt_params *p, const char **names, const char **types);
void gt_host_widen(const void *src, int32_t dtype, size_t n, float *dst);
// The doughnut's disk and ring offsets (i, j), i-major (kernels.rs:138-171).
void gt_host_doughnut_offsets(float radius, float bg_radius, float gap_radius,
                              std::vector<std::pair<int32_t, int32_t>> *sig,
                              std::vector<std::pair<int32_t, int32_t>> *ring);

// Which frames to run, in order, and (characterize_points) which positions
// each of them gets (FrameSubsetter, src/linking.rs:94-196; the keys/frames
// subset, src/execute_gpu.rs:1020-1056).
struct gt_host_plan {
    std::vector<uint64_t> frames;
    std::vector<size_t> pt_begin, pt_end;   // per planned frame, into pts
    std::vector<gt_result_row> pts;
};

// Records the message gt_last_error() returns on this thread and returns st
// (defined in gt_runtime.cpp).
gt_status gt_runtime_fail(gt_status st, const char *fmt, ...);

gt_status gt_host_plan_frames(int32_t characterize_points, uint64_t nframes, const uint64_t *keys,
                              uint64_t nkeys, const float *points, uint64_t npoints,
                              uint32_t point_cols, int32_t has_frames, int32_t has_r,
                              gt_host_plan *plan, std::string *why);

// Per-frame post-processing and output assembly (post_process,
// src/execute_gpu.rs:671-817, and the tail of execute_gpu, :1165-1181).
class gt_host_post {
  public:
    gt_host_post(const gt_params *p, const gt_derived *d, uint32_t H, uint32_t W,
                 const uint64_t *reset_points, uint64_t nreset);
    ~gt_host_post();
    gt_host_post(const gt_host_post &) = delete;
    gt_host_post &operator=(const gt_host_post &) = delete;
    bool ok() const { return error_.empty(); }
    const std::string &error() const { return error_; }
    bool needs_raw() const { return d_.doughnut != 0; }
    // rows: the frame's detections in any order (sorted here by seed).
    // post4, when given: per row (raw_mass, bg_median, corrected, keep) as the
    // GPU computed them (Trackpy style): filter_close and the doughnut are
    // taken from it instead of being computed here (raw may then be null).
    gt_status process(std::vector<gt_detection> &rows, uint64_t frame_index, const float *raw,
                      std::atomic<int32_t> *interrupt, std::string *err,
                      const float *post4 = nullptr);
    gt_status finish(gt_output *out, std::string *err);

  private:
    gt_status assemble(uint64_t frame_index, std::atomic<int32_t> *interrupt, std::string *err);
    // Linking runs on its own thread, frame by frame in order, beside the
    // per-frame work: each frame's positions and first output row go to it
    // as a job; the particle ids land in row_ids_ and are written into the
    // output by finish().
    struct LinkJob {
        uint64_t frame;
        size_t row0;
        std::vector<float> xy;
    };
    void link_loop();
    gt_status stop_linking(std::string *err);
    gt_params p_;
    gt_derived d_;
    uint32_t H_, W_;
    uint32_t ncols_ = 0;
    // The output block, row-major, grown with realloc (large blocks move
    // without copying) and handed to gt_output as it is.
    struct OutBuf {
        float *p = nullptr;
        size_t n = 0, cap = 0;
        OutBuf() = default;
        OutBuf(const OutBuf &) = delete;
        OutBuf &operator=(const OutBuf &) = delete;
        ~OutBuf() { std::free(p); }
        size_t size() const { return n; }
        float *data() { return p; }
        float &operator[](size_t i) { return p[i]; }
        bool resize(size_t m) {   // new elements are left unset
            if (m > cap) {
                const size_t c = std::max(m, cap + cap / 2 + 4096);
                float *q = static_cast<float *>(std::realloc(p, sizeof(float) * c));
                if (!q) return false;
                p = q;
                cap = c;
            }
            n = m;
            return true;
        }
        float *release() {
            float *q = p;
            p = nullptr;
            n = cap = 0;
            return q;
        }
    };
    OutBuf out_;
    std::string error_;
    std::vector<gt_result_row> tmp_rows_;
    std::vector<uint8_t> keep_;
    std::vector<float> dough_;
    std::vector<uint64_t> keys_;
    // linking (search_range): one linker per segment between reset points
    std::vector<uint64_t> reset_;          // sorted frame indices
    std::unique_ptr<gt_linker> linker_;
    std::vector<std::vector<gt_linker::Book>> seg_books_;   // finished segments
    std::vector<uint32_t> row_seg_;        // segment of each output row
    std::vector<uint64_t> row_ids_;        // particle id of each output row
    std::thread link_thread_;
    std::mutex link_mu_;
    std::condition_variable link_cv_;
    std::deque<LinkJob> link_q_;
    bool link_stop_ = false;
    gt_status link_status_ = GT_OK;        // under link_mu_
    std::string link_err_;                 // under link_mu_
    std::atomic<int32_t> *link_interrupt_ = nullptr;
};

// Implemented in gt_runtime.cpp (needs the device).
struct gt_context;
gt_status gt_host_characterize_points(gt_context *c, int slot, size_t base, uint32_t nf,
                                      int32_t dtype, const gt_host_plan &plan);
