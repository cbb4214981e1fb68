// gt_linker.h -- frame-to-frame particle linking on the host (trackpy-style
// optimal subnetwork linking with memory), the Linker of src/linking.rs.
#pragma once

#include <atomic>
#include <cstdint>
#include <deque>
#include <string>
#include <vector>

#include "gtrack.h"

class gt_linker {
  public:
    struct Entry {
        float x, y;      // row, column
        uint64_t id;     // particle id
    };
    struct Book {
        uint64_t start, duration;
    };

    gt_linker(float search_range, uint32_t memory);
    // Links one frame's points (xy: n x 2, row then column) to the previous
    // ones; ids[i] receives point i's particle id (Linker::advance,
    // src/linking.rs:435-498).
    gt_status advance(const float *xy, size_t n, uint64_t *ids, std::atomic<int32_t> *interrupt,
                      std::string *err);
    // Durations of every particle id so far (Linker::finish, :500-519).
    std::vector<Book> finish();
    // Links two point sets of one frame (Linker::connect, :405-432): set 1's
    // points take fresh ids in order (ids1), set 2's points the id of the
    // set-1 point they link to or a fresh one (ids2).  Leaves the frame
    // history alone.
    gt_status connect(const float *xy1, size_t n1, const float *xy2, size_t n2, uint64_t *ids1,
                      uint64_t *ids2, std::string *err);
    // Fresh ids for n points of a frame with no partner (fill_base, :914-919).
    void fresh(size_t n, uint64_t *ids) {
        for (size_t i = 0; i < n; ++i) ids[i] = part_idx_++;
    }

  private:
    gt_status link(size_t memory_start, uint64_t *ids, std::vector<Entry> *unused,
                   std::atomic<int32_t> *interrupt, std::string *err);
    bool solve(size_t progress, const uint32_t *dests, size_t nd_sub,
               std::atomic<int32_t> *interrupt);

    float radius_;
    uint32_t memory_;
    std::vector<Entry> prev_;
    size_t prev_n_ = 0;
    std::deque<std::vector<Entry>> memvec_;
    uint64_t frame_idx_ = 0;
    uint64_t part_idx_ = 0;
    std::vector<Book> books_;

    // per-link scratch, kept across frames (no per-frame allocation once the
    // largest frame has been seen)
    struct Cand {
        uint32_t idx;   // the other end (destination of a source, or source of a destination)
        float d2;       // squared distance
    };
    std::vector<Entry> dest_;
    // destinations ordered by strip of rows, then index: dkey_ (strip of each
    // destination), strip_ (the ordered points), soff_ (each strip's run when
    // the strips are few) or skey_ (the ordered strips, else)
    struct Pt {
        float x, y;
        uint32_t idx;
    };
    std::vector<double> dkey_, skey_;
    std::vector<Pt> strip_;
    std::vector<uint32_t> soff_;
    std::vector<std::pair<double, uint32_t>> order_;
    // candidate lists in compressed rows: s2d_ (per source, ascending
    // destination) and d2s_ (per destination, ascending source)
    std::vector<uint32_t> s2d_off_, d2s_off_, fill_;
    std::vector<Cand> s2d_, d2s_;
    std::vector<Cand> near_;
    std::vector<uint8_t> vis_src_, vis_dest_, taken_;
    std::vector<int64_t> out_;            // particle id per destination, -1 = none
    std::vector<int64_t> out_src_;        // source linked to each destination, -1 = none
    std::vector<uint32_t> sub_dests_, sub_off_;   // subnetworks: their destinations
    std::vector<uint32_t> psrc_, pdest_;
    struct Frame {
        uint32_t node;
        uint8_t kind;   // 0 source, 1 destination
        uint32_t next;
    };
    std::vector<Frame> stack_;
    std::vector<uint64_t> mark_;   // per particle id: tag of the frame that last listed it
    uint64_t mark_tag_ = 0;
    // subnetwork search state: the current partial assignment (destination,
    // source or -1 for none) and the best one so far
    float score_ = 0.0f, best_ = 0.0f;
    std::vector<std::pair<uint32_t, int64_t>> cur_, best_asg_;
    bool interrupted_ = false;
};
