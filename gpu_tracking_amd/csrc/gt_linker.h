// gt_linker.h -- frame-to-frame particle linking on the host (trackpy-style
// optimal subnetwork linking with memory), the Linker of src/linking.rs.
#pragma once

#include <atomic>
#include <cstdint>
#include <deque>
#include <string>
#include <unordered_map>
#include <unordered_set>
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
    bool solve(size_t progress, const std::vector<uint32_t> &dests,
               std::atomic<int32_t> *interrupt);

    float radius_;
    uint32_t memory_;
    std::vector<Entry> prev_;
    size_t prev_n_ = 0;
    std::deque<std::vector<Entry>> memvec_;
    uint64_t frame_idx_ = 0;
    uint64_t part_idx_ = 0;
    std::vector<Book> books_;

    // per-link scratch
    std::vector<Entry> dest_;
    std::vector<std::vector<std::pair<uint32_t, float>>> s2d_, d2s_;
    std::vector<uint8_t> vis_src_, vis_dest_;
    std::vector<int64_t> out_;            // particle id per destination, -1 = none
    // subnetwork search state
    float score_ = 0.0f, best_ = 0.0f;
    std::unordered_map<uint64_t, uint32_t> used_;
    std::unordered_set<uint32_t> nulls_;
    bool interrupted_ = false;
};
