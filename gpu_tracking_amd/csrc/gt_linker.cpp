// gt_linker.cpp -- the Linker of src/linking.rs (:346-521, link at :564-816)
// in C++, with a uniform-grid neighbour search in place of the kd-tree.
//
// A source's candidate destinations are listed in ascending destination
// index (the reference lists them in kd-tree order); that fixes the
// depth-first order of each subnetwork and so which of several equal-cost
// assignments wins.  Scores are f32 sums in the reference's order.
#include "gt_linker.h"

#include <algorithm>
#include <cmath>
#include <limits>

namespace {

int64_t cell_of(float v, double cell) {
    return static_cast<int64_t>(std::floor(static_cast<double>(v) / cell));
}

}  // namespace

gt_linker::gt_linker(float search_range, uint32_t memory)
    : radius_(search_range), memory_(memory) {
    for (uint32_t i = 0; i < memory_; ++i) memvec_.emplace_back();
}

gt_status gt_linker::link(size_t memory_start, uint64_t *ids, std::vector<Entry> *unused,
                          std::atomic<int32_t> *interrupt, std::string *err) {
    const std::vector<Entry> &src = prev_;
    const size_t ns = src.size(), nd = dest_.size();
    const float r2 = radius_ * radius_;
    s2d_.assign(ns, {});
    d2s_.assign(nd, {});
    // candidate pairs: squared distance (0 + dx^2) + dy^2 < r^2, dx = dest -
    // source (within_radius_rd2, linking.rs:208-247); grid cells slightly
    // wider than the radius so rounding cannot hide a pair one cell further
    if (ns && nd && radius_ > 0.0f) {
        const double cell = static_cast<double>(radius_) * 1.0001;
        std::unordered_map<int64_t, std::vector<uint32_t>> grid;
        grid.reserve(nd);
        auto key = [](int64_t a, int64_t b) { return (a << 32) ^ (b & 0xFFFFFFFFll); };
        for (size_t d = 0; d < nd; ++d) {
            if (!std::isfinite(dest_[d].x) || !std::isfinite(dest_[d].y)) continue;
            grid[key(cell_of(dest_[d].x, cell), cell_of(dest_[d].y, cell))].push_back(
                static_cast<uint32_t>(d));
        }
        for (size_t s = 0; s < ns; ++s) {
            if (!std::isfinite(src[s].x) || !std::isfinite(src[s].y)) continue;
            const int64_t cx = cell_of(src[s].x, cell), cy = cell_of(src[s].y, cell);
            auto &lst = s2d_[s];
            for (int64_t a = -1; a <= 1; ++a)
                for (int64_t b = -1; b <= 1; ++b) {
                    auto it = grid.find(key(cx + a, cy + b));
                    if (it == grid.end()) continue;
                    for (uint32_t d : it->second) {
                        const float dx = dest_[d].x - src[s].x, dy = dest_[d].y - src[s].y;
                        float d2 = 0.0f;
                        d2 = d2 + dx * dx;
                        d2 = d2 + dy * dy;
                        if (d2 < r2) lst.emplace_back(d, d2);
                    }
                }
            std::sort(lst.begin(), lst.end(),
                      [](const std::pair<uint32_t, float> &a, const std::pair<uint32_t, float> &b) {
                          return a.first < b.first;
                      });
            for (const auto &e : lst) d2s_[e.first].emplace_back(static_cast<uint32_t>(s), e.second);
        }
    }
    // connected components by depth-first search from each destination, in
    // the visiting order of the reference's recursion
    vis_src_.assign(ns, 0);
    vis_dest_.assign(nd, 0);
    out_.assign(nd, -1);
    std::vector<std::vector<uint32_t>> subnets;   // destinations of each subnetwork
    size_t load = 0;
    struct Frame {
        uint32_t node;
        uint8_t kind;   // 0 source, 1 destination
        uint32_t next;
    };
    std::vector<Frame> stack;
    std::vector<uint32_t> psrc, pdest;
    for (size_t i = 0; i < nd; ++i) {
        psrc.clear();
        pdest.clear();
        if (!vis_dest_[i]) {
            vis_dest_[i] = 1;
            pdest.push_back(static_cast<uint32_t>(i));
            stack.push_back({static_cast<uint32_t>(i), 1, 0});
            while (!stack.empty()) {
                Frame &f = stack.back();
                if (f.kind == 1) {
                    const auto &adj = d2s_[f.node];
                    if (f.next >= adj.size()) { stack.pop_back(); continue; }
                    const uint32_t s = adj[f.next++].first;
                    if (!vis_src_[s]) {
                        vis_src_[s] = 1;
                        psrc.push_back(s);
                        stack.push_back({s, 0, 0});
                    }
                } else {
                    const auto &adj = s2d_[f.node];
                    if (f.next >= adj.size()) { stack.pop_back(); continue; }
                    const uint32_t d = adj[f.next++].first;
                    if (!vis_dest_[d]) {
                        vis_dest_[d] = 1;
                        pdest.push_back(d);
                        stack.push_back({d, 1, 0});
                    }
                }
            }
        }
        if (pdest.size() == 1) {
            if (psrc.size() == 1) out_[i] = static_cast<int64_t>(src[psrc[0]].id);
            else if (psrc.size() > 1) subnets.push_back(pdest);
        } else if (pdest.size() > 1) {
            load += psrc.size() * pdest.size();
            subnets.push_back(pdest);
        }
    }
    if (load > 5000) {
        *err = "computational load during linking is abnormally high. try setting a smaller\n"
               " search range or track with settings that give fewer detections.";
        return GT_ERR_TOO_DENSE_TO_LINK;
    }
    // each subnetwork: exhaustive search for the assignment of minimum total
    // squared distance, r^2 for a destination left unlinked
    for (const auto &dests : subnets) {
        used_.clear();
        nulls_.clear();
        score_ = 0.0f;
        best_ = std::numeric_limits<float>::infinity();
        interrupted_ = false;
        solve(0, dests, interrupt);
        if (interrupted_) {
            *err = "Keyboard Interrupt, my custom version";
            return GT_ERR_INTERRUPTED;
        }
    }
    // unlinked destinations start new particles, in destination order
    std::unordered_set<uint64_t> taken;
    taken.reserve(nd * 2 + 1);
    for (size_t d = 0; d < nd; ++d) {
        if (out_[d] < 0) out_[d] = static_cast<int64_t>(part_idx_++);
        ids[d] = static_cast<uint64_t>(out_[d]);
        taken.insert(ids[d]);
    }
    unused->clear();
    const size_t lim = std::min(memory_start, ns);
    for (size_t s = 0; s < lim; ++s)
        if (!taken.count(src[s].id)) unused->push_back(src[s]);
    return GT_OK;
}

// recurse2 (linking.rs:694-774)
bool gt_linker::solve(size_t progress, const std::vector<uint32_t> &dests,
                      std::atomic<int32_t> *interrupt) {
    const uint32_t cur = dests[progress];
    const auto &cands = d2s_[cur];
    const float r2 = radius_ * radius_;
    for (size_t c = 0; c <= cands.size(); ++c) {
        if (interrupt && interrupt->load(std::memory_order_relaxed)) {
            interrupted_ = true;
            return false;
        }
        if (score_ > best_) return true;
        const bool is_null = c == cands.size();
        const float cost = is_null ? r2 : cands[c].second;
        score_ = score_ + cost;
        uint64_t ident = 0;
        if (!is_null) {
            ident = prev_[cands[c].first].id;
            if (used_.count(ident)) {
                score_ = score_ - cost;
                continue;
            }
            used_.emplace(ident, cur);
        } else {
            nulls_.insert(cur);
        }
        if (progress + 1 == dests.size()) {
            if (score_ < best_) {
                best_ = score_;
                for (const auto &kv : used_) out_[kv.second] = static_cast<int64_t>(kv.first);
                for (uint32_t v : nulls_) out_[v] = -1;
            }
        } else {
            if (!solve(progress + 1, dests, interrupt)) return false;
        }
        if (!is_null) used_.erase(ident);
        else nulls_.erase(cur);
        score_ = score_ - cost;
    }
    return true;
}

gt_status gt_linker::advance(const float *xy, size_t n, uint64_t *ids,
                             std::atomic<int32_t> *interrupt, std::string *err) {
    size_t memory_start = static_cast<size_t>(-1);
    if (memory_ > 0) {
        // remembered particles join the sources unless already among them
        std::unordered_set<uint64_t> seen;
        seen.reserve(prev_.size() * 2 + 1);
        for (const auto &e : prev_) seen.insert(e.id);
        for (const auto &vec : memvec_)
            for (const auto &e : vec)
                if (seen.insert(e.id).second) prev_.push_back(e);
        memory_start = prev_n_;
    }
    dest_.resize(n);
    for (size_t i = 0; i < n; ++i) dest_[i] = {xy[2 * i], xy[2 * i + 1], 0};
    const uint64_t before = part_idx_;
    std::vector<Entry> unused;
    gt_status st = link(memory_start, ids, &unused, interrupt, err);
    if (st != GT_OK) return st;
    for (uint64_t k = before; k < part_idx_; ++k) books_.push_back({frame_idx_, 0});
    prev_.resize(n);
    for (size_t i = 0; i < n; ++i) prev_[i] = {dest_[i].x, dest_[i].y, ids[i]};
    if (memory_ > 0) {
        memvec_.push_back(std::move(unused));
        std::vector<Entry> leaving = std::move(memvec_.front());
        memvec_.pop_front();
        for (const auto &e : leaving) {
            Book &b = books_[e.id];
            b.duration = (frame_idx_ - b.start) - memvec_.size();
        }
    } else {
        for (const auto &e : unused) {
            Book &b = books_[e.id];
            b.duration = frame_idx_ - b.start;
        }
    }
    prev_n_ = n;
    ++frame_idx_;
    return GT_OK;
}

gt_status gt_linker::connect(const float *xy1, size_t n1, const float *xy2, size_t n2,
                             uint64_t *ids1, uint64_t *ids2, std::string *err) {
    prev_.resize(n1);
    for (size_t i = 0; i < n1; ++i) {
        ids1[i] = part_idx_++;
        prev_[i] = {xy1[2 * i], xy1[2 * i + 1], ids1[i]};
    }
    dest_.resize(n2);
    for (size_t i = 0; i < n2; ++i) dest_[i] = {xy2[2 * i], xy2[2 * i + 1], 0};
    std::vector<Entry> unused;
    gt_status st = link(static_cast<size_t>(-1), ids2, &unused, nullptr, err);
    prev_.clear();
    return st;
}

std::vector<gt_linker::Book> gt_linker::finish() {
    std::vector<Book> books = books_;
    const size_t L = memvec_.size();
    size_t i = 0;
    for (const auto &vec : memvec_) {
        for (const auto &e : vec) {
            Book &b = books[e.id];
            b.duration = (frame_idx_ - b.start) - (L - i);
        }
        ++i;
    }
    for (const auto &e : prev_) {
        Book &b = books[e.id];
        b.duration = frame_idx_ - b.start;
    }
    return books;
}
