// gt_linker.cpp -- the Linker of src/linking.rs (:346-521, link at :564-816)
// in C++, with a search over strips of rows in place of the kd-tree.
//
// A source's candidate destinations are listed in ascending destination
// index (the reference lists them in kd-tree order); that fixes the
// depth-first order of each subnetwork and so which of several equal-cost
// assignments wins.  Scores are f32 sums in the reference's order.
#include "gt_linker.h"

#include <algorithm>
#include <cmath>
#include <limits>

gt_linker::gt_linker(float search_range, uint32_t memory)
    : radius_(search_range), memory_(memory) {
    for (uint32_t i = 0; i < memory_; ++i) memvec_.emplace_back();
}

gt_status gt_linker::link(size_t memory_start, uint64_t *ids, std::vector<Entry> *unused,
                          std::atomic<int32_t> *interrupt, std::string *err) {
    const std::vector<Entry> &src = prev_;
    const size_t ns = src.size(), nd = dest_.size();
    const float r2 = radius_ * radius_;
    s2d_off_.assign(ns + 1, 0u);
    s2d_.clear();
    // candidate pairs: squared distance (0 + dx^2) + dy^2 < r^2, dx = dest -
    // source (within_radius_rd2, linking.rs:208-247).  The destinations are
    // ordered by strip (floor of row / w, w slightly wider than the radius so
    // rounding cannot hide a pair one strip further) and then by index; a
    // source's partners lie in the one run covering its strip and the two
    // beside it.
    if (ns && nd && radius_ > 0.0f) {
        const double inv = 1.0 / (static_cast<double>(radius_) * 1.0001);
        double kmin = std::numeric_limits<double>::infinity(), kmax = -kmin;
        dkey_.resize(nd);
        size_t nfin = 0;
        for (size_t d = 0; d < nd; ++d) {
            if (!std::isfinite(dest_[d].x) || !std::isfinite(dest_[d].y)) {
                dkey_[d] = std::numeric_limits<double>::quiet_NaN();
                continue;
            }
            const double k = std::floor(static_cast<double>(dest_[d].x) * inv);
            dkey_[d] = k;
            kmin = std::min(kmin, k);
            kmax = std::max(kmax, k);
            ++nfin;
        }
        strip_.resize(nfin);
        const bool dense = nfin && kmax - kmin <= 4.0 * static_cast<double>(nd) + 4096.0;
        if (dense) {
            // counting sort by strip; strip k is strip_[soff_[k - kmin] .. soff_[k - kmin + 1])
            const size_t S = static_cast<size_t>(kmax - kmin) + 1;
            soff_.assign(S + 1, 0u);
            for (size_t d = 0; d < nd; ++d)
                if (dkey_[d] == dkey_[d]) ++soff_[static_cast<size_t>(dkey_[d] - kmin) + 1];
            for (size_t k = 0; k < S; ++k) soff_[k + 1] += soff_[k];
            fill_.assign(soff_.begin(), soff_.end() - 1);
            for (size_t d = 0; d < nd; ++d)
                if (dkey_[d] == dkey_[d])
                    strip_[fill_[static_cast<size_t>(dkey_[d] - kmin)]++] = {
                        dest_[d].x, dest_[d].y, static_cast<uint32_t>(d)};
        } else {
            order_.clear();
            for (size_t d = 0; d < nd; ++d)
                if (dkey_[d] == dkey_[d]) order_.emplace_back(dkey_[d], static_cast<uint32_t>(d));
            std::sort(order_.begin(), order_.end());
            skey_.resize(nfin);
            for (size_t i = 0; i < nfin; ++i) {
                const uint32_t d = order_[i].second;
                skey_[i] = order_[i].first;
                strip_[i] = {dest_[d].x, dest_[d].y, d};
            }
        }
        for (size_t s = 0; s < ns; ++s) {
            s2d_off_[s] = static_cast<uint32_t>(s2d_.size());
            if (!nfin || !std::isfinite(src[s].x) || !std::isfinite(src[s].y)) continue;
            const double k = std::floor(static_cast<double>(src[s].x) * inv);
            size_t lo, hi;
            if (dense) {
                const double a = std::max(k - 1.0, kmin), b = std::min(k + 1.0, kmax);
                if (a > b) continue;
                lo = soff_[static_cast<size_t>(a - kmin)];
                hi = soff_[static_cast<size_t>(b - kmin) + 1];
            } else {
                lo = static_cast<size_t>(std::lower_bound(skey_.begin(), skey_.end(), k - 1.0) -
                                         skey_.begin());
                hi = static_cast<size_t>(std::upper_bound(skey_.begin(), skey_.end(), k + 1.0) -
                                         skey_.begin());
            }
            near_.clear();
            for (size_t q = lo; q < hi; ++q) {
                const float dx = strip_[q].x - src[s].x, dy = strip_[q].y - src[s].y;
                float d2 = 0.0f;
                d2 = d2 + dx * dx;
                d2 = d2 + dy * dy;
                if (d2 < r2) near_.push_back({strip_[q].idx, d2});
            }
            // ascending destination index (a handful of entries)
            for (size_t i = 1; i < near_.size(); ++i)
                for (size_t j = i; j > 0 && near_[j - 1].idx > near_[j].idx; --j)
                    std::swap(near_[j - 1], near_[j]);
            s2d_.insert(s2d_.end(), near_.begin(), near_.end());
        }
    }
    s2d_off_[ns] = static_cast<uint32_t>(s2d_.size());
    // each destination's sources, ascending
    d2s_off_.assign(nd + 1, 0u);
    for (const Cand &c : s2d_) ++d2s_off_[c.idx + 1];
    for (size_t d = 0; d < nd; ++d) d2s_off_[d + 1] += d2s_off_[d];
    d2s_.resize(s2d_.size());
    fill_.assign(nd, 0u);
    for (size_t s = 0; s < ns; ++s)
        for (uint32_t q = s2d_off_[s]; q < s2d_off_[s + 1]; ++q) {
            const Cand &c = s2d_[q];
            d2s_[d2s_off_[c.idx] + fill_[c.idx]++] = {static_cast<uint32_t>(s), c.d2};
        }
    // connected components by depth-first search from each destination, in
    // the visiting order of the reference's recursion
    vis_src_.assign(ns, 0);
    vis_dest_.assign(nd, 0);
    out_.assign(nd, -1);
    out_src_.assign(nd, -1);
    sub_dests_.clear();
    sub_off_.assign(1, 0u);
    size_t load = 0;
    stack_.clear();
    for (size_t i = 0; i < nd; ++i) {
        if (vis_dest_[i]) continue;
        // the common components, without the stack: a destination with no
        // candidate (alone), and a destination and a source that are each
        // other's only candidate (linked); the search below finds the same
        const uint32_t nsrc = d2s_off_[i + 1] - d2s_off_[i];
        if (nsrc == 0) {
            vis_dest_[i] = 1;
            continue;
        }
        if (nsrc == 1) {
            const uint32_t s = d2s_[d2s_off_[i]].idx;
            if (s2d_off_[s + 1] - s2d_off_[s] == 1) {
                vis_dest_[i] = 1;
                vis_src_[s] = 1;
                out_[i] = static_cast<int64_t>(src[s].id);
                out_src_[i] = s;
                continue;
            }
        }
        psrc_.clear();
        pdest_.clear();
        {
            vis_dest_[i] = 1;
            pdest_.push_back(static_cast<uint32_t>(i));
            stack_.push_back({static_cast<uint32_t>(i), 1, 0});
            while (!stack_.empty()) {
                Frame &f = stack_.back();
                if (f.kind == 1) {
                    const uint32_t lo = d2s_off_[f.node], n = d2s_off_[f.node + 1] - lo;
                    if (f.next >= n) { stack_.pop_back(); continue; }
                    const uint32_t s = d2s_[lo + f.next++].idx;
                    if (!vis_src_[s]) {
                        vis_src_[s] = 1;
                        psrc_.push_back(s);
                        stack_.push_back({s, 0, 0});
                    }
                } else {
                    const uint32_t lo = s2d_off_[f.node], n = s2d_off_[f.node + 1] - lo;
                    if (f.next >= n) { stack_.pop_back(); continue; }
                    const uint32_t d = s2d_[lo + f.next++].idx;
                    if (!vis_dest_[d]) {
                        vis_dest_[d] = 1;
                        pdest_.push_back(d);
                        stack_.push_back({d, 1, 0});
                    }
                }
            }
        }
        if (pdest_.size() == 1) {
            if (psrc_.size() == 1) {
                out_[i] = static_cast<int64_t>(src[psrc_[0]].id);
                out_src_[i] = psrc_[0];
            } else if (psrc_.size() > 1) {
                sub_dests_.insert(sub_dests_.end(), pdest_.begin(), pdest_.end());
                sub_off_.push_back(static_cast<uint32_t>(sub_dests_.size()));
            }
        } else if (pdest_.size() > 1) {
            load += psrc_.size() * pdest_.size();
            sub_dests_.insert(sub_dests_.end(), pdest_.begin(), pdest_.end());
            sub_off_.push_back(static_cast<uint32_t>(sub_dests_.size()));
        }
    }
    if (load > 5000) {
        *err = "computational load during linking is abnormally high. try setting a smaller\n"
               " search range or track with settings that give fewer detections.";
        return GT_ERR_TOO_DENSE_TO_LINK;
    }
    // each subnetwork: exhaustive search for the assignment of minimum total
    // squared distance, r^2 for a destination left unlinked
    for (size_t k = 0; k + 1 < sub_off_.size(); ++k) {
        cur_.clear();
        score_ = 0.0f;
        best_ = std::numeric_limits<float>::infinity();
        interrupted_ = false;
        solve(0, sub_dests_.data() + sub_off_[k], sub_off_[k + 1] - sub_off_[k], interrupt);
        if (interrupted_) {
            *err = "Keyboard Interrupt, my custom version";
            return GT_ERR_INTERRUPTED;
        }
    }
    // unlinked destinations start new particles, in destination order; the
    // sources nobody linked to (ids are unique among the sources) are unused
    taken_.assign(ns, 0);
    for (size_t d = 0; d < nd; ++d) {
        if (out_[d] < 0) out_[d] = static_cast<int64_t>(part_idx_++);
        else taken_[static_cast<size_t>(out_src_[d])] = 1;
        ids[d] = static_cast<uint64_t>(out_[d]);
    }
    unused->clear();
    const size_t lim = std::min(memory_start, ns);
    for (size_t s = 0; s < lim; ++s)
        if (!taken_[s]) unused->push_back(src[s]);
    return GT_OK;
}

// recurse2 (linking.rs:694-774); cur_ holds the partial assignment (a
// destination and its source, or -1 for none), and a better complete one
// is written to out_ / out_src_
bool gt_linker::solve(size_t progress, const uint32_t *dests, size_t nd_sub,
                      std::atomic<int32_t> *interrupt) {
    const uint32_t cur = dests[progress];
    const uint32_t lo = d2s_off_[cur], nc = d2s_off_[cur + 1] - lo;
    const float r2 = radius_ * radius_;
    for (uint32_t c = 0; c <= nc; ++c) {
        if (interrupt && interrupt->load(std::memory_order_relaxed)) {
            interrupted_ = true;
            return false;
        }
        if (score_ > best_) return true;
        const bool is_null = c == nc;
        const float cost = is_null ? r2 : d2s_[lo + c].d2;
        score_ = score_ + cost;
        int64_t s = -1;
        if (!is_null) {
            s = d2s_[lo + c].idx;
            bool used = false;
            for (const auto &e : cur_) used = used || e.second == s;
            if (used) {
                score_ = score_ - cost;
                continue;
            }
        }
        cur_.emplace_back(cur, s);
        if (progress + 1 == nd_sub) {
            if (score_ < best_) {
                best_ = score_;
                for (const auto &e : cur_) {
                    out_[e.first] = e.second >= 0 ? static_cast<int64_t>(prev_[e.second].id) : -1;
                    out_src_[e.first] = e.second;
                }
            }
        } else {
            if (!solve(progress + 1, dests, nd_sub, interrupt)) return false;
        }
        cur_.pop_back();
        score_ = score_ - cost;
    }
    return true;
}

gt_status gt_linker::advance(const float *xy, size_t n, uint64_t *ids,
                             std::atomic<int32_t> *interrupt, std::string *err) {
    size_t memory_start = static_cast<size_t>(-1);
    if (memory_ > 0) {
        // remembered particles join the sources unless already among them
        // (every id is below part_idx_; a particle's mark is the tag of the
        // last frame that listed it)
        if (mark_.size() < part_idx_) mark_.resize(part_idx_, 0);
        const uint64_t tag = ++mark_tag_;
        for (const auto &e : prev_) mark_[e.id] = tag;
        for (const auto &vec : memvec_)
            for (const auto &e : vec)
                if (mark_[e.id] != tag) {
                    mark_[e.id] = tag;
                    prev_.push_back(e);
                }
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
