// Warm-start cache bookkeeping: WarmStartCache (rans-uq/include/ForwardModel.hpp:21-157)
// restated for flow states that stay in device memory. An entry carries the
// active parameter vector, the payload (the owner's handle of the stored state)
// and its byte count; lookups, replacement and eviction follow the reference,
// and every payload the cache lets go of is handed back to the owner.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstddef>
#include <deque>
#include <limits>
#include <stdexcept>
#include <vector>

namespace rq {

struct WarmCache {
    struct Entry {
        std::vector<double> theta;
        void *payload;
        size_t bytes;
    };

    // ForwardModel.hpp:33-37 (the defaults are ForwardModel's, ForwardModel.cpp:17)
    explicit WarmCache(int maxSize = 50, double threshold = 0.5,
                       size_t byteBudget = 64 * 1024 * 1024)
        : max_size_(maxSize), threshold_(threshold), byte_budget_(byteBudget) {
        if (maxSize < 0 || !std::isfinite(threshold) || threshold < 0)
            throw std::invalid_argument("Invalid warm-start cache limits");
    }

    // WarmStartCache::entryBytes (ForwardModel.hpp:41-50): the parameter
    // vector plus interior and boundary values of U (3 doubles) and of the
    // eight scalar fields p, k, omega, nuT, F1, F2, Pk, CDkw
    static size_t entry_bytes(size_t n_theta, int n_cells, int n_boundary) {
        const size_t n = (size_t)n_cells + (size_t)n_boundary;
        return n_theta * sizeof(double) + n * 3 * sizeof(double) + 8 * n * sizeof(double);
    }

    // WarmStartCache::nearest (ForwardModel.hpp:130-141): the first entry
    // strictly closer than every earlier one and than the threshold
    const Entry *nearest(const double *theta, size_t n) const {
        const Entry *best = nullptr;
        double best_dist = threshold_;
        for (const Entry &e : entries_) {
            const double d = euclidean(theta, n, e.theta);
            if (d < best_dist) {
                best_dist = d;
                best = &e;
            }
        }
        return best;
    }

    // WarmStartCache::store (ForwardModel.hpp:72-91). Returns false when the
    // new state is not retained (its payload stays with the caller); payloads
    // of replaced or evicted entries are appended to `freed`.
    bool store(const double *theta, size_t n, void *payload, size_t bytes,
               std::vector<void *> &freed) {
        for (size_t i = 0; i < n; ++i)
            if (!std::isfinite(theta[i])) return false;
        for (auto it = entries_.begin(); it != entries_.end(); ++it) {
            if (it->theta.size() == n && std::equal(it->theta.begin(), it->theta.end(), theta)) {
                retained_ -= it->bytes;
                freed.push_back(it->payload);
                entries_.erase(it);
                break;
            }
        }
        if (max_size_ == 0 || bytes > byte_budget_) return false;
        while (!entries_.empty() && (entries_.size() >= (size_t)max_size_ ||
                                     retained_ > byte_budget_ - bytes))
            evict_front(freed);
        entries_.push_back({std::vector<double>(theta, theta + n), payload, bytes});
        retained_ += bytes;
        return true;
    }

    // WarmStartCache::setByteBudget (ForwardModel.hpp:111-116)
    void set_byte_budget(size_t bytes, std::vector<void *> &freed) {
        byte_budget_ = bytes;
        while (!entries_.empty() && retained_ > byte_budget_) evict_front(freed);
    }

    void clear(std::vector<void *> &freed) {
        for (const Entry &e : entries_) freed.push_back(e.payload);
        entries_.clear();
        retained_ = 0;
    }

    int size() const { return (int)entries_.size(); }
    size_t retained_bytes() const { return retained_; }
    size_t byte_budget() const { return byte_budget_; }

  private:
    int max_size_;
    double threshold_;
    size_t byte_budget_;
    size_t retained_ = 0;
    std::deque<Entry> entries_;

    void evict_front(std::vector<void *> &freed) {
        retained_ -= entries_.front().bytes;
        freed.push_back(entries_.front().payload);
        entries_.pop_front();
    }

    // ForwardModel.hpp:144-153
    static double euclidean(const double *a, size_t n, const std::vector<double> &b) {
        if (n != b.size()) return std::numeric_limits<double>::infinity();
        double s = 0;
        for (size_t i = 0; i < n; ++i) {
            const double d = a[i] - b[i];
            s += d * d;
        }
        return std::sqrt(s);
    }
};

}  // namespace rq
