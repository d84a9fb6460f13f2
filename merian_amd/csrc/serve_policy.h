// Host-only policy of the serving loop (no HIP): first-come-first-served admission of jobs
// by free slots and an HBM byte budget, and latency percentiles. Kept free of device code so
// the policy is unit-tested on the CPU (tests/native/test_serve_policy.cpp).
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <deque>
#include <vector>

namespace mrt {

// Jobs are admitted strictly in submission order: the head of the queue is admitted once a
// slot is free and its bytes fit what the jobs in flight leave of the budget. No job
// overtakes the head, so a large job cannot be starved by a stream of small ones.
class Admission {
  public:
    Admission(uint64_t budget, uint32_t slots) : budget_(budget), slots_(slots) {}

    uint64_t budget() const { return budget_; }
    uint64_t in_use() const { return in_use_; }
    uint32_t slots() const { return slots_; }
    uint32_t busy() const { return busy_; }
    size_t queued() const { return q_.size(); }

    // A job larger than the whole budget can never be admitted.
    bool can_ever_fit(uint64_t bytes) const { return bytes <= budget_; }

    void push(uint64_t id, uint64_t bytes) { q_.push_back(Entry{id, bytes}); }

    // Removes a queued job; false if it is not queued.
    bool remove(uint64_t id) {
        for (auto it = q_.begin(); it != q_.end(); ++it) {
            if (it->id == id) {
                q_.erase(it);
                return true;
            }
        }
        return false;
    }

    // Pops and charges the head of the queue if it can start now.
    bool admit(uint64_t* id, uint64_t* bytes) {
        if (q_.empty() || busy_ >= slots_) return false;
        const Entry e = q_.front();
        if (in_use_ + e.bytes > budget_) return false;
        q_.pop_front();
        in_use_ += e.bytes;
        busy_++;
        *id = e.id;
        *bytes = e.bytes;
        return true;
    }

    // A job in flight finished (or failed) and returns its bytes and slot.
    void release(uint64_t bytes) {
        in_use_ = bytes <= in_use_ ? in_use_ - bytes : 0;
        if (busy_ > 0) busy_--;
    }

  private:
    struct Entry {
        uint64_t id, bytes;
    };
    std::deque<Entry> q_;
    uint64_t budget_ = 0, in_use_ = 0;
    uint32_t slots_ = 1, busy_ = 0;
};

// Nearest-rank percentile (p in [0, 100]) of a sample; 0 for an empty sample.
inline double percentile(std::vector<double> v, double p) {
    if (v.empty()) return 0.0;
    std::sort(v.begin(), v.end());
    const double rank = std::ceil(p / 100.0 * double(v.size()));
    const size_t k = rank < 1.0 ? 0 : std::min(v.size() - 1, size_t(rank) - 1);
    return v[k];
}

// Device bytes a job holds while in flight. Every allocation is rounded to 256 bytes.
struct JobShape {
    uint64_t npix;      // width * height
    uint32_t spp;       // samples per pixel (0 allowed)
    uint32_t L;         // max path length
    bool nee;           // shadow queue needed
    bool demod;         // "demodulate albedo": per-pixel demodulation albedo
    uint64_t max_paths; // paths in flight per job
    uint32_t trace_threads; // resident trace-kernel threads of the device
    uint32_t stack_entries; // global traversal-stack entries per thread the scene needs
};

struct JobLayout {
    uint64_t chunk;     // samples per pass
    uint64_t n_chunks;  // passes
    uint64_t paths;     // npix * chunk
    uint64_t counters;  // u32 counters (4 arrays of n_chunks * (L + 1))
    uint32_t threads;   // traversal-stack columns
    uint64_t bytes;     // total device bytes
};

inline uint64_t round256(uint64_t b) { return (std::max<uint64_t>(b, 1) + 255) & ~uint64_t(255); }

inline JobLayout job_layout(const JobShape& s) {
    JobLayout l{};
    const uint64_t spp = s.spp;
    l.chunk = std::max<uint64_t>(1, std::min<uint64_t>(spp ? spp : 1, s.max_paths / std::max<uint64_t>(1, s.npix)));
    l.n_chunks = spp ? (spp + l.chunk - 1) / l.chunk : 1;
    l.paths = s.npix * l.chunk;
    l.counters = 4 * l.n_chunks * uint64_t(s.L + 1);
    const uint64_t want = ((std::max<uint64_t>(l.paths, s.npix) + 127) / 128) * 128;
    l.threads = uint32_t(std::min<uint64_t>(s.trace_threads, want));
    uint64_t b = 0;
    b += 2 * round256(16 * s.npix);               // gbuffer, irradiance
    b += round256(32 * l.paths);                  // path state
    b += round256(4 * l.paths);                   // signature
    b += 4 * round256(16 * l.paths);              // two ray queues (origin, direction)
    b += round256(16 * l.paths);                  // hits
    b += round256(4 * l.paths);                   // aux (pdf of the last direction)
    if (s.nee) b += 3 * round256(16 * l.paths);   // shadow queue
    b += round256(4 * l.counters);                // queue counters
    b += round256(16 * s.npix) + round256(4 * s.npix); // accumulation, signature sum
    if (s.demod) b += round256(16 * s.npix);      // demodulation albedo
    b += round256(uint64_t(l.threads) * std::max<uint32_t>(1, s.stack_entries) * 4); // stack overflow
    b += round256(64);                            // ray cursor
    l.bytes = b;
    return l;
}

} // namespace mrt
