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

    // The head of the queue, if a slot is free and its bytes fit beside the jobs in flight.
    bool peek(uint64_t* id, uint64_t* bytes) const {
        if (q_.empty() || busy_ >= slots_) return false;
        const Entry e = q_.front();
        if (in_use_ + e.bytes > budget_) return false;
        *id = e.id;
        *bytes = e.bytes;
        return true;
    }

    // Pops and charges the head of the queue if it can start now.
    bool admit(uint64_t* id, uint64_t* bytes) {
        if (!peek(id, bytes)) return false;
        q_.pop_front();
        in_use_ += *bytes;
        busy_++;
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

// Slots keep their device block from job to job, so the steady state allocates nothing (a
// stream-ordered pool that has to grow synchronises with the other streams' work: measured, a
// 1080p job's allocation then waited up to 2.5 s for the jobs in flight). plan_slot picks the
// slot a job of `need` bytes runs in:
//   1. a free slot whose block fits, the smallest such;
//   2. else the free slot with the largest block, grown to `need`; while the blocks held would
//      then exceed the budget, the blocks of other free slots are released, largest first.
// False when no free slot exists or the job cannot fit beside the blocks busy slots hold (it
// waits until one of them finishes).
struct SlotPlan {
    int slot = -1;
    uint64_t block = 0;       // the slot's block for this job (its current one if it fits)
    std::vector<int> release; // other free slots whose blocks are released first
};

inline bool plan_slot(const std::vector<uint64_t>& held, const std::vector<bool>& busy, uint64_t budget,
                      uint64_t need, SlotPlan* plan) {
    *plan = SlotPlan{};
    int fit = -1, largest = -1;
    uint64_t total = 0;
    for (size_t k = 0; k < held.size(); k++) {
        total += held[k];
        if (busy[k]) continue;
        if (held[k] >= need && (fit < 0 || held[k] < held[size_t(fit)])) fit = int(k);
        if (largest < 0 || held[k] > held[size_t(largest)]) largest = int(k);
    }
    if (fit >= 0) {
        plan->slot = fit;
        plan->block = held[size_t(fit)];
        return true;
    }
    if (largest < 0) return false;
    uint64_t after = total - held[size_t(largest)] + need;
    std::vector<int> idle;
    for (size_t k = 0; k < held.size(); k++)
        if (!busy[k] && int(k) != largest && held[k] > 0) idle.push_back(int(k));
    std::sort(idle.begin(), idle.end(), [&](int a, int b) { return held[size_t(a)] > held[size_t(b)]; });
    for (int k : idle) {
        if (after <= budget) break;
        after -= held[size_t(k)];
        plan->release.push_back(k);
    }
    if (after > budget) {
        plan->release.clear();
        return false;
    }
    plan->slot = largest;
    plan->block = need;
    return true;
}

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
    uint64_t bytes;     // device bytes the job's buffers need
    uint64_t block;     // bytes allocated and charged: `bytes` rounded up to a size class
};

inline uint64_t round256(uint64_t b) { return (std::max<uint64_t>(b, 1) + 255) & ~uint64_t(255); }

// Size classes of job blocks: 1 MiB, then 2^k x {1, 1.25, 1.5, 1.75} (at most 25 % over the
// need). A job's buffers are one block from the server's pool, so blocks of a class freed by
// finished jobs serve later jobs of that class without the pool mapping new memory.
inline uint64_t size_class(uint64_t b) {
    const uint64_t min_block = uint64_t(1) << 20;
    if (b <= min_block) return min_block;
    uint64_t p = min_block;
    while (p * 2 < b) p *= 2; // p < b <= 2p
    for (uint64_t q = 5; q <= 8; q++) // 1.25p, 1.5p, 1.75p, 2p
        if (b <= p * q / 4) return p * q / 4;
    return 2 * p;
}

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
    b += 8 * round256(16 * l.paths);              // two ray queues (origin, direction, path state)
    b += round256(16 * l.paths);                  // hits
    if (s.nee) b += 3 * round256(16 * l.paths);   // shadow queue
    b += round256(4 * l.counters);                // queue counters
    b += round256(16 * s.npix) + round256(4 * s.npix); // accumulation, signature sum
    if (s.demod) b += round256(16 * s.npix);      // demodulation albedo
    b += round256(uint64_t(l.threads) * std::max<uint32_t>(1, s.stack_entries) * 4); // stack overflow
    b += round256(64);                            // ray cursor
    l.bytes = b;
    l.block = size_class(b);
    return l;
}

} // namespace mrt
