// CPU unit test of the serving loop's admission policy and latency percentiles
// (merian_amd/csrc/serve_policy.h). Built and run by tests/test_serve_policy.py with g++.
#include "../../merian_amd/csrc/serve_policy.h"

#include <cstdio>
#include <cstdlib>

static int failures = 0;
#define CHECK(c)                                                         \
    do {                                                                 \
        if (!(c)) {                                                      \
            std::printf("FAILED %s:%d: %s\n", __FILE__, __LINE__, #c);   \
            failures++;                                                  \
        }                                                                \
    } while (0)

int main() {
    using namespace mrt;
    uint64_t id = 0, bytes = 0;
    {   // slots bound the jobs in flight
        Admission a(1000, 2);
        for (uint64_t k = 1; k <= 4; k++) a.push(k, 100);
        CHECK(a.admit(&id, &bytes) && id == 1);
        CHECK(a.admit(&id, &bytes) && id == 2);
        CHECK(!a.admit(&id, &bytes));
        CHECK(a.busy() == 2 && a.in_use() == 200 && a.queued() == 2);
        a.release(100);
        CHECK(a.admit(&id, &bytes) && id == 3);
    }
    {   // bytes bound the jobs in flight; strict submission order (no overtaking)
        Admission a(1000, 8);
        a.push(1, 600);
        a.push(2, 600);
        a.push(3, 100);
        CHECK(a.admit(&id, &bytes) && id == 1 && bytes == 600);
        CHECK(!a.admit(&id, &bytes)); // 2 does not fit and 3 may not overtake it
        a.release(600);
        CHECK(a.in_use() == 0 && a.busy() == 0);
        CHECK(a.admit(&id, &bytes) && id == 2);
        CHECK(a.admit(&id, &bytes) && id == 3);
        CHECK(a.in_use() == 700);
    }
    {   // never-fitting jobs, removal of queued jobs
        Admission a(1000, 1);
        CHECK(!a.can_ever_fit(1001) && a.can_ever_fit(1000));
        a.push(1, 10);
        a.push(2, 10);
        CHECK(a.remove(2) && !a.remove(2) && a.queued() == 1);
        CHECK(a.admit(&id, &bytes) && id == 1);
        CHECK(!a.remove(1)); // in flight, not queued
        a.release(10);
        a.release(10); // extra release does not underflow
        CHECK(a.in_use() == 0 && a.busy() == 0);
    }
    {   // nearest-rank percentiles
        std::vector<double> v;
        for (int k = 100; k >= 1; k--) v.push_back(double(k));
        CHECK(percentile(v, 50) == 50.0);
        CHECK(percentile(v, 99) == 99.0);
        CHECK(percentile(v, 100) == 100.0);
        CHECK(percentile(v, 0) == 1.0);
        CHECK(percentile({}, 99) == 0.0);
        CHECK(percentile({7.0}, 99) == 7.0);
        CHECK(percentile({1.0, 2.0, 3.0}, 50) == 2.0);
    }
    {   // job layout: chunking by the per-job path limit, byte totals
        JobShape s{};
        s.npix = 1920 * 1080;
        s.spp = 64;
        s.L = 8;
        s.max_paths = 1u << 22;
        s.trace_threads = 524288;
        s.stack_entries = 30;
        JobLayout l = job_layout(s);
        CHECK(l.chunk == 2 && l.n_chunks == 32 && l.paths == 2 * s.npix);
        CHECK(l.counters == 4 * 32 * 9);
        CHECK(l.threads == 524288);
        s.nee = true;
        const JobLayout ln = job_layout(s);
        CHECK(ln.bytes == l.bytes + 3 * round256(16 * l.paths));
        s.spp = 0;
        l = job_layout(s);
        CHECK(l.chunk == 1 && l.n_chunks == 1);
        s.npix = 100;
        s.spp = 3;
        s.max_paths = 1000;
        l = job_layout(s);
        CHECK(l.chunk == 3 && l.n_chunks == 1 && l.threads == 384);
        s.max_paths = 50; // fewer than one sample of every pixel: still one sample per pass
        l = job_layout(s);
        CHECK(l.chunk == 1 && l.n_chunks == 3 && l.paths == 100);
        CHECK(l.block == (uint64_t(1) << 20) && l.block >= l.bytes);
    }
    {   // slot blocks kept from job to job
        SlotPlan p;
        // best fit among free slots whose block fits; busy slots are never chosen
        CHECK(plan_slot({100, 40, 60, 0}, {false, false, false, true}, 1000, 50, &p) && p.slot == 2 && p.block == 60 &&
              p.release.empty());
        CHECK(plan_slot({100, 40, 60, 0}, {true, false, false, false}, 1000, 50, &p) && p.slot == 2);
        // nothing fits: the largest free block grows
        CHECK(plan_slot({10, 40, 0}, {false, false, true}, 1000, 50, &p) && p.slot == 1 && p.block == 50 &&
              p.release.empty());
        // growing would pass the budget: other free blocks are released, largest first
        CHECK(plan_slot({30, 40, 20, 25}, {false, false, false, true}, 120, 70, &p) && p.slot == 1 && p.block == 70 &&
              p.release.size() == 1 && p.release[0] == 0);
        CHECK(plan_slot({30, 40, 20, 25}, {false, false, false, true}, 100, 70, &p) && p.release.size() == 2);
        CHECK(!plan_slot({30, 40, 20, 25}, {false, false, false, true}, 90, 70, &p)); // 25 busy + 70 > 90
        // busy blocks leave too little: wait
        CHECK(!plan_slot({80, 10}, {true, false}, 100, 30, &p) && p.slot == -1 && p.release.empty());
        // no free slot
        CHECK(!plan_slot({10, 10}, {true, true}, 100, 5, &p));
        // an empty free slot with nothing held
        CHECK(plan_slot({0, 0}, {false, false}, 100, 5, &p) && p.block == 5);
    }
    {   // size classes: 1 MiB, then 2^k x {1, 1.25, 1.5, 1.75}; never below the need, at most 25 % above
        const uint64_t M = uint64_t(1) << 20;
        CHECK(size_class(1) == M && size_class(M) == M);
        CHECK(size_class(M + 1) == M * 5 / 4 && size_class(M * 5 / 4) == M * 5 / 4);
        CHECK(size_class(M * 5 / 4 + 1) == M * 3 / 2 && size_class(M * 7 / 4 + 1) == 2 * M);
        CHECK(size_class(3 * M) == 3 * M && size_class(12ull << 30) == 12ull << 30);
        for (uint64_t b = 1; b < (uint64_t(1) << 36); b = b * 3 + 7) {
            const uint64_t c = size_class(b);
            CHECK(c >= b && (b <= M || double(c) <= 1.25 * double(b) + 1.0));
            CHECK(size_class(c) == c); // a class is its own class
        }
    }
    if (failures) return 1;
    std::printf("serve policy: all checks passed\n");
    return 0;
}
