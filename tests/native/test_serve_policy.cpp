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
    }
    if (failures) return 1;
    std::printf("serve policy: all checks passed\n");
    return 0;
}
