// autotune.cpp -- empirical choice of the sweep's launch shape and method
// threshold for one prepared query batch.
//
// Coordinate descent over the knobs that change the sweep's cost without
// changing its result: the sparse/dense threshold, the tile size (DB elements
// per thread), the work-queue chunk and the resident blocks per CU.  Each
// candidate is planned, swept once to warm up and then timed over `reps`
// sweeps with HIP events on the index's stream; the fastest value of a knob
// is kept before the next knob is tried.
#include <cstdio>
#include <string>
#include <vector>

#include "engine.hpp"

namespace chk {

double time_sweeps(Query& q, uint64_t* d_counts, uint64_t* d_bp, int reps) {
    Index& ix = *q.idx;
    hipStream_t st = ix.stream;
    query_plan(q);
    sweep_run(q, d_counts, d_bp, st);
    hipEvent_t a, b;
    CHK_HIP(hipEventCreate(&a));
    CHK_HIP(hipEventCreate(&b));
    CHK_HIP(hipEventRecord(a, st));
    for (int r = 0; r < reps; r++) sweep_run(q, d_counts, d_bp, st);
    CHK_HIP(hipEventRecord(b, st));
    CHK_HIP(hipEventSynchronize(b));
    float ms = 0.f;
    CHK_HIP(hipEventElapsedTime(&ms, a, b));
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    return (double)ms / reps;
}

namespace {

double autotune_body(Query& q, uint64_t* d_counts, uint64_t* d_bp, int reps, std::string* report);

}  // namespace

double autotune(Query& q, uint64_t* d_counts, uint64_t* d_bp, int reps, std::string* report) {
    // a candidate that throws (allocation failure while re-planning, a plan
    // limit) leaves the index's tuning and profiling as they were on entry
    Index& ix = *q.idx;
    const Tuning saved = ix.tune;
    const bool prof = ix.prof.on;
    try {
        const double ms = autotune_body(q, d_counts, d_bp, reps, report);
        ix.prof.on = prof;
        return ms;
    } catch (...) {
        ix.tune = saved;
        ix.tune_version++;
        ix.prof.on = prof;
        throw;
    }
}

namespace {

double autotune_body(Query& q, uint64_t* d_counts, uint64_t* d_bp, int reps, std::string* report) {
    Index& ix = *q.idx;
    reps = reps > 0 ? reps : 3;
    ix.prof.on = false;
    Tuning& t = ix.tune;
    char line[160];
    auto run = [&](const char* knob, uint64_t v) {
        ix.tune_version++;
        const double ms = time_sweeps(q, d_counts, d_bp, reps);
        if (report) {
            std::snprintf(line, sizeof line, "%s=%llu: %.4f ms\n", knob, (unsigned long long)v, ms);
            *report += line;
        }
        return ms;
    };
    double best = run("start", 0);
    auto descend = [&](const char* knob, auto get, auto set, const std::vector<uint64_t>& vals) {
        const uint64_t keep = get();
        uint64_t bv = keep;
        for (uint64_t v : vals) {
            if (v == keep) continue;
            set(v);
            const double ms = run(knob, v);
            if (ms < best) {
                best = ms;
                bv = v;
            }
        }
        set(bv);
        ix.tune_version++;  // the next plan follows the kept value, not the last candidate
    };
    // method threshold (files above it go dense): only matters when the
    // batch has files on both sides of it; 1 << 62 sends every file sparse
    {
        const uint64_t n = std::max<uint64_t>(ix.n, 1);
        std::vector<uint64_t> vals = {n / 16384 + 1, n / 4096 + 1, n / 1024 + 1, n / 256 + 1,
                                      1ull << 62};
        descend("sweep.sparse_max", [&] { return t.sparse_max; },
                [&](uint64_t v) { t.sparse_max = v; }, vals);
    }
    // dense launch shape (no effect when every file went sparse)
    query_plan(q);
    if (!q.jobs_host.empty()) {
        descend("sweep.items", [&] { return (uint64_t)t.items; },
                [&](uint64_t v) { t.items = (uint32_t)v; }, {4, 8, 12, 16});
        descend("sweep.chunk", [&] { return (uint64_t)t.chunk; },
                [&](uint64_t v) { t.chunk = (uint32_t)v; }, {2, 8, 32});
        descend("sweep.blocks_per_cu", [&] { return (uint64_t)t.blocks_per_cu; },
                [&](uint64_t v) { t.blocks_per_cu = (uint32_t)v; }, {2, 3});
    }
    ix.tune_version++;
    query_plan(q);
    return best;
}

}  // namespace

}  // namespace chk
