// server.cpp -- serving form: a request loop over an HBM-resident index with
// admission by HBM footprint and continuous batching.
//
// Clients submit query files (one file per request) from any thread.  One
// scheduler thread owns the device work.  Whenever a batch slot is free it
// takes the waiting requests in arrival order, as many as the HBM budget and
// the per-batch file limit allow, places them as one multi-file query batch
// and enqueues the sweep (and the p-values) on that slot's own stream, so
// batches admitted while others run overlap on the device.  Completed slots
// are found by polling their events; each request's row of counts, bp and
// p-values is copied out and its waiter woken.  Slots keep their device and
// pinned host buffers between batches, so the steady state allocates
// nothing.  A request's row depends only on its own intervals and the index,
// never on the batch it shared, so results are identical alone and in a mix.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>

#include "dbfile.hpp"
#include "engine.hpp"

namespace chk {

using Clock = std::chrono::steady_clock;

// A request's intervals (shard, start and end of up to cap intervals in one
// allocation).  Left uninitialised: the submit copy writes every element it
// keeps.  Buffers go back to the server's spare list when a request
// completes, so steady serving neither zero-fills nor page-faults them.
struct IvBuf {
    std::unique_ptr<uint32_t[]> p;
    uint64_t cap = 0;
    uint32_t* shard() const { return p.get(); }
    uint32_t* start() const { return p.get() + cap; }
    uint32_t* end() const { return p.get() + 2 * cap; }
};

struct Request {
    uint64_t id = 0;
    IvBuf iv;        // on shards of the DB only
    uint64_t n = 0;  // intervals kept
    std::vector<uint64_t> fs_n;               // intervals per shard
    double total_len = 0.0;
    uint32_t max_end = 0;
    uint64_t n_given = 0;
    uint64_t footprint = 0;
    Clock::time_point t_submit, t_done;
    std::vector<uint64_t> counts, bp;
    std::vector<double> p, llr;
    std::vector<uint8_t> flags;
    bool done = false;
    std::string error;
};

struct Slot {
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;
    std::unique_ptr<Query> q;
    DevBuf<uint64_t> dc, db;
    DevBuf<double> dp, dl;
    DevBuf<uint8_t> df;
    PinnedBuf<uint64_t> hc, hb;
    PinnedBuf<double> hp, hl;
    PinnedBuf<uint8_t> hf;
    std::vector<Request*> reqs;
    uint64_t footprint = 0;
    bool busy = false;
};

struct ServerOpts {
    uint32_t slots = 2;
    uint32_t max_batch_files = 64;
    uint64_t hbm_budget = 0;
    bool pvalues = false;
};

struct Server {
    Index* ix = nullptr;
    ServerOpts opt;
    std::vector<std::unique_ptr<Slot>> slots;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<Request*> pending;
    std::vector<IvBuf> spare;  // interval buffers of completed requests, at most MAX_SPARE
    static constexpr size_t MAX_SPARE = 8;
    std::unordered_map<uint64_t, std::unique_ptr<Request>> reqs;
    uint64_t next_id = 1, inflight_bytes = 0;
    bool stop = false;
    std::string fatal;
    // statistics over completed requests
    std::vector<double> lat_ms;
    uint64_t swept = 0, batches = 0, batched_files = 0;
    bool have_first = false;
    Clock::time_point first_submit, last_done;
    std::thread th;

    // device bytes a request needs while in flight: uploaded intervals and
    // events, the event sort's scratch, its merge-path partition (dense
    // files) and its output row
    uint64_t footprint(uint64_t n) const {
        const uint64_t per_iv = 20 + 2 * (8 + 4 + 4 + 8) + 2 * 12;
        const uint64_t part = (2 * ix->n + 2 * n) / 1024 * 4 + 4096;
        const uint64_t row = (uint64_t)ix->nfiles * (8 + 8 + 8 + 8 + 1);
        return per_iv * n + part + row;
    }

    void start() {
        for (uint32_t i = 0; i < opt.slots; i++) {
            auto s = std::make_unique<Slot>();
            CHK_HIP(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking));
            CHK_HIP(hipEventCreateWithFlags(&s->done, hipEventDisableTiming));
            s->q = std::make_unique<Query>();
            s->q->idx = ix;
            s->q->stream = s->stream;
            slots.push_back(std::move(s));
        }
        th = std::thread([this] { loop(); });
    }

    ~Server() {
        {
            std::lock_guard<std::mutex> g(mu);
            stop = true;
        }
        cv_work.notify_all();
        if (th.joinable()) th.join();
        for (auto& s : slots) {
            if (s->stream) (void)hipStreamSynchronize(s->stream);
            if (s->done) (void)hipEventDestroy(s->done);
            if (s->stream) (void)hipStreamDestroy(s->stream);
        }
    }

    // Launch one batch on a free slot (scheduler thread, lock not held): the
    // requests' intervals go straight into the slot query's pinned staging
    // as its canonical layout (one file per request).
    void launch(Slot& s) {
        const uint32_t B = (uint32_t)s.reqs.size(), S = ix->nshards;
        Query& q = *s.q;
        q.nfiles = B;
        q.file_n.assign(B, 0);
        q.file_len.assign(B, 0.0);
        q.fs_n.assign((size_t)B * S, 0);
        q.file_off.assign(B + 1, 0);
        q.max_end = 0;
        for (uint32_t b = 0; b < B; b++) {
            const Request* r = s.reqs[b];
            q.max_end = std::max(q.max_end, r->max_end);
            q.file_n[b] = r->n;
            q.file_len[b] = r->total_len;
            std::copy(r->fs_n.begin(), r->fs_n.end(), q.fs_n.begin() + (size_t)b * S);
            q.file_off[b + 1] = q.file_off[b] + r->n;
        }
        const uint64_t n = q.file_off[B];
        q.n = q.n_in = n;
        q.st_shard.reserve(n);
        q.st_start.reserve(n);
        q.st_end.reserve(n);
        // staging copy of positions [a, b) of the batch layout; large batches
        // are split over host threads by position
        auto stage = [&](uint64_t a, uint64_t e) {
            uint32_t b = (uint32_t)(std::upper_bound(q.file_off.begin(), q.file_off.end(), a) -
                                    q.file_off.begin()) - 1;
            for (; a < e && b < B; b++) {
                const Request* r = s.reqs[b];
                const uint64_t o = q.file_off[b], k = a - o;
                const uint64_t m = std::min<uint64_t>(q.file_off[b + 1], e) - a;
                if (!m) continue;
                std::memcpy(q.st_shard.p + a, r->iv.shard() + k, m * 4);
                std::memcpy(q.st_start.p + a, r->iv.start() + k, m * 4);
                std::memcpy(q.st_end.p + a, r->iv.end() + k, m * 4);
                a += m;
            }
        };
        const unsigned T = (unsigned)std::max<uint64_t>(1, std::min<uint64_t>(db::host_threads(), n >> 17));
        db::host_run(T, [&](unsigned t) { stage(n * t / T, n * (t + 1) / T); });
        query_stage_commit(q);
        if (opt.pvalues) {
            std::vector<uint32_t> qf(n);
            for (uint32_t b = 0; b < B; b++)
                std::fill(qf.begin() + q.file_off[b], qf.begin() + q.file_off[b + 1], b);
            query_groups(q, q.st_shard.p, q.st_start.p, q.st_end.p, qf.data());
        }
        const size_t cells = (size_t)B * ix->nfiles;
        s.dc.reserve(cells);
        s.db.reserve(cells);
        s.hc.reserve(cells);
        s.hb.reserve(cells);
        sweep_run(q, s.dc.p, s.db.p, s.stream);
        CHK_HIP(hipMemcpyAsync(s.hc.p, s.dc.p, cells * 8, hipMemcpyDeviceToHost, s.stream));
        CHK_HIP(hipMemcpyAsync(s.hb.p, s.db.p, cells * 8, hipMemcpyDeviceToHost, s.stream));
        if (opt.pvalues) {
            s.dp.reserve(cells);
            s.dl.reserve(cells);
            s.df.reserve(cells);
            s.hp.reserve(cells);
            s.hl.reserve(cells);
            s.hf.reserve(cells);
            pvalues_run(q, s.dc.p, s.db.p, s.dp.p, s.dl.p, s.df.p, s.stream);
            CHK_HIP(hipMemcpyAsync(s.hp.p, s.dp.p, cells * 8, hipMemcpyDeviceToHost, s.stream));
            CHK_HIP(hipMemcpyAsync(s.hl.p, s.dl.p, cells * 8, hipMemcpyDeviceToHost, s.stream));
            CHK_HIP(hipMemcpyAsync(s.hf.p, s.df.p, cells, hipMemcpyDeviceToHost, s.stream));
        }
        CHK_HIP(hipEventRecord(s.done, s.stream));
    }

    // Keep a completed request's interval buffer for later requests (lock
    // held): at most MAX_SPARE of them, the larger ones preferred.
    void give_back(IvBuf&& b) {
        if (!b.cap) return;
        if (spare.size() < MAX_SPARE) {
            spare.push_back(std::move(b));
            return;
        }
        auto small = std::min_element(spare.begin(), spare.end(),
                                      [](const IvBuf& x, const IvBuf& y) { return x.cap < y.cap; });
        if (small->cap < b.cap) *small = std::move(b);
    }

    // The smallest spare buffer that holds n intervals, or a new one.
    IvBuf take_buffer(uint64_t n) {
        {
            std::lock_guard<std::mutex> g(mu);
            auto best = spare.end();
            for (auto it = spare.begin(); it != spare.end(); ++it)
                if (it->cap >= n && (best == spare.end() || it->cap < best->cap)) best = it;
            if (best != spare.end()) {
                IvBuf b = std::move(*best);
                spare.erase(best);
                return b;
            }
        }
        IvBuf b;
        b.p.reset(new uint32_t[3 * std::max<uint64_t>(n, 1)]);
        b.cap = std::max<uint64_t>(n, 1);
        return b;
    }

    // Copy each request's row out of a completed slot (lock held).
    void complete(Slot& s, const std::string& err) {
        const uint32_t F = ix->nfiles;
        const auto now = Clock::now();
        for (size_t b = 0; b < s.reqs.size(); b++) {
            Request* r = s.reqs[b];
            if (err.empty()) {
                r->counts.assign(s.hc.p + b * F, s.hc.p + (b + 1) * F);
                r->bp.assign(s.hb.p + b * F, s.hb.p + (b + 1) * F);
                if (opt.pvalues) {
                    r->p.assign(s.hp.p + b * F, s.hp.p + (b + 1) * F);
                    r->llr.assign(s.hl.p + b * F, s.hl.p + (b + 1) * F);
                    r->flags.assign(s.hf.p + b * F, s.hf.p + (b + 1) * F);
                }
                lat_ms.push_back(std::chrono::duration<double, std::milli>(now - r->t_submit).count());
                swept += r->n_given + ix->n;
            } else {
                r->error = err;
            }
            r->t_done = now;
            r->done = true;
            give_back(std::move(r->iv));
        }
        last_done = now;
        inflight_bytes -= s.footprint;
        s.footprint = 0;
        s.reqs.clear();
        s.busy = false;
    }

    void loop() {
        try {
            CHK_HIP(hipSetDevice(ix->device));
        } catch (const Error& e) {
            std::lock_guard<std::mutex> g(mu);
            fatal = e.what();
            cv_done.notify_all();
            return;
        }
        std::unique_lock<std::mutex> lk(mu);
        for (;;) {
            bool any_busy = false;
            for (auto& s : slots) any_busy |= s->busy;
            if (stop && !any_busy) {
                // requests never admitted fail rather than leave their waiters blocked
                const auto now = Clock::now();
                for (Request* r : pending) {
                    r->error = "server stopped before the request was admitted";
                    r->t_done = now;
                    r->done = true;
                }
                pending.clear();
                cv_done.notify_all();
                break;
            }
            // completions
            for (auto& s : slots) {
                if (!s->busy) continue;
                const hipError_t e = hipEventQuery(s->done);
                if (e == hipSuccess) {
                    complete(*s, "");
                } else if (e != hipErrorNotReady) {
                    (void)hipGetLastError();
                    complete(*s, std::string("device error: ") + hipGetErrorString(e));
                }
            }
            cv_done.notify_all();
            // admission: fill every free slot from the head of the queue
            for (auto& s : slots) {
                if (s->busy || pending.empty() || stop) continue;
                uint64_t fp = 0;
                while (!pending.empty() && s->reqs.size() < opt.max_batch_files) {
                    Request* r = pending.front();
                    const bool fits = inflight_bytes + fp + r->footprint <= opt.hbm_budget;
                    // a request larger than the whole budget runs alone
                    const bool alone = inflight_bytes == 0 && s->reqs.empty();
                    if (!fits && !alone) break;
                    fp += r->footprint;
                    s->reqs.push_back(r);
                    pending.pop_front();
                }
                if (s->reqs.empty()) continue;
                s->busy = true;
                s->footprint = fp;
                inflight_bytes += fp;
                batches++;
                batched_files += s->reqs.size();
                lk.unlock();
                std::string err;
                try {
                    launch(*s);
                } catch (const std::exception& e) {
                    err = e.what();
                }
                lk.lock();
                if (!err.empty()) {
                    (void)hipStreamSynchronize(s->stream);
                    complete(*s, err);
                    cv_done.notify_all();
                }
            }
            any_busy = false;
            for (auto& s : slots) any_busy |= s->busy;
            if (any_busy) {
                // poll the device: short sleeps keep the completion latency low
                lk.unlock();
                std::this_thread::sleep_for(std::chrono::microseconds(20));
                lk.lock();
            } else {
                cv_work.wait(lk, [&] { return stop || !pending.empty(); });
            }
        }
    }

    // Validate and keep one query file (client thread): intervals on shards
    // absent from the DB are dropped here, as by query_upload.
    uint64_t submit(uint64_t n, const uint32_t* shard, const uint32_t* start, const uint32_t* end) {
        // the request arrives now: its validation and copy count toward its latency
        const Clock::time_point t_arrive = Clock::now();
        auto r = std::make_unique<Request>();
        const uint32_t S = ix->nshards;
        r->n_given = n;
        // large requests are validated and copied over host threads, each a
        // contiguous range of the input written after the earlier ranges
        const unsigned T = (unsigned)std::max<uint64_t>(1, std::min<uint64_t>(db::host_threads(), n >> 17));
        struct Part {
            std::vector<uint64_t> fs;
            uint64_t kept = 0, len = 0, bad = ~0ull;
            uint32_t max_end = 0;
            int why = 0;
        };
        std::vector<Part> parts(T);
        auto range = [&](unsigned t) { return std::make_pair(n * t / T, n * (t + 1) / T); };
        auto run = [&](auto&& f) { db::host_run(T, f); };
        run([&](unsigned t) {
            Part& P = parts[t];
            P.fs.assign(S, 0);
            const auto [i0, i1] = range(t);
            for (uint64_t i = i0; i < i1; i++) {
                const int why = start[i] >= end[i] ? 1 : (end[i] > 0x7FFFFFFFu ? 2 : 0);
                if (why) {
                    P.bad = i;
                    P.why = why;
                    return;
                }
                if (shard[i] >= S || !ix->shard_n[shard[i]]) continue;
                P.kept++;
                P.fs[shard[i]]++;
                P.max_end = std::max(P.max_end, end[i]);
                P.len += end[i] - start[i];
            }
        });
        for (const Part& P : parts) {  // the first invalid interval in input order
            if (P.bad == ~0ull) continue;
            CHK_REQUIRE(P.why != 1, CHK_ERR_INVALID, "query interval start must be less than end");
            CHK_REQUIRE(false, CHK_ERR_INVALID,
                        "query coordinate exceeds 2^31-1 (device coordinates are int32)");
        }
        r->fs_n.assign(S, 0);
        uint64_t kept = 0, len = 0;
        std::vector<uint64_t> off(T);
        for (unsigned t = 0; t < T; t++) {
            off[t] = kept;
            kept += parts[t].kept;
            len += parts[t].len;
            r->max_end = std::max(r->max_end, parts[t].max_end);
            for (uint32_t c = 0; c < S; c++) r->fs_n[c] += parts[t].fs[c];
        }
        r->total_len = (double)len;
        r->iv = take_buffer(kept);
        r->n = kept;
        uint32_t* const ksh = r->iv.shard();
        uint32_t* const kst = r->iv.start();
        uint32_t* const ken = r->iv.end();
        run([&](unsigned t) {
            uint64_t k = off[t];
            const auto [i0, i1] = range(t);
            for (uint64_t i = i0; i < i1; i++) {
                if (shard[i] >= S || !ix->shard_n[shard[i]]) continue;
                ksh[k] = shard[i];
                kst[k] = start[i];
                ken[k] = end[i];
                k++;
            }
        });
        r->footprint = footprint(kept);
        std::lock_guard<std::mutex> g(mu);
        CHK_REQUIRE(fatal.empty(), CHK_ERR_HIP, "server thread failed: " + fatal);
        r->id = next_id++;
        r->t_submit = t_arrive;
        if (!have_first) {
            have_first = true;
            first_submit = r->t_submit;
        }
        Request* raw = r.get();
        reqs[raw->id] = std::move(r);
        pending.push_back(raw);
        cv_work.notify_one();
        return raw->id;
    }

    std::unique_ptr<Request> wait(uint64_t id) {
        std::unique_lock<std::mutex> lk(mu);
        auto it = reqs.find(id);
        CHK_REQUIRE(it != reqs.end(), CHK_ERR_NOT_FOUND, "unknown or already collected ticket");
        Request* r = it->second.get();
        cv_done.wait(lk, [&] { return r->done || !fatal.empty(); });
        CHK_REQUIRE(r->done, CHK_ERR_HIP, "server thread failed: " + fatal);
        std::unique_ptr<Request> out = std::move(it->second);
        reqs.erase(it);
        return out;
    }
};

Server* server_create(Index& ix, uint32_t slots, uint32_t max_batch_files, uint64_t hbm_budget,
                      bool pvalues) {
    CHK_REQUIRE(slots >= 1 && slots <= 8, CHK_ERR_INVALID, "slots must be 1..8");
    CHK_REQUIRE(max_batch_files >= 1, CHK_ERR_INVALID, "max_batch_files must be positive");
    if (pvalues) {
        CHK_REQUIRE(ix.mom.present, CHK_ERR_NOT_FOUND, "p-values need the index's moment tables");
        CHK_REQUIRE(ix.mom.genome, CHK_ERR_INVALID, "p-values need the genome table");
    }
    auto* s = new Server();
    s->ix = &ix;
    s->opt.slots = slots;
    s->opt.max_batch_files = max_batch_files;
    s->opt.pvalues = pvalues;
    if (!hbm_budget) {
        size_t fr = 0, tot = 0;
        CHK_HIP(hipMemGetInfo(&fr, &tot));
        hbm_budget = fr / 2;  // the other half stays for the slots' reuse and the runtime
    }
    s->opt.hbm_budget = hbm_budget;
    try {
        s->start();
    } catch (...) {
        delete s;
        throw;
    }
    return s;
}

void server_free(Server* s) { delete s; }

uint64_t server_submit(Server& s, uint64_t n, const uint32_t* shard, const uint32_t* start,
                       const uint32_t* end) {
    return s.submit(n, shard, start, end);
}

void server_wait(Server& s, uint64_t ticket, uint64_t* counts, uint64_t* bp, double* p, double* llr,
                 uint8_t* flags, double* latency_ms) {
    std::unique_ptr<Request> r = s.wait(ticket);
    CHK_REQUIRE(r->error.empty(), CHK_ERR_HIP, "request failed: " + r->error);
    const uint32_t F = s.ix->nfiles;
    if (counts) std::copy(r->counts.begin(), r->counts.end(), counts);
    if (bp) std::copy(r->bp.begin(), r->bp.end(), bp);
    if (s.opt.pvalues) {
        if (p) std::copy(r->p.begin(), r->p.end(), p);
        if (llr) std::copy(r->llr.begin(), r->llr.end(), llr);
        if (flags) std::copy(r->flags.begin(), r->flags.end(), flags);
    } else {
        if (p) std::fill(p, p + F, 1.0);
        if (llr) std::fill(llr, llr + F, 0.0);
        if (flags) std::fill(flags, flags + F, (uint8_t)0);
    }
    if (latency_ms)
        *latency_ms = std::chrono::duration<double, std::milli>(r->t_done - r->t_submit).count();
}

void server_stats(Server& s, uint64_t* completed, double* p50, double* p99, double* mean,
                  double* swept_per_s, double* files_per_batch) {
    std::lock_guard<std::mutex> g(s.mu);
    std::vector<double> v = s.lat_ms;
    std::sort(v.begin(), v.end());
    auto pct = [&](double q) {
        if (v.empty()) return 0.0;
        const size_t i = std::min(v.size() - 1, (size_t)std::ceil(q * (double)v.size()) - 1);
        return v[i];
    };
    double sum = 0.0;
    for (double x : v) sum += x;
    if (completed) *completed = v.size();
    if (p50) *p50 = pct(0.50);
    if (p99) *p99 = pct(0.99);
    if (mean) *mean = v.empty() ? 0.0 : sum / (double)v.size();
    if (swept_per_s) {
        const double secs =
            s.have_first ? std::chrono::duration<double>(s.last_done - s.first_submit).count() : 0.0;
        *swept_per_s = secs > 0 ? (double)s.swept / secs : 0.0;
    }
    if (files_per_batch) *files_per_batch = s.batches ? (double)s.batched_files / s.batches : 0.0;
}

void server_reserve(Server& s, uint64_t max_batch_intervals) {
    std::lock_guard<std::mutex> g(s.mu);
    for (auto& sl : s.slots)
        CHK_REQUIRE(!sl->busy, CHK_ERR_INVALID, "reserve needs an idle server");
    CHK_REQUIRE(s.pending.empty(), CHK_ERR_INVALID, "reserve needs an idle server");
    const uint32_t B = s.opt.max_batch_files;
    const size_t cells = (size_t)B * s.ix->nfiles;
    for (auto& sl : s.slots) {
        query_reserve(*sl->q, max_batch_intervals, B);
        sl->dc.reserve(cells);
        sl->db.reserve(cells);
        sl->hc.reserve(cells);
        sl->hb.reserve(cells);
        if (s.opt.pvalues) {
            sl->dp.reserve(cells);
            sl->dl.reserve(cells);
            sl->df.reserve(cells);
            sl->hp.reserve(cells);
            sl->hl.reserve(cells);
            sl->hf.reserve(cells);
            // query length groups: one per distinct 100-bp length per (file,
            // shard), typically far fewer than the intervals
            Query& q = *sl->q;
            const size_t nseg = (size_t)B * s.ix->nshards;
            const size_t ng = std::min<size_t>(max_batch_intervals, nseg * 1024);
            q.g_off.reserve(nseg + 1);
            q.g_lint.reserve(ng);
            q.g_ci.reserve(ng);
            q.g_shard.reserve(ng);
            q.g_ss.reserve(ng);
            q.g_ss2.reserve(ng);
            q.g_cnt.reserve(ng);
        }
    }
}

void server_reset_stats(Server& s) {
    std::lock_guard<std::mutex> g(s.mu);
    s.lat_ms.clear();
    s.swept = 0;
    s.batches = 0;
    s.batched_files = 0;
    s.have_first = false;
}

}  // namespace chk
