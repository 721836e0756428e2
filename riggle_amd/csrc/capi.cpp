// capi.cpp -- extern "C" entry points declared in include/chuckle_hip.h.
// Each catches every exception, records the message for chk_last_error() and
// returns a status code.
#include <cstring>
#include <new>
#include <set>
#include <unordered_map>

#include "dbfile.hpp"
#include "engine.hpp"

struct chk_index {
    chk::Index ix;
};
struct chk_query {
    chk::Query q;
};
struct chk_server {
    chk_index* idx = nullptr;
    chk::Server* s = nullptr;
};
struct chk_graph {
    chk_index* idx = nullptr;
    chk_query* q = nullptr;
    uint32_t plan_version = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    ~chk_graph() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
    }
};

namespace chk {
static thread_local std::string g_last_error;
void set_last_error(const std::string& msg) { g_last_error = msg; }
}  // namespace chk

namespace {

template <typename F>
int guarded(F&& f) {
    try {
        f();
        return CHK_OK;
    } catch (const chk::Error& e) {
        chk::set_last_error(e.what());
        return e.code;
    } catch (const std::bad_alloc&) {
        chk::set_last_error("host allocation failed");
        return CHK_ERR_NOMEM;
    } catch (const std::exception& e) {
        chk::set_last_error(e.what());
        return CHK_ERR_INTERNAL;
    } catch (...) {
        chk::set_last_error("unknown error");
        return CHK_ERR_INTERNAL;
    }
}

void set_device(int device) { CHK_HIP(hipSetDevice(device)); }

}  // namespace

extern "C" {

const char* chk_version(void) { return "chuckle-mi355x 0.1.0 (gfx950)"; }

const char* chk_last_error(void) { return chk::g_last_error.c_str(); }

int chk_device_count(int* count) {
    return guarded([&] {
        CHK_REQUIRE(count, CHK_ERR_INVALID, "count is NULL");
        int n = 0;
        hipError_t e = hipGetDeviceCount(&n);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            n = 0;
        }
        *count = n;
    });
}

int chk_index_create(int device, uint32_t nshards, uint32_t nfiles, uint64_t n,
                     const uint32_t* shard, const uint32_t* start, const uint32_t* end,
                     const uint32_t* fid, chk_index** out) {
    return guarded([&] {
        CHK_REQUIRE(out, CHK_ERR_INVALID, "out is NULL");
        *out = nullptr;
        CHK_REQUIRE(n == 0 || (shard && start && end && fid), CHK_ERR_INVALID,
                    "interval arrays must be non-NULL");
        CHK_REQUIRE(nfiles > 0, CHK_ERR_INVALID, "nfiles must be positive");
        set_device(device);
        auto* h = new chk_index();
        try {
            chk::index_init(h->ix, device, nshards, nfiles, n);
            chk::index_build(h->ix, shard, start, end, fid);
        } catch (...) {
            delete h;
            throw;
        }
        *out = h;
    });
}

int chk_index_open(const char* db_path, int device, chk_index** out) {
    return guarded([&] {
        CHK_REQUIRE(db_path && out, CHK_ERR_INVALID, "NULL argument");
        *out = nullptr;
        const std::string db(db_path);
        const chk::db::Meta meta = chk::db::Meta::load(db);
        const chk::db::Flat f = chk::db::load_intervals(db, meta);
        for (uint32_t v : f.sid)
            CHK_REQUIRE(meta.sid_map.count(v), CHK_ERR_IO, "layer file holds a D_SID missing from meta.json");
        const uint32_t S = (uint32_t)meta.shards.size(), F = std::max<uint32_t>(1, meta.next_sid());
        set_device(device);
        auto* h = new chk_index();
        try {
            chk::index_init(h->ix, device, S, F, f.start.size());
            chk::index_build(h->ix, f.shard.data(), f.start.data(), f.end.data(), f.sid.data());
            h->ix.shard_names = meta.shards;
            h->ix.file_names.assign(F, std::string());
            for (auto& kv : meta.sid_map)
                if (kv.first < F) h->ix.file_names[kv.first] = kv.second;
            const auto g = chk::db::genome_for(meta.shards, nullptr);
            chk::set_genome(h->ix, g.data());
            chk::db::load_moments(db, h->ix, meta);
        } catch (...) {
            delete h;
            throw;
        }
        *out = h;
    });
}

static int copy_name(const std::vector<std::string>& names, uint32_t id, char* buf, uint64_t len,
                     uint64_t* needed, const char* what) {
    return guarded([&] {
        CHK_REQUIRE(!names.empty(), CHK_ERR_NOT_FOUND,
                    std::string("the index has no ") + what + " names (it was not opened from a database)");
        CHK_REQUIRE(id < names.size(), CHK_ERR_INVALID, std::string(what) + " id out of range");
        const std::string& n = names[id];
        if (needed) *needed = n.size() + 1;
        if (buf && len) {
            const size_t k = std::min<size_t>(n.size(), (size_t)len - 1);
            std::memcpy(buf, n.data(), k);
            buf[k] = 0;
        }
    });
}

int chk_index_shard_name(const chk_index* idx, uint32_t shard, char* buf, uint64_t len,
                         uint64_t* needed) {
    if (!idx) return guarded([] { throw chk::Error(CHK_ERR_INVALID, "NULL index"); });
    return copy_name(idx->ix.shard_names, shard, buf, len, needed, "shard");
}

int chk_index_file_name(const chk_index* idx, uint32_t fid, char* buf, uint64_t len,
                        uint64_t* needed) {
    if (!idx) return guarded([] { throw chk::Error(CHK_ERR_INVALID, "NULL index"); });
    return copy_name(idx->ix.file_names, fid, buf, len, needed, "file");
}

int chk_index_set_names(chk_index* idx, const char* const* shard_names, const char* const* file_names) {
    return guarded([&] {
        CHK_REQUIRE(idx && (idx->ix.nshards == 0 || shard_names) && file_names, CHK_ERR_INVALID,
                    "NULL argument");
        std::vector<std::string> sn, fn;
        for (uint32_t s = 0; s < idx->ix.nshards; s++) {
            CHK_REQUIRE(shard_names[s], CHK_ERR_INVALID, "NULL shard name");
            sn.emplace_back(shard_names[s]);
        }
        for (uint32_t f = 0; f < idx->ix.nfiles; f++) {
            CHK_REQUIRE(file_names[f], CHK_ERR_INVALID, "NULL file name");
            fn.emplace_back(file_names[f]);
        }
        idx->ix.shard_names = std::move(sn);
        idx->ix.file_names = std::move(fn);
    });
}

int chk_db_add(const char* input, const char* db_path, uint64_t batch_size, int stats,
               const char* const* chroms, uint32_t nchroms, int sorted_sids, int device) {
    return guarded([&] {
        CHK_REQUIRE(input && db_path && (nchroms == 0 || chroms), CHK_ERR_INVALID, "NULL argument");
        std::set<std::string> wl;
        for (uint32_t k = 0; k < nchroms; k++) wl.insert(chroms[k]);
        chk::db::add(input, db_path, batch_size, stats != 0, wl, sorted_sids != 0, device);
    });
}

int chk_write_results(const char* path, uint32_t nquery, const char* const* query_names,
                      uint32_t nfiles, const char* const* db_names, const uint64_t* counts,
                      const uint64_t* bp, const double* p, const double* llr, const uint8_t* flags,
                      int stats) {
    return guarded([&] {
        CHK_REQUIRE(path && (nquery == 0 || query_names) && (nfiles == 0 || db_names), CHK_ERR_INVALID,
                    "NULL argument");
        const size_t cells = (size_t)nquery * nfiles;
        CHK_REQUIRE(cells == 0 || (counts && (!stats || (bp && p && llr && flags))), CHK_ERR_INVALID,
                    "NULL result array");
        std::vector<std::string> q(query_names, query_names + nquery), d(db_names, db_names + nfiles);
        chk::db::write_results(path, q, d, counts, bp, p, llr, flags, stats != 0);
    });
}

int chk_index_free(chk_index* idx) {
    return guarded([&] {
        if (!idx) return;
        set_device(idx->ix.device);
        (void)hipStreamSynchronize(idx->ix.stream);
        delete idx;
    });
}

int chk_index_info(const chk_index* idx, uint32_t* nshards, uint32_t* nfiles, uint64_t* n,
                   uint64_t* device_bytes) {
    return guarded([&] {
        CHK_REQUIRE(idx, CHK_ERR_INVALID, "index is NULL");
        if (nshards) *nshards = idx->ix.nshards;
        if (nfiles) *nfiles = idx->ix.nfiles;
        if (n) *n = idx->ix.n;
        if (device_bytes) *device_bytes = idx->ix.device_bytes();
    });
}

int chk_query_create(chk_index* idx, uint32_t nfiles, uint64_t n, const uint32_t* shard,
                     const uint32_t* start, const uint32_t* end, const uint32_t* qfile,
                     chk_query** out) {
    return guarded([&] {
        CHK_REQUIRE(idx && out, CHK_ERR_INVALID, "index / out is NULL");
        *out = nullptr;
        CHK_REQUIRE(nfiles > 0, CHK_ERR_INVALID, "nfiles must be positive");
        CHK_REQUIRE(n == 0 || (shard && start && end), CHK_ERR_INVALID,
                    "interval arrays must be non-NULL");
        CHK_REQUIRE(nfiles == 1 || qfile || n == 0, CHK_ERR_INVALID,
                    "qfile is required when nfiles > 1");
        set_device(idx->ix.device);
        auto* h = new chk_query();
        try {
            h->q.idx = &idx->ix;
            h->q.nfiles = nfiles;
            h->q.n_in = n;
            h->q.stream = idx->ix.stream;
            chk::query_upload(h->q, shard, start, end, qfile);
            chk::query_groups(h->q, shard, start, end, qfile);
        } catch (...) {
            delete h;
            throw;
        }
        *out = h;
    });
}

int chk_query_reset(chk_query* q, uint32_t nfiles, uint64_t n, const uint32_t* shard,
                    const uint32_t* start, const uint32_t* end, const uint32_t* qfile) {
    return guarded([&] {
        CHK_REQUIRE(q, CHK_ERR_INVALID, "NULL query");
        CHK_REQUIRE(nfiles > 0, CHK_ERR_INVALID, "nfiles must be positive");
        CHK_REQUIRE(n == 0 || (shard && start && end), CHK_ERR_INVALID,
                    "interval arrays must be non-NULL");
        CHK_REQUIRE(nfiles == 1 || qfile || n == 0, CHK_ERR_INVALID,
                    "qfile is required when nfiles > 1");
        set_device(q->q.idx->device);
        // the previous batch's copies read the staging buffers: let them finish
        CHK_HIP(hipStreamSynchronize(q->q.stream));
        q->q.nfiles = nfiles;
        q->q.n_in = n;
        chk::query_upload(q->q, shard, start, end, qfile);
        chk::query_groups(q->q, shard, start, end, qfile);
    });
}

int chk_query_open(chk_index* idx, uint32_t npaths, const char* const* paths, uint32_t* row_of_path,
                   chk_query** out) {
    return guarded([&] {
        CHK_REQUIRE(idx && out && (npaths == 0 || paths), CHK_ERR_INVALID, "NULL argument");
        *out = nullptr;
        chk::Index& ix = idx->ix;
        CHK_REQUIRE(!ix.shard_names.empty() || ix.nshards == 0, CHK_ERR_INVALID,
                    "the index has no shard names: open it with chk_index_open");
        std::vector<std::string> ps(paths, paths + npaths);
        const auto files = chk::db::parse_beds(ps, chk::db::host_threads());
        std::unordered_map<std::string, uint32_t> id;
        for (uint32_t s = 0; s < ix.shard_names.size(); s++) id[ix.shard_names[s]] = s;
        // empty files take no row; rows follow the path order (files.rs:79-101)
        uint32_t rows = 0;
        uint64_t n = 0;
        for (uint32_t k = 0; k < npaths; k++) {
            if (row_of_path) row_of_path[k] = files[k].n ? rows : 0xFFFFFFFFu;
            rows += files[k].n ? 1 : 0;
            n += files[k].n;
        }
        if (rows == 0) return;  // every file empty: no query (*out stays NULL)
        std::vector<uint32_t> sh(n), st(n), en(n), qf(n);
        uint64_t i = 0;
        uint32_t r = 0;
        for (auto& f : files) {
            if (!f.n) continue;
            for (size_t s = 0; s < f.shards.size(); s++) {
                auto it = id.find(f.shards[s]);
                const uint32_t sid = it == id.end() ? ix.nshards : it->second;  // absent: dropped
                for (auto& iv : f.ivs[s]) {
                    sh[i] = sid;
                    st[i] = iv.start;
                    en[i] = iv.end;
                    qf[i++] = r;
                }
            }
            r++;
        }
        set_device(ix.device);
        auto* h = new chk_query();
        try {
            h->q.idx = &ix;
            h->q.nfiles = rows;
            h->q.n_in = n;
            h->q.stream = ix.stream;
            chk::query_upload(h->q, sh.data(), st.data(), en.data(), qf.data());
            chk::query_groups(h->q, sh.data(), st.data(), en.data(), qf.data());
        } catch (...) {
            delete h;
            throw;
        }
        *out = h;
    });
}

int chk_query_info(chk_query* q, uint32_t* nfiles, uint64_t* n, uint32_t* dense_files,
                   uint64_t* dense_intervals) {
    return guarded([&] {
        CHK_REQUIRE(q, CHK_ERR_INVALID, "NULL query");
        set_device(q->q.idx->device);
        chk::query_plan(q->q);
        if (nfiles) *nfiles = q->q.nfiles;
        if (n) *n = q->q.n;
        if (dense_files) *dense_files = (uint32_t)q->q.dense_rows.size();
        if (dense_intervals) *dense_intervals = q->q.n_dense;
    });
}

int chk_query_free(chk_query* q) {
    return guarded([&] {
        if (!q) return;
        set_device(q->q.idx->device);
        (void)hipStreamSynchronize(q->q.idx->stream);
        delete q;
    });
}

int chk_sweep_async(chk_index* idx, chk_query* q, uint64_t* d_counts, uint64_t* d_bp) {
    return guarded([&] {
        CHK_REQUIRE(idx && q && d_counts && d_bp, CHK_ERR_INVALID, "NULL argument");
        CHK_REQUIRE(q->q.idx == &idx->ix, CHK_ERR_INVALID, "query belongs to another index");
        set_device(idx->ix.device);
        chk::query_plan(q->q);
        chk::sweep_run(q->q, d_counts, d_bp, idx->ix.stream);
    });
}

int chk_sweep(chk_index* idx, chk_query* q, uint64_t* counts, uint64_t* bp) {
    return guarded([&] {
        CHK_REQUIRE(idx && q && counts && bp, CHK_ERR_INVALID, "NULL argument");
        CHK_REQUIRE(q->q.idx == &idx->ix, CHK_ERR_INVALID, "query belongs to another index");
        set_device(idx->ix.device);
        size_t cells = (size_t)q->q.nfiles * idx->ix.nfiles;
        chk::DevBuf<uint64_t> dc, db;
        dc.reserve(cells);
        db.reserve(cells);
        chk::query_plan(q->q);
        chk::sweep_run(q->q, dc.p, db.p, idx->ix.stream);
        CHK_HIP(hipMemcpyAsync(counts, dc.p, cells * 8, hipMemcpyDeviceToHost, idx->ix.stream));
        CHK_HIP(hipMemcpyAsync(bp, db.p, cells * 8, hipMemcpyDeviceToHost, idx->ix.stream));
        CHK_HIP(hipStreamSynchronize(idx->ix.stream));
    });
}

int chk_index_set_genome(chk_index* idx, const uint64_t* shard_size) {
    return guarded([&] {
        CHK_REQUIRE(idx && shard_size, CHK_ERR_INVALID, "NULL argument");
        set_device(idx->ix.device);
        chk::set_genome(idx->ix, shard_size);
    });
}

int chk_moments_upload(chk_index* idx, const int64_t* slot_off, const uint32_t* slot_len,
                       uint64_t total, const double* mean, const double* var) {
    return guarded([&] {
        CHK_REQUIRE(idx && slot_off && slot_len && (total == 0 || (mean && var)), CHK_ERR_INVALID,
                    "NULL argument");
        set_device(idx->ix.device);
        chk::moments_upload(idx->ix, slot_off, slot_len, total, mean, var);
    });
}

int chk_moments_build(chk_index* idx, int32_t* rows, int32_t* exact_rows) {
    return guarded([&] {
        CHK_REQUIRE(idx, CHK_ERR_INVALID, "NULL index");
        set_device(idx->ix.device);
        chk::moments_build(idx->ix);
        if (rows) *rows = idx->ix.mom.stat_rows;
        if (exact_rows) *exact_rows = idx->ix.mom.stat_exact_rows;
    });
}

int chk_moments_widths(chk_index* idx, uint32_t nwidths, const uint32_t* widths, double* mean,
                       double* var, uint32_t* inexact_rows, uint64_t* bins, double* lags,
                       uint64_t* exact_widths) {
    return guarded([&] {
        CHK_REQUIRE(idx, CHK_ERR_INVALID, "NULL index");
        set_device(idx->ix.device);
        chk::moments_widths(idx->ix, widths, nwidths, mean, var, inexact_rows, bins, lags, exact_widths);
    });
}

int chk_moments_download(chk_index* idx, uint64_t* total, int64_t* slot_off, uint32_t* slot_len,
                         double* mean, double* var) {
    return guarded([&] {
        CHK_REQUIRE(idx, CHK_ERR_INVALID, "NULL index");
        set_device(idx->ix.device);
        CHK_REQUIRE(idx->ix.mom.present, CHK_ERR_NOT_FOUND, "the index has no moment tables");
        if (total) *total = idx->ix.mom.total;
        chk::moments_download(idx->ix, slot_off, slot_len, mean, var);
    });
}

int chk_pvalues_async(chk_index* idx, chk_query* q, const uint64_t* d_counts, const uint64_t* d_bp,
                      double* d_p, double* d_llr, uint8_t* d_flags) {
    return guarded([&] {
        CHK_REQUIRE(idx && q && d_counts && d_bp && d_p && d_llr && d_flags, CHK_ERR_INVALID,
                    "NULL argument");
        CHK_REQUIRE(q->q.idx == &idx->ix, CHK_ERR_INVALID, "query belongs to another index");
        set_device(idx->ix.device);
        chk::pvalues_run(q->q, d_counts, d_bp, d_p, d_llr, d_flags, idx->ix.stream);
    });
}

int chk_query_stats(chk_index* idx, chk_query* q, uint64_t* counts, uint64_t* bp, double* p,
                    double* llr, uint8_t* flags) {
    return guarded([&] {
        CHK_REQUIRE(idx && q && counts && bp && p && llr && flags, CHK_ERR_INVALID, "NULL argument");
        CHK_REQUIRE(q->q.idx == &idx->ix, CHK_ERR_INVALID, "query belongs to another index");
        set_device(idx->ix.device);
        hipStream_t st = idx->ix.stream;
        size_t cells = (size_t)q->q.nfiles * idx->ix.nfiles;
        chk::DevBuf<uint64_t> dc, db;
        chk::DevBuf<double> dp, dl;
        chk::DevBuf<uint8_t> df;
        dc.reserve(cells);
        db.reserve(cells);
        dp.reserve(cells);
        dl.reserve(cells);
        df.reserve(cells);
        chk::query_plan(q->q);
        chk::sweep_run(q->q, dc.p, db.p, idx->ix.stream);
        chk::pvalues_run(q->q, dc.p, db.p, dp.p, dl.p, df.p, st);
        CHK_HIP(hipMemcpyAsync(counts, dc.p, cells * 8, hipMemcpyDeviceToHost, st));
        CHK_HIP(hipMemcpyAsync(bp, db.p, cells * 8, hipMemcpyDeviceToHost, st));
        CHK_HIP(hipMemcpyAsync(p, dp.p, cells * 8, hipMemcpyDeviceToHost, st));
        CHK_HIP(hipMemcpyAsync(llr, dl.p, cells * 8, hipMemcpyDeviceToHost, st));
        CHK_HIP(hipMemcpyAsync(flags, df.p, cells, hipMemcpyDeviceToHost, st));
        CHK_HIP(hipStreamSynchronize(st));
    });
}

int chk_dev_alloc(chk_index* idx, uint64_t bytes, void** ptr) {
    return guarded([&] {
        CHK_REQUIRE(idx && ptr, CHK_ERR_INVALID, "NULL argument");
        set_device(idx->ix.device);
        void* p = nullptr;
        hipError_t e = hipMalloc(&p, bytes ? bytes : 1);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            throw chk::Error(CHK_ERR_NOMEM, std::string("hipMalloc failed: ") + hipGetErrorString(e));
        }
        *ptr = p;
    });
}

int chk_dev_free(chk_index* idx, void* ptr) {
    return guarded([&] {
        CHK_REQUIRE(idx, CHK_ERR_INVALID, "NULL index");
        set_device(idx->ix.device);
        CHK_HIP(hipStreamSynchronize(idx->ix.stream));
        if (ptr) CHK_HIP(hipFree(ptr));
    });
}

int chk_dev_to_host(chk_index* idx, void* dst, const void* src, uint64_t bytes) {
    return guarded([&] {
        CHK_REQUIRE(idx && dst && src, CHK_ERR_INVALID, "NULL argument");
        set_device(idx->ix.device);
        CHK_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, idx->ix.stream));
        CHK_HIP(hipStreamSynchronize(idx->ix.stream));
    });
}

int chk_sync(chk_index* idx) {
    return guarded([&] {
        CHK_REQUIRE(idx, CHK_ERR_INVALID, "NULL index");
        set_device(idx->ix.device);
        CHK_HIP(hipStreamSynchronize(idx->ix.stream));
    });
}

int chk_profile_enable(chk_index* idx, int on) {
    return guarded([&] {
        CHK_REQUIRE(idx, CHK_ERR_INVALID, "NULL index");
        idx->ix.prof.on = on != 0;
    });
}

int chk_profile_only(chk_index* idx, const char* regions) {
    return guarded([&] {
        CHK_REQUIRE(idx, CHK_ERR_INVALID, "NULL index");
        std::set<std::string> only;
        if (regions) {
            std::string cur;
            for (const char* c = regions;; c++) {
                if (*c == ',' || *c == 0) {
                    if (!cur.empty()) only.insert(cur);
                    cur.clear();
                    if (!*c) break;
                } else {
                    cur.push_back(*c);
                }
            }
        }
        idx->ix.prof.only = only;
    });
}

int chk_profile_read(chk_index* idx, const char* name, uint64_t* launches, double* total_ms) {
    return guarded([&] {
        CHK_REQUIRE(idx && name, CHK_ERR_INVALID, "NULL argument");
        set_device(idx->ix.device);
        CHK_HIP(hipStreamSynchronize(idx->ix.stream));
        uint64_t cnt = 0;
        double ms = 0.0;
        auto it = idx->ix.prof.ev.find(name);
        if (it != idx->ix.prof.ev.end()) {
            for (auto& p : it->second) {
                float t = 0.f;
                CHK_HIP(hipEventElapsedTime(&t, p.first, p.second));
                ms += t;
                cnt++;
                idx->ix.prof.pool.push_back(p.first);
                idx->ix.prof.pool.push_back(p.second);
            }
            it->second.clear();
        }
        if (launches) *launches = cnt;
        if (total_ms) *total_ms = ms;
    });
}

int chk_tune(chk_index* idx, const char* key, int64_t value) {
    return guarded([&] {
        CHK_REQUIRE(idx && key, CHK_ERR_INVALID, "NULL argument");
        CHK_REQUIRE(value >= 0 && value <= 0xFFFFFFFFll, CHK_ERR_INVALID, "tuning value out of range");
        std::string k(key);
        chk::Tuning& t = idx->ix.tune;
        const uint32_t v = (uint32_t)value;
        if (k == "sweep.blocks_per_cu") t.blocks_per_cu = v;
        else if (k == "sweep.chunk") t.chunk = v;
        else if (k == "sweep.items") {
            CHK_REQUIRE(v == 0 || v == 4 || v == 8 || v == 12 || v == 16, CHK_ERR_INVALID,
                        "sweep.items must be 0, 4, 8, 12 or 16");
            t.items = v;
        } else if (k == "sweep.sparse_max") t.sparse_max = (uint64_t)value;
        else if (k == "sweep.method") {
            CHK_REQUIRE(v <= 2, CHK_ERR_INVALID, "sweep.method must be 0, 1 or 2");
            t.method = v;
        } else if (k == "sweep.sparse_items") t.sparse_items = v;
        else if (k == "moments.fft_mib") t.fft_mib = v;
        else if (k == "moments.fft_block") {
            CHK_REQUIRE(v == 0 || v == 4096 || v == 8192, CHK_ERR_INVALID,
                        "moments.fft_block must be 0, 4096 or 8192");
            t.fft_block = v;
        } else if (k == "moments.round") {
            CHK_REQUIRE(v <= 2, CHK_ERR_INVALID, "moments.round must be 0, 1 or 2");
            t.mom_round = v;
        } else if (k == "moments.eval_seg") {
            CHK_REQUIRE(v % 4096 == 0, CHK_ERR_INVALID, "moments.eval_seg must be a multiple of 4096");
            t.eval_seg = v;
        } else if (k == "moments.fft_pipe") {
            CHK_REQUIRE(v <= 1, CHK_ERR_INVALID, "moments.fft_pipe must be 0 or 1");
            t.fft_pipe = v;
        } else throw chk::Error(CHK_ERR_INVALID, "unknown tuning key: " + k);
        idx->ix.tune_version++;
    });
}

int chk_tune_get(chk_index* idx, const char* key, int64_t* value) {
    return guarded([&] {
        CHK_REQUIRE(idx && key && value, CHK_ERR_INVALID, "NULL argument");
        std::string k(key);
        const chk::Tuning& t = idx->ix.tune;
        if (k == "sweep.blocks_per_cu") *value = t.blocks_per_cu;
        else if (k == "sweep.chunk") *value = t.chunk;
        else if (k == "sweep.items") *value = t.items;
        else if (k == "sweep.sparse_max") *value = (int64_t)t.sparse_max;
        else if (k == "sweep.method") *value = t.method;
        else if (k == "sweep.sparse_items") *value = t.sparse_items;
        else if (k == "moments.fft_mib") *value = t.fft_mib;
        else if (k == "moments.fft_block") *value = t.fft_block;
        else if (k == "moments.round") *value = t.mom_round;
        else if (k == "moments.eval_seg") *value = t.eval_seg;
        else if (k == "moments.fft_pipe") *value = t.fft_pipe;
        else throw chk::Error(CHK_ERR_INVALID, "unknown tuning key: " + k);
    });
}

int chk_autotune(chk_index* idx, chk_query* q, uint64_t* d_counts, uint64_t* d_bp, int reps,
                 double* best_ms, char* report, uint64_t report_len) {
    return guarded([&] {
        CHK_REQUIRE(idx && q && d_counts && d_bp, CHK_ERR_INVALID, "NULL argument");
        CHK_REQUIRE(q->q.idx == &idx->ix, CHK_ERR_INVALID, "query belongs to another index");
        set_device(idx->ix.device);
        std::string rep;
        const double ms = chk::autotune(q->q, d_counts, d_bp, reps, &rep);
        if (best_ms) *best_ms = ms;
        if (report && report_len) {
            const size_t n = std::min<size_t>(rep.size(), (size_t)report_len - 1);
            std::memcpy(report, rep.data(), n);
            report[n] = 0;
        }
    });
}

int chk_graph_create(chk_index* idx, chk_query* q, uint64_t* d_counts, uint64_t* d_bp,
                     double* d_p, double* d_llr, uint8_t* d_flags, chk_graph** out) {
    return guarded([&] {
        CHK_REQUIRE(idx && q && d_counts && d_bp && out, CHK_ERR_INVALID, "NULL argument");
        CHK_REQUIRE(q->q.idx == &idx->ix, CHK_ERR_INVALID, "query belongs to another index");
        CHK_REQUIRE((d_p && d_llr && d_flags) || (!d_p && !d_llr && !d_flags), CHK_ERR_INVALID,
                    "p-value outputs must be all given or all NULL");
        *out = nullptr;
        set_device(idx->ix.device);
        chk::Index& ix = idx->ix;
        hipStream_t st = ix.stream;
        // one eager pass sizes every work buffer, so the capture allocates nothing
        chk::query_plan(q->q);
        chk::sweep_run(q->q, d_counts, d_bp, st);
        if (d_p) chk::pvalues_run(q->q, d_counts, d_bp, d_p, d_llr, d_flags, st);
        CHK_HIP(hipStreamSynchronize(st));
        auto* g = new chk_graph();
        const bool prof = ix.prof.on;
        ix.prof.on = false;
        try {
            g->idx = idx;
            g->q = q;
            g->plan_version = q->q.plan_version;
            CHK_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
            try {
                chk::sweep_run(q->q, d_counts, d_bp, st);
                if (d_p) chk::pvalues_run(q->q, d_counts, d_bp, d_p, d_llr, d_flags, st);
            } catch (...) {
                hipGraph_t dead = nullptr;
                (void)hipStreamEndCapture(st, &dead);
                if (dead) (void)hipGraphDestroy(dead);
                throw;
            }
            CHK_HIP(hipStreamEndCapture(st, &g->graph));
            CHK_HIP(hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0));
        } catch (...) {
            ix.prof.on = prof;
            delete g;
            throw;
        }
        ix.prof.on = prof;
        *out = g;
    });
}

int chk_graph_launch(chk_graph* g) {
    return guarded([&] {
        CHK_REQUIRE(g && g->exec, CHK_ERR_INVALID, "graph is NULL");
        CHK_REQUIRE(g->q->q.plan_version == g->idx->ix.tune_version &&
                        g->plan_version == g->idx->ix.tune_version,
                    CHK_ERR_INVALID, "the index's tuning changed after capture: recreate the graph");
        set_device(g->idx->ix.device);
        CHK_HIP(hipGraphLaunch(g->exec, g->idx->ix.stream));
    });
}

int chk_graph_free(chk_graph* g) {
    return guarded([&] {
        if (!g) return;
        set_device(g->idx->ix.device);
        (void)hipStreamSynchronize(g->idx->ix.stream);
        delete g;
    });
}

int chk_server_create(chk_index* idx, uint32_t slots, uint32_t max_batch_files,
                      uint64_t hbm_budget, int pvalues, chk_server** out) {
    return guarded([&] {
        CHK_REQUIRE(idx && out, CHK_ERR_INVALID, "NULL argument");
        *out = nullptr;
        set_device(idx->ix.device);
        auto* h = new chk_server();
        try {
            h->idx = idx;
            h->s = chk::server_create(idx->ix, slots, max_batch_files, hbm_budget, pvalues != 0);
        } catch (...) {
            delete h;
            throw;
        }
        *out = h;
    });
}

int chk_server_submit(chk_server* s, uint64_t n, const uint32_t* shard, const uint32_t* start,
                      const uint32_t* end, uint64_t* ticket) {
    return guarded([&] {
        CHK_REQUIRE(s && s->s && ticket, CHK_ERR_INVALID, "NULL argument");
        CHK_REQUIRE(n == 0 || (shard && start && end), CHK_ERR_INVALID, "interval arrays must be non-NULL");
        *ticket = chk::server_submit(*s->s, n, shard, start, end);
    });
}

int chk_server_wait(chk_server* s, uint64_t ticket, uint64_t* counts, uint64_t* bp, double* p,
                    double* llr, uint8_t* flags, double* latency_ms) {
    return guarded([&] {
        CHK_REQUIRE(s && s->s, CHK_ERR_INVALID, "NULL server");
        chk::server_wait(*s->s, ticket, counts, bp, p, llr, flags, latency_ms);
    });
}

int chk_server_stats(chk_server* s, uint64_t* completed, double* p50_ms, double* p99_ms,
                     double* mean_ms, double* swept_per_s, double* files_per_batch) {
    return guarded([&] {
        CHK_REQUIRE(s && s->s, CHK_ERR_INVALID, "NULL server");
        chk::server_stats(*s->s, completed, p50_ms, p99_ms, mean_ms, swept_per_s, files_per_batch);
    });
}

int chk_server_reserve(chk_server* s, uint64_t max_batch_intervals) {
    return guarded([&] {
        CHK_REQUIRE(s && s->s, CHK_ERR_INVALID, "NULL server");
        set_device(s->idx->ix.device);
        chk::server_reserve(*s->s, max_batch_intervals);
    });
}

int chk_server_reset_stats(chk_server* s) {
    return guarded([&] {
        CHK_REQUIRE(s && s->s, CHK_ERR_INVALID, "NULL server");
        chk::server_reset_stats(*s->s);
    });
}

int chk_server_free(chk_server* s) {
    return guarded([&] {
        if (!s) return;
        set_device(s->idx->ix.device);
        chk::server_free(s->s);
        delete s;
    });
}

int chk_sweep_noalloc(chk_index* idx, uint64_t n, const uint32_t* shard, const uint32_t* start,
                      const uint32_t* end, uint64_t* d_counts, uint64_t* d_bp) {
    return guarded([&] {
        CHK_REQUIRE(idx && d_counts && d_bp, CHK_ERR_INVALID, "NULL argument");
        CHK_REQUIRE(n == 0 || (shard && start && end), CHK_ERR_INVALID,
                    "interval arrays must be non-NULL");
        set_device(idx->ix.device);
        chk::sweep_noalloc(idx->ix, n, shard, start, end, d_counts, d_bp);
    });
}

int chk_calibrate_read(chk_index* idx, uint32_t reps, double* best_ms, uint64_t* bytes) {
    return guarded([&] {
        CHK_REQUIRE(idx, CHK_ERR_INVALID, "NULL index");
        set_device(idx->ix.device);
        chk::calibrate_read(idx->ix, reps, best_ms, bytes);
    });
}

int chk_calibrate_copy(chk_index* idx, uint64_t bytes) {
    return guarded([&] {
        CHK_REQUIRE(idx, CHK_ERR_INVALID, "NULL index");
        set_device(idx->ix.device);
        chk::calibrate_copy(idx->ix, bytes);
    });
}

}  // extern "C"
