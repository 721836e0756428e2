// engine.hpp -- the HBM-resident index, prepared query batches and the
// merge-path overlap sweep.
#pragma once

#include <map>
#include <string>
#include <utility>

#include "common.hpp"
#include "sort.hpp"

namespace chk {

// HIP-event timing of named kernels on one stream.
struct Profiler {
    bool on = false;
    std::map<std::string, std::vector<std::pair<hipEvent_t, hipEvent_t>>> ev;
    std::vector<hipEvent_t> pool;
    hipEvent_t get() {
        if (!pool.empty()) {
            hipEvent_t e = pool.back();
            pool.pop_back();
            return e;
        }
        hipEvent_t e;
        CHK_HIP(hipEventCreate(&e));
        return e;
    }
    // Returns the start event (or nullptr when profiling is off).
    hipEvent_t begin(hipStream_t st) {
        if (!on) return nullptr;
        hipEvent_t a = get();
        CHK_HIP(hipEventRecord(a, st));
        return a;
    }
    void end(const char* name, hipEvent_t a, hipStream_t st) {
        if (!a) return;
        hipEvent_t b = get();
        CHK_HIP(hipEventRecord(b, st));
        ev[name].push_back({a, b});
    }
    ~Profiler() {
        for (auto& kv : ev)
            for (auto& p : kv.second) {
                (void)hipEventDestroy(p.first);
                (void)hipEventDestroy(p.second);
            }
        for (auto e : pool) (void)hipEventDestroy(e);
    }
};

// Launch-shape knobs of the sweep (0 = choose automatically).
struct Tuning {
    uint32_t blocks_per_cu = 0;    // resident blocks per CU the grid is sized for
    uint32_t tiles_per_block = 0;  // contiguous tiles per block
    uint32_t part_window = 0;      // partition entries per LDS window (<= 1024)
};

// Compact sliding-window moment tables of the DB files (FP64, device).
struct Moments {
    bool genome = false;
    std::vector<uint32_t> nbins;       // per shard; 0 = shard not in the genome table
    DevBuf<uint32_t> d_nbins;
    bool present = false;
    std::vector<int64_t> slot_off;     // [nfiles * nshards]: first entry or -1
    std::vector<uint32_t> slot_len;    // entries per slot
    DevBuf<int64_t> d_slot_off;
    DevBuf<uint32_t> d_slot_len;
    DevBuf<uint8_t> d_has_entry;       // [nfiles]: file has a stored entry
    DevBuf<double> mean, var;
    uint64_t total = 0;
    int stat_rows = 0, stat_exact_rows = 0;  // signals built / whose FFT rounded exactly
};

// Per-shard sorted streams of the DB (structure of arrays in HBM).
struct Index {
    Tuning tune;
    Moments mom;
    int device = 0;
    hipStream_t stream = nullptr;
    int ncu = 256;
    uint32_t nshards = 0, nfiles = 0;
    uint64_t n = 0;
    std::vector<uint64_t> shard_off;  // nshards+1, offsets into the streams
    std::vector<uint32_t> fs_count;   // [nfiles * nshards] intervals per (file, shard)
    DevBuf<int32_t> s_coord, e_coord;  // starts / ends sorted per shard
    DevBuf<uint32_t> s_fid, e_fid;     // file id in the same order (u32 form)
    DevBuf<uint16_t> s_fid16, e_fid16; // ... u16 form, used when nfiles <= 65536
    bool fid16 = false;
    SortScratch scratch;
    Profiler prof;
    ~Index() {
        if (stream) (void)hipStreamDestroy(stream);
    }
    uint64_t device_bytes() const {
        return s_coord.bytes() + e_coord.bytes() + s_fid.bytes() + e_fid.bytes() +
               s_fid16.bytes() + e_fid16.bytes();
    }
};

// One unit of sweep work: a DB stream of one shard merged with one query
// file's event segment on that shard.
struct Job {
    uint64_t db_off;   // first element in the stream arrays
    uint32_t n_db;     // DB elements
    uint32_t q_off;    // first event of the segment in the sorted event array
    uint32_t m;        // events in the segment
    uint32_t qfile;    // output row
    uint32_t is_end;   // 0 = start stream, 1 = end stream
    uint32_t tile0;    // first global tile of this job
    uint32_t part0;    // first partition boundary of this job
    uint32_t pad;
};

struct Query {
    Index* idx = nullptr;
    uint32_t nfiles = 0;
    uint64_t n = 0;
    DevBuf<uint32_t> shard, start, end, qfile;  // uploaded intervals
    // prepared form (rebuilt by every sweep)
    DevBuf<uint64_t> keys;     // 2n event keys, sorted
    DevBuf<uint32_t> key32;    // their low 32 bits (2*coord + is_start)
    DevBuf<uint32_t> rs;       // 2n+1 exclusive prefix: # start events
    DevBuf<int64_t> w;         // 2n+1 exclusive prefix: sum(start) - sum(end)
    DevBuf<int64_t> w_sums;
    DevBuf<uint32_t> seg;      // [2][nfiles*nshards] first / one-past-last event
    std::vector<uint32_t> seg_host;
    DevBuf<Job> jobs;
    DevBuf<uint32_t> part;     // merge-path partition (DB index at each boundary)
    std::vector<Job> jobs_host;
    // query length groups per (query file, shard), built when the index has a
    // genome table (src/fourier/moments.rs:138-177)
    bool has_groups = false;
    DevBuf<uint32_t> g_off;    // [nfiles * nshards + 1]
    DevBuf<uint32_t> g_lint, g_ci, g_shard;
    DevBuf<double> g_ss, g_ss2, g_cnt;
    uint64_t ngroups = 0;
};

// Host helpers shared with the stats code.
size_t compact_index(size_t l);
size_t compact_ls(uint32_t n_bins, std::vector<uint32_t>* ls);
void query_groups(Query& q, const uint32_t* shard, const uint32_t* start, const uint32_t* end,
                  const uint32_t* qfile);
void moments_upload(Index& ix, const int64_t* slot_off, const uint32_t* slot_len, uint64_t total,
                    const double* mean, const double* var);
// Compact moment tables of every (file, shard) from the index's intervals.
void moments_build(Index& ix);
void moments_download(Index& ix, int64_t* slot_off, uint32_t* slot_len, double* mean, double* var);
// p-values for every cell with count > 0 (device arrays of nfiles(q) x F).
void pvalues_run(Query& q, const uint64_t* d_counts, const uint64_t* d_bp, double* d_p,
                 double* d_llr, uint8_t* d_flags);

void index_build(Index& ix, const uint32_t* shard, const uint32_t* start, const uint32_t* end,
                 const uint32_t* fid);
void query_upload(Query& q, const uint32_t* shard, const uint32_t* start, const uint32_t* end,
                  const uint32_t* qfile);
// Sort events, prefix sums and segment table (device), then the host job list.
void query_prepare(Query& q);
// Merge-path partition + sweep into device count / bp matrices (zeroed here).
void sweep_run(Query& q, uint64_t* d_counts, uint64_t* d_bp);
// Plain copy of `bytes` (reads and writes that many bytes once).
void calibrate_copy(Index& ix, uint64_t bytes);

}  // namespace chk
