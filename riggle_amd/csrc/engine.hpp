// engine.hpp -- the HBM-resident index, prepared query batches and the
// merge-path overlap sweep.
#pragma once

#include <map>
#include <set>
#include <string>
#include <utility>

#include "common.hpp"
#include "sort.hpp"

namespace chk {

// HIP-event timing of named kernels on one stream.
struct Profiler {
    bool on = false;
    std::set<std::string> only;  // regions to record; empty = all
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
    // Returns the start event (or nullptr when the region is not recorded).
    hipEvent_t begin(const char* name, hipStream_t st) {
        if (!on || (!only.empty() && !only.count(name))) return nullptr;
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

// DB files per group: the dense sweep keeps one LDS histogram of a group's
// files per block, so the DB streams are split by (shard, file group) and
// carry group-local 16-bit file ids.
constexpr uint32_t FG = 1024;

// Strides of the stream samples that the merge-path split search starts
// from: coarse (every PSTRIDE2-th element), then fine (every PSTRIDE-th).
constexpr uint32_t PSTRIDE = 64;
constexpr uint32_t PSTRIDE2 = 4096;
// Size classes of the layered start-sorted table (src/io/meta.rs:33-81,
// LayerConfig::default_genomic: min_size 128, growth factor 4).
constexpr uint32_t LAYER_MIN = 128, LAYER_GF = 4;

// Launch-shape and method knobs of the sweep (0 = choose automatically).
struct Tuning {
    uint32_t blocks_per_cu = 0;  // resident dense-sweep blocks per CU
    uint32_t chunk = 0;          // merge-path tiles per work-queue item
    uint32_t items = 0;          // DB elements per thread per tile: 4, 8 or 16
    uint64_t sparse_max = 0;     // largest query file (intervals) swept by the sparse path
    uint32_t method = 0;         // 0 auto by size, 1 dense only, 2 sparse only
    uint32_t sparse_items = 0;   // query intervals per sparse work item
    uint32_t fft_mib = 0;        // moment FFT workspace per launch (MiB), 0 = default
    uint32_t fft_block = 0;      // complex values per four-step row block: 0 (= 4096), 4096 or 8192
    uint32_t mom_round = 0;      // autocorrelation rounding: 0 auto, 1 never, 2 only exact values
    uint32_t eval_seg = 0;       // positions per evaluation segment (multiple of 4096), 0 = default
    uint32_t fft_pipe = 0;       // forward column pass: 0 persistent pipelined blocks, 1 one block per group
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

// One (shard, file group, layer) segment of the layered start-sorted table with its jump
// table: jt[t] = first element whose start >= t << shift (src/io/jump.rs:66-113
// with a power-of-two bucket sized to the segment).
struct LayerSeg {
    uint64_t off;       // first element in l_start / l_end / l_fid
    uint64_t jt;        // first entry in the jump table
    uint32_t n;         // elements
    uint32_t nb;        // jump entries
    uint32_t shift;     // bucket = start >> shift
    uint32_t lookback;  // layer_max_size: every length in the layer is below it
};

// HBM-resident index.  Dense form: per (shard, file group) a start stream
// (int32 start ascending + u16 group-local file id times 8) and an end stream
// (int32 end ascending + u16 id times 8), 12 bytes per interval.  Sparse form: per (shard,
// file group, size layer) the intervals sorted by start as int32 start /
// int32 end / u32 file id plus a jump table, 12 bytes per interval + tables.
struct Index {
    Tuning tune;
    std::vector<std::string> shard_names, file_names;  // set when opened from a DB directory
    uint32_t tune_version = 0;
    Moments mom;
    int device = 0;
    hipStream_t stream = nullptr;
    int ncu = 256;
    uint32_t nshards = 0, nfiles = 0, ngroups = 0;
    uint64_t n = 0;
    std::vector<uint32_t> fs_count;   // [nfiles * nshards] intervals per (file, shard)
    std::vector<uint64_t> shard_n;    // intervals per shard
    std::vector<uint64_t> str_off;    // [nshards * ngroups + 1] stream segment offsets
    DevBuf<int32_t> s_coord, e_coord;
    DevBuf<int32_t> s_samp, e_samp;   // every PSTRIDE-th stream coordinate (split search)
    DevBuf<int32_t> s_samp2, e_samp2; // every PSTRIDE2-th stream coordinate
    DevBuf<uint16_t> s_lfid, e_lfid;  // group-local file id * 8 (histogram slot byte offset)
    uint32_t nlayers = 0;
    uint64_t span = 0;                // sum over shards of the largest start
    std::vector<uint64_t> layer_n;    // intervals per size layer (all shards)
    std::vector<LayerSeg> lseg;       // [(nshards * ngroups + group) * nlayers + layer]
    DevBuf<LayerSeg> d_lseg;
    DevBuf<int32_t> l_start, l_end;
    DevBuf<uint32_t> l_fid;
    DevBuf<uint32_t> jt;
    SortScratch scratch;
    Profiler prof;
    ~Index() {
        if (stream) (void)hipStreamDestroy(stream);
    }
    uint64_t device_bytes() const {
        return s_coord.bytes() + e_coord.bytes() + s_samp.bytes() + e_samp.bytes() +
               s_samp2.bytes() + e_samp2.bytes() +
               s_lfid.bytes() + e_lfid.bytes() +
               l_start.bytes() + l_end.bytes() + l_fid.bytes() + jt.bytes() + d_lseg.bytes();
    }
};

// One unit of dense sweep work: the start or end stream of one (shard, file
// group) merged with one query file's event segment on that shard.
struct Job {
    uint64_t db_off;  // first element in the stream arrays
    uint32_t n_db;    // DB elements
    uint32_t q_off;   // first event of the segment in the sorted event array
    uint32_t m;       // events in the segment
    uint32_t row;     // output row (query file)
    uint32_t group;   // DB file group
    uint32_t is_end;  // 0 = start stream, 1 = end stream
    uint32_t tile0;   // first global tile
    uint32_t part0;   // first merge-path boundary
};

// One query file's block in the canonical and the placed layouts.
struct PlaceFile {
    uint64_t src, dst;            // first interval in the canonical / placed arrays
    uint32_t n, row, drank, pad;  // drank = rank among the dense files, or ~0
};

// Sparse work item: query intervals [begin, end) (placed order) of one file
// against the DB files of one group.
struct SparseItem {
    uint32_t row, begin, end, group;
};

// A second stream for work that may overlap the caller's, with the events
// that fork it from and join it back to the caller's stream (created on first
// use, on the current device).
struct ForkStream {
    hipStream_t s = nullptr;
    hipEvent_t fork = nullptr, join = nullptr;
    ForkStream() = default;
    ForkStream(const ForkStream&) = delete;
    ForkStream& operator=(const ForkStream&) = delete;
    ~ForkStream() {
        if (fork) (void)hipEventDestroy(fork);
        if (join) (void)hipEventDestroy(join);
        if (s) (void)hipStreamDestroy(s);
    }
    void init() {
        if (s) return;
        CHK_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
        CHK_HIP(hipEventCreateWithFlags(&fork, hipEventDisableTiming));
        CHK_HIP(hipEventCreateWithFlags(&join, hipEventDisableTiming));
    }
};

struct Query {
    Index* idx = nullptr;
    uint32_t nfiles = 0;
    hipStream_t stream = nullptr;     // uploads and plan copies (the index's stream by default)
    ForkStream aux;                   // the sparse sweep beside the dense one
    uint64_t n_in = 0;                // intervals given
    uint32_t max_end = 0;             // largest end coordinate kept (sizes the event keys)
    uint64_t n = 0;                   // intervals kept (on shards of the DB)
    std::vector<double> file_len;     // total length per query file
    std::vector<double> file_visits;  // estimated sparse-path visits per query file
    DevBuf<uint64_t> vis_off, vis;    // their device-side inputs and counters
    std::vector<uint64_t> fs_n;       // [nfiles * nshards] intervals per (file, shard)
    std::vector<uint64_t> file_off;   // canonical layout: first interval of each file
    PinnedBuf<uint32_t> st_shard, st_start, st_end;  // canonical layout, host staging
    DevBuf<uint32_t> c_shard, c_start, c_end;         // canonical layout on the device
    DevBuf<PlaceFile> place;
    bool staged = false, placed = false;
    uint64_t n_dense = 0;             // the first n_dense belong to dense files
    std::vector<uint64_t> file_n;     // intervals per query file
    std::vector<uint8_t> file_dense;  // method per query file
    std::vector<uint32_t> dense_rows; // rows of the dense files, in upload order
    std::vector<uint32_t> seg_n;      // [ndense * nshards] intervals per (dense file, shard)
    // uploaded intervals, grouped by file (dense files first)
    DevBuf<uint32_t> shard, start, end, row, segid;
    uint32_t plan_version = 0xFFFFFFFFu;
    // dense plan
    uint32_t tile = 0, chunk = 0, items = 0, ntiles = 0, nbounds = 0, nchunks = 0, grid = 0;
    std::vector<Job> jobs_host;
    std::vector<uint32_t> chunk_job_host;
    DevBuf<Job> jobs;
    DevBuf<uint32_t> chunk_job;
    DevBuf<uint32_t> part;            // merge-path split (DB index) at each tile boundary
    DevBuf<uint32_t> queue;           // per-XCD work-queue counters (zeroed by each sweep)
    // sparse plan
    std::vector<SparseItem> items_host;
    DevBuf<SparseItem> sitems;
    // prepared events (rebuilt by every sweep)
    DevBuf<uint64_t> keys;            // 2 n_dense event keys, sorted
    DevBuf<uint32_t> key32;           // their low 32 bits (2*coord + is_start)
    DevBuf<uint32_t> rs;              // exclusive prefix: # start events
    DevBuf<int64_t> w;                // exclusive prefix: sum(start) - sum(end)
    SortScratch scratch;              // per query, so batches on other streams do not share it
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
// FP32 lag-limited route: mean / var of every (file, shard) at block sizes
// widths[0..nw) <= 1024 bins, [F x S x nw] host arrays (NaN where a slot has
// no intervals or L >= n_bins); lags_out (optional, [F x S x 1024]) receives
// the FP64 values of the autocorrelation lags; exact_widths the (signal,
// width) pairs evaluated by exact window sums (short signals, or an FP32
// error bound above 5e-7 of the variance sum).
void moments_widths(Index& ix, const uint32_t* widths, uint32_t nw, double* mean, double* var,
                    uint32_t* inexact_rows, uint64_t* bins, double* lags_out, uint64_t* exact_widths);
// p-values for every cell with count > 0 (device arrays of nfiles(q) x F).
void pvalues_run(Query& q, const uint64_t* d_counts, const uint64_t* d_bp, double* d_p,
                 double* d_llr, uint8_t* d_flags, hipStream_t st);

// Binds an empty index to a device (its stream, CU count) for n intervals of
// nfiles files on nshards shards.
void index_init(Index& ix, int device, uint32_t nshards, uint32_t nfiles, uint64_t n);
void index_build(Index& ix, const uint32_t* shard, const uint32_t* start, const uint32_t* end,
                 const uint32_t* fid);
// Genome table of the statistics: shard_size[s] bp, 0 = not in the table.
void set_genome(Index& ix, const uint64_t* shard_size);
void query_upload(Query& q, const uint32_t* shard, const uint32_t* start, const uint32_t* end,
                  const uint32_t* qfile);
// Canonical layout (grouped by file) in the staging buffers -> device, then plan.
void query_stage_commit(Query& q);
// Whether a query file of n intervals and estimated sparse visits takes the
// dense path.
bool file_is_dense(const Index& ix, uint64_t n, double visits);
// Estimated sparse-path visits per query file (device, from the jump tables).
void query_visits(Query& q);
// Method per file and device placement for the current tuning.
void query_place(Query& q);
// Placement plus the host work plan of the dense and sparse paths for the
// current tuning (no-op when the tuning is unchanged).
void query_plan(Query& q);
// One sweep: zero the outputs, sort the dense files' events, prefix sums,
// merge-path partition, dense sweep, sparse sweep.  Enqueued on `st` with
// sizes known on the host only, so the sequence can be graph-captured.
void sweep_run(Query& q, uint64_t* d_counts, uint64_t* d_bp, hipStream_t st);
// The sparse pair sweep over `nitems` work items of query intervals given by
// (shard, start, end) device arrays; rows come from the items.  Intervals on
// shard ids >= the index's nshards are skipped.
void sparse_launch(Index& ix, const SparseItem* items, uint32_t nitems, const uint32_t* shard,
                   const uint32_t* start, const uint32_t* end, uint64_t* d_counts, uint64_t* d_bp,
                   hipStream_t st);
// Allocation-free sweep of one query file through the sparse pair path with
// the library's static device scratch array (noalloc.hip).
void sweep_noalloc(Index& ix, uint64_t n, const uint32_t* shard, const uint32_t* start,
                   const uint32_t* end, uint64_t* d_counts, uint64_t* d_bp);
// Fill n u64 with zero from a kernel (graph-safe clearing).
void zero_u64(uint64_t* p, size_t n, hipStream_t st);
// zero fills of three buffers (a kernel, not a memset: sequences may be graph-captured)
void zero3_u64(uint64_t* p0, size_t n0, uint64_t* p1, size_t n1, uint64_t* p2, size_t n2,
               hipStream_t st);
// Mean time of `reps` sweeps (after one warm-up) with HIP events.
double time_sweeps(Query& q, uint64_t* d_counts, uint64_t* d_bp, int reps);
// Coordinate-descent autotuner over the sweep knobs for this batch; leaves the
// best values in the index's tuning and returns the best time per sweep.
double autotune(Query& q, uint64_t* d_counts, uint64_t* d_bp, int reps, std::string* report);
// Serving form (server.cpp).
struct Server;
Server* server_create(Index& ix, uint32_t slots, uint32_t max_batch_files, uint64_t hbm_budget,
                      bool pvalues);
void server_free(Server* s);
uint64_t server_submit(Server& s, uint64_t n, const uint32_t* shard, const uint32_t* start,
                       const uint32_t* end);
void server_wait(Server& s, uint64_t ticket, uint64_t* counts, uint64_t* bp, double* p, double* llr,
                 uint8_t* flags, double* latency_ms);
void server_stats(Server& s, uint64_t* completed, double* p50, double* p99, double* mean,
                  double* swept_per_s, double* files_per_batch);
void server_reset_stats(Server& s);
void server_reserve(Server& s, uint64_t max_batch_intervals);
// Grows a query's buffers for batches of up to n intervals in `files` files
// (the plan's for the current tuning), so that such batches allocate nothing.
void query_reserve(Query& q, uint64_t n, uint32_t files);
void plan_reserve(Query& q, uint64_t n, uint32_t files);
// Plain copy of `bytes` (reads and writes that many bytes once).
void calibrate_copy(Index& ix, uint64_t bytes);
// Read-only streaming of the index's dense streams (what the dense sweep
// reads), best time over `reps` runs at three grid sizes.
void calibrate_read(Index& ix, uint32_t reps, double* best_ms, uint64_t* bytes);

}  // namespace chk
