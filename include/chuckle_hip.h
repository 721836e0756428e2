/*
 * chuckle_hip.h -- C ABI of the MI355X implementation of Chuckle's
 * file-to-file search hot path (overlap sweep, coverage moments, analytic
 * p-values).  Plain C types only; every function returns CHK_OK (0) or a
 * CHK_ERR_* code and leaves a thread-local message for chk_last_error().
 * No exception or abort crosses this boundary.
 *
 * Handles are bound to one device.  Calls on one handle are serialised on
 * that handle's HIP stream; different handles may be used from different
 * threads.
 *
 * Reference interfaces replaced (simondoesstuff/riggle, crate `chuckle`):
 *   chk_index_create      -> the layered on-disk index opened by query_database
 *                            (src/io/layer.rs:34-105 MappedLayer,
 *                             src/io/jump.rs:19-52 MappedJumpTable,
 *                             src/tasks/query/mod.rs:204-241)
 *   chk_query_create      -> parse_file_batch's per-shard query intervals
 *                            (src/tasks/query/files.rs:40-111, mod.rs:193-202)
 *   chk_sweep             -> query_sweep_windowed + the window merge
 *                            (src/sweep/query.rs:84-117, overlap predicate
 *                             :105-108,
 *                             src/tasks/query/mod.rs:243-298,
 *                             src/matrix/dense.rs:4-160)
 */
#ifndef CHUCKLE_HIP_H
#define CHUCKLE_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CHK_OK 0
#define CHK_ERR_INVALID 1   /* bad argument / malformed input */
#define CHK_ERR_HIP 2       /* HIP runtime error */
#define CHK_ERR_NOMEM 3     /* device or host allocation failed */
#define CHK_ERR_IO 4        /* file read / write failed */
#define CHK_ERR_NOT_FOUND 5 /* database / file / entry missing */
#define CHK_ERR_INTERNAL 6

typedef struct chk_index chk_index;
typedef struct chk_query chk_query;

/* Library version string and the last error message of the calling thread. */
const char* chk_version(void);
const char* chk_last_error(void);

/* Number of visible HIP devices. */
int chk_device_count(int* count);

/*
 * Build an HBM-resident index from host arrays of n intervals.
 *   shard[i] < nshards  -- shard (chromosome) id of interval i
 *   start[i] < end[i] <= 2^31-1  -- half-open coordinates
 *   fid[i]   < nfiles   -- DB file id (D_SID)
 * Device layout: per (shard, group of 1024 files), a start stream (int32
 * start ascending + u16 group-local file id, stored times 8) and an end
 * stream (int32 end ascending + the same u16) for the dense merge-path
 * sweep; per (shard, file group, size layer), the intervals sorted by start
 * (int32 start, int32 end, u32 file id) with a jump table for the sparse pair
 * path -- the reference's layer files and jump tables
 * (src/io/layer.rs:34-128, src/io/jump.rs:19-113) in HBM.
 */
int chk_index_create(int device, uint32_t nshards, uint32_t nfiles, uint64_t n,
                     const uint32_t* shard, const uint32_t* start, const uint32_t* end,
                     const uint32_t* fid, chk_index** out);
int chk_index_free(chk_index* idx);

/*
 * Open a database directory in the reference's format and build its HBM
 * index (replaces query_database's opening of meta.json and the layer files,
 * src/tasks/query/mod.rs:123-241, src/io/meta.rs:124-127,
 * src/io/layer.rs:34-105): DB/meta.json, DB/shards/<shard>/layer_K.{pos,sid}
 * read over host threads.  Shard ids follow meta.json's "shards" order, file
 * ids are the D_SIDs.  The genome table is set to hg38 by shard name
 * (src/fourier/genome.rs:7-34), and DB/momentmap.chk is loaded when it
 * matches the database (otherwise build the tables with chk_moments_build).
 */
int chk_index_open(const char* db_path, int device, chk_index** out);
/* Names of an opened database's shards and files (the meta.json sid_map),
 * NUL-terminated into buf (truncated to len - 1 bytes); needed (optional)
 * receives the full size with the NUL.  CHK_ERR_NOT_FOUND for an index made
 * by chk_index_create. */
int chk_index_shard_name(const chk_index* idx, uint32_t shard, char* buf, uint64_t len,
                         uint64_t* needed);
int chk_index_file_name(const chk_index* idx, uint32_t fid, char* buf, uint64_t len,
                        uint64_t* needed);
/* Attach shard and file names to an index built by chk_index_create (nshards
 * and nfiles strings, copied), so that chk_query_open and the name queries
 * work on it as on an opened database. */
int chk_index_set_names(chk_index* idx, const char* const* shard_names, const char* const* file_names);

/*
 * `chuckle add` (replaces add_to_database, src/tasks/build.rs:86-276): the
 * BED / BED.gz files of `input` (a file or a directory) are parsed over host
 * threads in batches of batch_size files (0: all), every batch's intervals
 * are sorted by (shard, size layer, start) on `device`, and the layer files,
 * jump tables and meta.json of db_path are written or extended.  D_SIDs
 * follow the directory's order, as the reference's read_dir does, or the
 * sorted path order when sorted_sids != 0.  With stats != 0 the moment
 * tables of every file are built on the device (chromosomes limited to the
 * nchroms names when nchroms > 0) and saved as DB/momentmap.chk.
 */
int chk_db_add(const char* input, const char* db_path, uint64_t batch_size, int stats,
               const char* const* chroms, uint32_t nchroms, int sorted_sids, int device);
int chk_index_info(const chk_index* idx, uint32_t* nshards, uint32_t* nfiles, uint64_t* n,
                   uint64_t* device_bytes);

/*
 * Upload a batch of query files: n intervals with shard ids in the index's
 * shard numbering (ids >= nshards mark shards absent from the DB; their
 * intervals are ignored, as the reference skips such shards), coordinates as
 * above and qfile[i] < nfiles the query-file row (Q_SID).  qfile may be NULL
 * when nfiles == 1.
 */
int chk_query_create(chk_index* idx, uint32_t nfiles, uint64_t n, const uint32_t* shard,
                     const uint32_t* start, const uint32_t* end, const uint32_t* qfile,
                     chk_query** out);
int chk_query_free(chk_query* q);
/* Replace a query handle's batch by a new one (arguments as chk_query_create),
 * reusing its host staging and device buffers: a loop that uploads batch
 * after batch allocates only when a batch outgrows the earlier ones.  Waits
 * for work still reading the handle's buffers.  No counterpart in the
 * reference. */
int chk_query_reset(chk_query* q, uint32_t nfiles, uint64_t n, const uint32_t* shard,
                    const uint32_t* start, const uint32_t* end, const uint32_t* qfile);
/*
 * Parse BED / BED.gz query files over host threads and upload them as one
 * batch (replaces parse_file_batch, src/tasks/query/files.rs:40-111, with
 * src/io/parse.rs:53-141 rules) on an index opened with chk_index_open:
 * chromosome names map to its shards (intervals on other names are
 * dropped), empty files take no row, and the other files take rows in path
 * order; row_of_path (optional, npaths entries) receives each path's row or
 * 0xFFFFFFFF.  When every file is empty, *out is set to NULL.
 */
int chk_query_open(chk_index* idx, uint32_t npaths, const char* const* paths, uint32_t* row_of_path,
                   chk_query** out);
/*
 * The batch's sweep plan under the index's current tuning: query files and
 * intervals kept (on shards of the DB), and how many files (and their
 * intervals) take the dense merge-path sweep, which streams the whole DB per
 * file; the others take the sparse pair sweep.  No counterpart in the
 * reference (its sweep has a single method).
 */
int chk_query_info(chk_query* q, uint32_t* nfiles, uint64_t* n, uint32_t* dense_files,
                   uint64_t* dense_intervals);

/*
 * Overlap sweep: for every (query file r, DB file f) cell,
 *   counts[r*F + f] = #{(q, d) : same shard, d.start < q.end && d.end > q.start}
 *   bp[r*F + f]     = sum over those pairs of |q intersect d| in base pairs,
 * exact u64 (the reference's observed_bins is bp / 100).  Host outputs of
 * nfiles(query) x nfiles(index) elements; blocks until done.
 */
int chk_sweep(chk_index* idx, chk_query* q, uint64_t* counts, uint64_t* bp);

/* Same, into device buffers (chk_dev_alloc), enqueued on the index's stream
 * without a final synchronisation. */
int chk_sweep_async(chk_index* idx, chk_query* q, uint64_t* d_counts, uint64_t* d_bp);

/*
 * Genome table used by the statistics (src/fourier/genome.rs:7-34 is the
 * reference's fixed hg38 table): shard_size[s] = length in bp of shard s, or 0
 * for a shard outside the table (no moments, no p-value terms), as the
 * reference ignores chromosomes absent from its table.  n_bins = ceil(size/100).
 * Must be set before queries whose p-values are wanted are created.
 */
int chk_index_set_genome(chk_index* idx, const uint64_t* shard_size);

/*
 * Install compact moment tables (the reference's momentmap store,
 * src/io/momentstore.rs:20-106), one per (DB file f, shard s):
 * slot_off[f*nshards+s] = first entry in mean/var, or -1 when the file stores
 * no table for the shard; slot_len = entries in the slot; entry i of a slot is
 * block size L with compact_index(L) == i (src/fourier/moments.rs:58-66).
 */
int chk_moments_upload(chk_index* idx, const int64_t* slot_off, const uint32_t* slot_len,
                       uint64_t total, const double* mean, const double* var);

/*
 * Build the compact moment tables of every (DB file, shard) with intervals on
 * a genome shard on the device (replaces build_chrom_moments and the moment
 * store of `chuckle add --stats`, src/tasks/build.rs:160-194,
 * src/fourier/moments.rs:129-254): exact integer coverage per 100-bp bin,
 * FFT autocorrelation of the centred signal (radix-8 LDS passes, FP64),
 * double-double prefix sums, mean/var per stored block size.
 * exact_rows (optional) receives the number of signals whose autocorrelation
 * was rounded to exact integers, rows the number of signals.
 */
int chk_moments_build(chk_index* idx, int32_t* rows, int32_t* exact_rows);

/*
 * Precision form of the moments (no counterpart in the reference, whose
 * tables come from one FP64 FFT per chromosome, src/fourier/moments.rs:421-466):
 * mean and variance of every (DB file, shard) at nwidths block sizes of 1..1024
 * bins, from FP32 FFTs of 4096-point blocks (hop 3072) of the centred
 * coverage that give the autocorrelation lags below 1024, accumulated in FP32
 * two-sum (compensated) arithmetic; FP64 only in the variance step
 * E[x^2] - E[x]^2.  Host outputs mean / var of nfiles x nshards x nwidths
 * (NaN where the file has no intervals on the shard, the shard is outside the
 * genome table, or L >= n_bins).  inexact_rows (optional) receives the number
 * of signals with a centred coverage value of 2^24 or more (not exact in
 * FP32), bins (optional) the bins processed, lags (optional, nfiles x
 * nshards x 1024) the lags A[0..1023] as FP64.  Signals of at most 8192 bins,
 * and widths whose FP32 error bound (2.5e-7 A[0] L^1.5) exceeds 5e-7 of the
 * signal's variance sum, are evaluated by exact integer window sums instead;
 * exact_widths (optional) receives how many (signal, width) pairs were.
 * Needs the genome table.
 */
int chk_moments_widths(chk_index* idx, uint32_t nwidths, const uint32_t* widths, double* mean,
                       double* var, uint32_t* inexact_rows, uint64_t* bins, double* lags,
                       uint64_t* exact_widths);

/* Copy the installed tables to the host: *total receives the entry count;
 * arrays may be NULL (slot arrays hold nfiles*nshards entries). */
int chk_moments_download(chk_index* idx, uint64_t* total, int64_t* slot_off, uint32_t* slot_len,
                         double* mean, double* var);

/*
 * Analytic p-values (replaces compute_analytic_pvalues +
 * compute_analytic_stats, src/tasks/query/pvalues.rs:22-77,
 * src/fourier/stats.rs:34-126) for every cell with count > 0 of a sweep
 * result.  Per cell: p, llr and flags (bit 0: a record is emitted -- count > 0,
 * the DB file has moments and shares a chromosome with the query; bit 1: the
 * LLR exists).  Cells without bit 0 carry no record (the reference's absent
 * entry) and are written p = 1, llr = 0.  Device arrays; enqueued without
 * final synchronisation.
 */
int chk_pvalues_async(chk_index* idx, chk_query* q, const uint64_t* d_counts, const uint64_t* d_bp,
                      double* d_p, double* d_llr, uint8_t* d_flags);

/* Sweep + p-values into host arrays of nfiles(query) x nfiles(index); blocks. */
int chk_query_stats(chk_index* idx, chk_query* q, uint64_t* counts, uint64_t* bp, double* p,
                    double* llr, uint8_t* flags);

/*
 * Write a result matrix as the reference's StatsOutput JSON (replaces the
 * output step of src/main.rs:146-211, layout of src/stats/mod.rs:31-60):
 * with stats != 0 one record per cell with flags bit 0, sorted by p
 * ascending, with observed_bins = bp / 100, p_value and llr (when flags bit
 * 1); otherwise one record per nonzero count, sorted by count descending.
 */
int chk_write_results(const char* path, uint32_t nquery, const char* const* query_names,
                      uint32_t nfiles, const char* const* db_names, const uint64_t* counts,
                      const uint64_t* bp, const double* p, const double* llr, const uint8_t* flags,
                      int stats);

/* Device memory helpers for callers without a device allocator. */
int chk_dev_alloc(chk_index* idx, uint64_t bytes, void** ptr);
int chk_dev_free(chk_index* idx, void* ptr);
int chk_dev_to_host(chk_index* idx, void* dst, const void* src, uint64_t bytes);
int chk_sync(chk_index* idx);

/*
 * Kernel timing with HIP events recorded on the index's stream around each
 * named kernel ("sweep", "partition", "qprep", ...).  chk_profile_read
 * synchronises and returns the number of launches and their summed duration.
 * chk_profile_only restricts recording to a comma-separated list of region
 * names (NULL or "" records every region); each recorded region costs the
 * stream a few microseconds.
 */
int chk_profile_enable(chk_index* idx, int on);
int chk_profile_only(chk_index* idx, const char* regions);
int chk_profile_read(chk_index* idx, const char* name, uint64_t* launches, double* total_ms);

/*
 * Launch-shape knobs of the sweep, set by the autotuner or by tests
 * (value 0 restores the automatic choice):
 *   "sweep.blocks_per_cu"  resident dense-sweep blocks per CU (persistent grid)
 *   "sweep.chunk"          merge-path tiles per work-queue item (1..64)
 *   "sweep.items"          DB elements per thread per tile: 4, 8, 12 or 16
 *                          (tile = 1024, 2048 or 4096 merged items)
 *   "sweep.sparse_items"   query intervals per sparse work item
 *   "sweep.sparse_max"     largest query file (intervals) swept by the sparse
 *                          pair path; larger files take the dense path
 *                          (0: each file's estimated sparse-path work, from
 *                          the jump tables, against the cost of a DB pass)
 *   "sweep.method"         0 choose per file (sweep.sparse_max or the work
 *                          estimate), 1 dense only, 2 sparse only
 *   "moments.fft_mib"      FFT workspace of one launch of the moment build's
 *                          four-step FFT, in MiB (0: the default, 128)
 *   "moments.fft_block"    complex values one four-step row block holds in
 *                          LDS (a row and its mirror): 4096 (64 KiB, two
 *                          blocks per CU) or 8192 (128 KiB, where the device
 *                          allows it); 0: 4096
 *   "moments.round"        0: round the FFT autocorrelations to the exact
 *                          integers they are when every value lies within
 *                          0.25 of one (default); 1: never round; 2: round
 *                          only when every value already is an integer (a
 *                          row that is not is scanned again unrounded)
 *   "moments.eval_seg"     bins per independent segment of the moment
 *                          evaluation's prefix scans, a multiple of 4096
 *                          (0: 65536; raised so that a chromosome has at
 *                          most 256 segments)
 *   "moments.fft_pipe"     forward column pass of the four-step FFT: 0
 *                          persistent blocks that load the next column
 *                          group while transforming the current one
 *                          (default); 1 one block per column group
 * The sweep keys apply to every later sweep, chk_query_info and graph
 * creation (an existing query handle is planned again at its next use); the
 * moments keys apply to later moment builds.  A graph captured before any
 * chk_tune call refuses to launch after it.
 */
int chk_tune(chk_index* idx, const char* key, int64_t value);
/* Current value of a tuning key (0 = automatic choice). */
int chk_tune_get(chk_index* idx, const char* key, int64_t* value);

/*
 * Autotuner (no counterpart in the reference, whose sweep has no launch
 * shape): coordinate descent over sweep.sparse_max, sweep.items, sweep.chunk
 * and sweep.blocks_per_cu on this query batch.  Each candidate is swept once
 * to warm up, then `reps` times, timed with HIP events; the fastest value of
 * each knob is kept in the index's tuning (read it back with chk_tune_get).
 * d_counts / d_bp are device outputs of the batch's size (overwritten).
 * best_ms receives the best time per sweep; report (optional, report_len
 * bytes) one line "knob=value: ms" per candidate.
 */
int chk_autotune(chk_index* idx, chk_query* q, uint64_t* d_counts, uint64_t* d_bp, int reps,
                 double* best_ms, char* report, uint64_t report_len);

/*
 * Batched form as one HIP graph: the whole sweep of a query batch (output
 * zeroing, event sort, prefix sums, merge-path partition, the persistent
 * work-queue dense sweep and the sparse pair sweep) and, when d_p is given,
 * the p-value kernel, captured on the index's stream.  Replays need no host
 * work and no allocation.  The graph replays the batch it was created for
 * (its launch plan -- tiles, chunks, event key widths -- is that batch's); a
 * new batch takes a new query and graph.  It must also be recreated after
 * the index's tuning changes (chk_graph_launch then reports an error).
 * The graph holds the query handle and the device buffers it was captured
 * with by address only: the query and every output buffer must outlive the
 * graph (free the graph first).
 */
typedef struct chk_graph chk_graph;
int chk_graph_create(chk_index* idx, chk_query* q, uint64_t* d_counts, uint64_t* d_bp,
                     double* d_p, double* d_llr, uint8_t* d_flags, chk_graph** out);
int chk_graph_launch(chk_graph* g);
int chk_graph_free(chk_graph* g);

/*
 * Serving form (no counterpart in the reference, which runs one `chuckle
 * query` per process): a request loop that keeps the index resident, admits
 * query files by HBM footprint and batches them as they arrive.  One
 * scheduler thread owns the device work; `slots` batches (1..8) may be in
 * flight at once, each on its own stream, so requests that arrive while
 * others run are admitted as soon as a slot and budget free up and overlap
 * on the device.  A batch takes at most max_batch_files waiting requests;
 * hbm_budget bounds the device bytes of in-flight requests (0: half of the
 * free memory at creation; a request above the budget runs alone).  With
 * pvalues != 0 the index needs its genome and moment tables.  A request's
 * results do not depend on the batch it shared.  While a server runs, do not
 * enable profiling on its index.
 */
typedef struct chk_server chk_server;
int chk_server_create(chk_index* idx, uint32_t slots, uint32_t max_batch_files,
                      uint64_t hbm_budget, int pvalues, chk_server** out);
/* Queue one query file (validated and copied, over host threads when large;
 * returns once it is queued) and get its ticket. */
int chk_server_submit(chk_server* s, uint64_t n, const uint32_t* shard, const uint32_t* start,
                      const uint32_t* end, uint64_t* ticket);
/* Block until the request is done; outputs are rows of nfiles(index) entries
 * (any may be NULL).  latency_ms = completion - the chk_server_submit call
 * (its validation and copy included). */
int chk_server_wait(chk_server* s, uint64_t ticket, uint64_t* counts, uint64_t* bp, double* p,
                    double* llr, uint8_t* flags, double* latency_ms);
/* Latency percentiles (ms) over completed requests; swept_per_s =
 * sum(|Q| + |DB|) over them / (last completion - first submission). */
int chk_server_stats(chk_server* s, uint64_t* completed, double* p50_ms, double* p99_ms,
                     double* mean_ms, double* swept_per_s, double* files_per_batch);
int chk_server_reset_stats(chk_server* s);
/* Grow every slot's buffers for batches of up to max_batch_intervals
 * intervals (and max_batch_files files) while the server is idle, so that
 * later batches up to that size allocate nothing: growing a device buffer
 * mid-stream frees the old one, which waits for the whole device and stalls
 * every other slot.  No counterpart in the reference. */
int chk_server_reserve(chk_server* s, uint64_t max_batch_intervals);
/* Waits for in-flight batches, then stops the loop.  Requests still waiting
 * for admission are failed ("server stopped ...").  Do not call
 * chk_server_wait or chk_server_submit on the server while, or after, it is
 * freed: collect every ticket first. */
int chk_server_free(chk_server* s);

/*
 * Allocation-free sweep of one query file.  No counterpart in the reference:
 * it is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  The same
 * results as chk_query_create (nfiles = 1) + chk_sweep_async: n intervals in
 * host arrays (shard ids as in chk_query_create; ids >= nshards are
 * skipped), outputs d_counts / d_bp = device buffers of nfiles(index) u64
 * each, overwritten.  The file is swept by the sparse pair kernel with its
 * intervals and work items in a static device array of
 * CHK_NOALLOC_SCRATCH_BYTES per device (held until the process ends); no
 * hipMalloc, hipFree or other allocating call is made.  n above
 * CHK_NOALLOC_MAX_INTERVALS is rejected with CHK_ERR_INVALID.  Calls on one
 * device share its scratch array: the library serialises them with a
 * per-device lock, so overlapping calls from several threads of a process
 * wait for each other (work a caller enqueues itself cannot touch the
 * array).  Blocks until the outputs are written.
 */
#define CHK_NOALLOC_MAX_INTERVALS (1u << 20)
#define CHK_NOALLOC_SCRATCH_BYTES (12ull * CHK_NOALLOC_MAX_INTERVALS + 16ull * 65536)
int chk_sweep_noalloc(chk_index* idx, uint64_t n, const uint32_t* shard, const uint32_t* start,
                      const uint32_t* end, uint64_t* d_counts, uint64_t* d_bp);

/* Read-only ceiling of the sweep's input: a streaming kernel that reads the
 * index's dense streams (int32 start / end coordinates, u16 file ids: what
 * the dense sweep reads per query file) once with 16-byte loads and nothing
 * else; best time over reps runs at three grid sizes, and the bytes read. */
int chk_calibrate_read(chk_index* idx, uint32_t reps, double* best_ms, uint64_t* bytes);

/* Launch a plain device copy kernel ("calib_copy") that reads and writes
 * `bytes` bytes once; used to scale hardware byte counters. */
int chk_calibrate_copy(chk_index* idx, uint64_t bytes);

#ifdef __cplusplus
}
#endif
#endif
