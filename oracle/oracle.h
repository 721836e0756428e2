/*
 * oracle.h -- CPU restatement of Chuckle's file-to-file search path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product library (riggle_amd/)
 * links, loads or calls this code.  It is used by tests/, by
 * __graft_entry__.smoke() and by bench.py's cpu_baseline leg, and only there,
 * as the checker / the CPU baseline -- never as the thing measured or shipped.
 *
 * Parts (each function cites the reference file:line it restates):
 *   rng.c      rand 0.8.5 StdRng (= rand_chacha 0.3.1 ChaCha12Rng seeded by
 *              rand_core 0.6.4 seed_from_u64) and the synthetic BED generator
 *              of src/bench.rs:75-104.
 *   sweep.c    layered index + jump tables + windowed per-query sweep
 *              (src/io/meta.rs, src/io/jump.rs, src/sweep/query.rs,
 *              src/tasks/query/mod.rs), plus two exact int64 forms
 *              (brute force, rank identities).
 *   moments.c  depth signal (src/fourier/depth.rs), compact L table
 *              (src/fourier/moments.rs:46-66, 215-247), reference-method FP64
 *              moments (d*, d**, d2, FFT autocorrelation) and exact-integer
 *              moments.
 *   stats.c    query length grouping (src/fourier/moments.rs:138-177) and the
 *              analytic NB / Poisson statistics with the Abramowitz & Stegun
 *              7.1.26 erfc (src/fourier/stats.rs:34-126).
 */
#ifndef CHUCKLE_ORACLE_H
#define CHUCKLE_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------ rng.c */
typedef struct {
    uint32_t key[8];
    uint64_t ctr;      /* next block counter */
    uint32_t buf[16];  /* current keystream block */
    uint32_t idx;      /* next unread word in buf (16 = empty) */
    int rounds;
} orc_rng;

void orc_chacha_block(const uint32_t key[8], uint64_t ctr, uint64_t stream,
                      int rounds, uint32_t out[16]);
void orc_rng_from_seed(orc_rng* r, const uint8_t seed[32], int rounds);
void orc_rng_seed_u64(orc_rng* r, uint64_t seed);
uint32_t orc_rng_next_u32(orc_rng* r);
uint64_t orc_rng_next_u64(orc_rng* r);
uint32_t orc_rng_range_u32_incl(orc_rng* r, uint32_t lo, uint32_t hi);

/* generate_bed_file (unsorted branch): chrom index into the 24-name table
 * chr1..chr22, chrX, chrY; start; end. */
void orc_generate_bed(uint64_t seed, size_t n, uint32_t genome_size,
                      uint32_t min_len, uint32_t max_len,
                      uint8_t* chrom, uint32_t* start, uint32_t* end);

/* ---------------------------------------------------------------- sweep.c */
uint32_t orc_layer_for_size(uint32_t min_size, uint32_t gf, uint32_t size);
uint32_t orc_layer_max_size(uint32_t min_size, uint32_t gf, uint32_t layer);
uint32_t orc_tile_size(uint32_t min_size, uint32_t gf, uint32_t layer);
size_t orc_build_jump_table(const uint32_t* sorted_starts, size_t n,
                            uint32_t tile, uint64_t* out);

typedef struct orc_layered_db orc_layered_db;

/* Build the layered, start-sorted index (src/tasks/build.rs:196-247).
 * shard[i] in [0, nshards); sid[i] in [0, nsources). */
orc_layered_db* orc_layered_build(size_t n, const uint32_t* shard,
                                  const uint32_t* start, const uint32_t* end,
                                  const uint32_t* sid, uint32_t nshards,
                                  uint32_t nsources, uint32_t min_size,
                                  uint32_t gf, int nthreads);
void orc_layered_free(orc_layered_db* db);
size_t orc_layered_layer_len(const orc_layered_db* db, uint32_t shard,
                             uint32_t layer);
uint32_t orc_layered_num_layers(const orc_layered_db* db);

/* query_database's parallel core (src/tasks/query/mod.rs:193-298): queries
 * (shard, start, end, q_sid) of nq files; counts u32 [nq x F] and bins f32
 * [nq x F] (row-major).  Returns the number of worker threads used. */
int orc_layered_query(const orc_layered_db* db, size_t nquery,
                      const uint32_t* shard, const uint32_t* start,
                      const uint32_t* end, const uint32_t* qsid, uint32_t nq,
                      uint32_t* counts, float* bins, int nthreads);
/* The same windowed scan with exact int64 counts and overlap bp [nq x F]
 * (the reference accumulates u32 counts and f32 bins). */
int orc_layered_query_exact(const orc_layered_db* db, size_t nquery,
                            const uint32_t* shard, const uint32_t* start,
                            const uint32_t* end, const uint32_t* qsid, uint32_t nq,
                            int64_t* counts, int64_t* bp, int nthreads);

/* Exact oracles: int64 count and overlap-bp matrices [nq x F]. */
void orc_sweep_bruteforce(size_t ndb, const uint32_t* dshard,
                          const uint32_t* dstart, const uint32_t* dend,
                          const uint32_t* dsid, uint32_t F, size_t nquery,
                          const uint32_t* qshard, const uint32_t* qstart,
                          const uint32_t* qend, const uint32_t* qsid,
                          uint32_t nq, int64_t* counts, int64_t* bp);
void orc_sweep_ranks(size_t ndb, const uint32_t* dshard,
                     const uint32_t* dstart, const uint32_t* dend,
                     const uint32_t* dsid, uint32_t F, size_t nquery,
                     const uint32_t* qshard, const uint32_t* qstart,
                     const uint32_t* qend, const uint32_t* qsid, uint32_t nq,
                     uint32_t nshards, int64_t* counts, int64_t* bp);

/* -------------------------------------------------------------- moments.c */
#define ORC_BIN_SIZE 100u
size_t orc_compact_index(size_t l);
size_t orc_compact_ls(uint32_t n_bins, uint32_t* ls);
void orc_depth_signal_f32(size_t n, const uint32_t* start, const uint32_t* end,
                          uint32_t n_bins, float* g);
void orc_depth_exact(size_t n, const uint32_t* start, const uint32_t* end,
                     uint32_t n_bins, int64_t* G);
void orc_autocorr_fft(const double* x, size_t n, double* r);
void orc_moments_reference(const float* g, uint32_t n_bins,
                           const uint32_t* ls, size_t nl, double* mean,
                           double* var);
void orc_moments_exact(const int64_t* G, uint32_t n_bins, const uint32_t* ls,
                       size_t nl, double* mean, double* var, int64_t* s1,
                       double* s2);
int orc_moments_exact_fft(const int64_t* G, uint32_t n_bins, const uint32_t* ls,
                          size_t nl, double* mean, double* var);

/* ---------------------------------------------------------------- stats.c */
double orc_erfc_as(double x);
int orc_nb_llr(double mu, double var, double obs, double* llr);
int orc_poisson_llr(double mu, double obs, double* llr);
size_t orc_group_lengths(size_t n, const uint32_t* len_bp, uint32_t* l_int,
                         double* sum_scale, double* sum_scale2,
                         uint64_t* count);
int orc_analytic_stats(size_t ngroups, const uint32_t* l_int,
                       const double* sum_scale, const double* sum_scale2,
                       const uint64_t* count, const int32_t* chrom_slot,
                       const uint32_t* slot_n_bins, const uint64_t* slot_off,
                       const uint64_t* slot_len, const double* tab_mean,
                       const double* tab_var, double observed, double* p,
                       double* llr, int* has_llr);

#ifdef __cplusplus
}
#endif
#endif
