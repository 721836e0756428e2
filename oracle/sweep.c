/*
 * sweep.c -- CPU restatement of the reference's file-to-file overlap sweep
 * (TEST INFRASTRUCTURE ONLY; see oracle.h).
 *
 * Three forms:
 *  1. orc_layered_build / orc_layered_query restate the reference CPU path:
 *     size-class layers (src/io/meta.rs:33-81), start-sorted layer buckets
 *     (src/tasks/build.rs:196-211), jump tables (src/io/jump.rs:46-51,
 *     66-113), the per-query look-back scan (src/sweep/query.rs:84-117,
 *     predicate d.start < q.end && d.end > q.start, u32 count += 1 and f32
 *     bins += overlap/100 per hit), 64-row Q_SID windows run in parallel and
 *     merged serially (src/tasks/query/mod.rs:193-285).  Shards are processed
 *     one after another and only the windows of one shard run concurrently,
 *     exactly as the reference schedules them.  This is also the CPU baseline.
 *  2. orc_sweep_bruteforce: every (q, d) pair, int64 count and bp.
 *  3. orc_sweep_ranks: the pair-free identities
 *        count = sum_d #{q.s < d.e} - #{q.e <= d.s}
 *        bp    = sum_d CQ(d.e) - CQ(d.s),
 *        CQ(x) = x*#{q.s<x} - sum_{q.s<x} q.s - x*#{q.e<x} + sum_{q.e<x} q.e,
 *     evaluated with binary searches over the sorted query starts / ends.
 */
#include "oracle.h"

#include <pthread.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

/* ---------------------------------------------------------- layer config */

/* src/io/meta.rs:33-58 */
uint32_t orc_layer_for_size(uint32_t min_size, uint32_t gf, uint32_t size) {
    if (size < min_size) return 0;
    if ((min_size & (min_size - 1)) == 0 && (gf & (gf - 1)) == 0 && gf >= 2) {
        uint32_t log2_size = 31u - (uint32_t)__builtin_clz(size);
        uint32_t log2_min = 31u - (uint32_t)__builtin_clz(min_size);
        uint32_t log2_gf = (uint32_t)__builtin_ctz(gf);
        uint32_t d = log2_size > log2_min ? log2_size - log2_min : 0;
        return d / log2_gf + 1;
    }
    uint64_t threshold = min_size;
    uint32_t layer = 0;
    while (threshold <= size) {
        threshold = threshold * gf;
        if (threshold > UINT64_MAX / 2) threshold = UINT64_MAX / 2;
        layer++;
    }
    return layer;
}

/* src/io/meta.rs:63-72 */
uint32_t orc_layer_max_size(uint32_t min_size, uint32_t gf, uint32_t layer) {
    uint64_t v = min_size;
    for (uint32_t i = 0; i < layer; i++) {
        v *= gf;
        if (v > 0xFFFFFFFFull) v = 0xFFFFFFFFull;
    }
    return (uint32_t)v;
}

/* src/io/meta.rs:78-80 */
uint32_t orc_tile_size(uint32_t min_size, uint32_t gf, uint32_t layer) {
    uint32_t t = orc_layer_max_size(min_size, gf, layer) / 2;
    return t ? t : 1;
}

/* src/io/jump.rs:66-113: table[t] = #{starts < t*tile}, t = 0..=max_tile;
 * an empty layer gets the single entry [0].  Returns the table length. */
size_t orc_build_jump_table(const uint32_t* s, size_t n, uint32_t tile, uint64_t* out) {
    if (n == 0) {
        if (out) out[0] = 0;
        return 1;
    }
    size_t max_tile = s[n - 1] / tile;
    if (out) {
        size_t i = 0;
        for (size_t t = 0; t <= max_tile; t++) {
            uint64_t bound = (uint64_t)t * tile;
            while (i < n && s[i] < bound) i++;
            out[t] = i;
        }
    }
    return max_tile + 1;
}

/* ------------------------------------------------------- layered index */

typedef struct {
    uint32_t* start;
    uint32_t* end;
    uint32_t* sid;
    size_t n;
    uint64_t* jt;
    size_t njt;
    uint32_t max_size, tile;
} orc_layer;

struct orc_layered_db {
    uint32_t nshards, nlayers, nsources, min_size, gf;
    orc_layer* L; /* [shard * nlayers + layer] */
};

/* Stable LSD radix sort of 64-bit words by their high 32 bits. */
static void sort_by_hi32(uint64_t* a, uint64_t* tmp, size_t n) {
    size_t* cnt = (size_t*)malloc(sizeof(size_t) * 65536);
    for (int pass = 0; pass < 2; pass++) {
        int shift = 32 + 16 * pass;
        memset(cnt, 0, sizeof(size_t) * 65536);
        for (size_t i = 0; i < n; i++) cnt[(a[i] >> shift) & 0xFFFF]++;
        size_t acc = 0;
        for (size_t d = 0; d < 65536; d++) {
            size_t c = cnt[d];
            cnt[d] = acc;
            acc += c;
        }
        for (size_t i = 0; i < n; i++) tmp[cnt[(a[i] >> shift) & 0xFFFF]++] = a[i];
        memcpy(a, tmp, n * sizeof(uint64_t));
    }
    free(cnt);
}

typedef struct {
    orc_layered_db* db;
    const uint32_t *start, *end, *sid;
    size_t** bucket_idx; /* per bucket: input indices in input order */
    int next;            /* shared work counter */
    pthread_mutex_t mu;
} build_ctx;

static void* build_worker(void* p) {
    build_ctx* c = (build_ctx*)p;
    size_t nb = (size_t)c->db->nshards * c->db->nlayers;
    for (;;) {
        pthread_mutex_lock(&c->mu);
        int b = c->next++;
        pthread_mutex_unlock(&c->mu);
        if ((size_t)b >= nb) break;
        orc_layer* L = &c->db->L[b];
        size_t n = L->n;
        if (n == 0) continue;
        uint64_t* key = (uint64_t*)malloc(n * sizeof(uint64_t));
        uint64_t* tmp = (uint64_t*)malloc(n * sizeof(uint64_t));
        for (size_t i = 0; i < n; i++)
            key[i] = ((uint64_t)c->start[c->bucket_idx[b][i]] << 32) | (uint64_t)i;
        sort_by_hi32(key, tmp, n);
        for (size_t i = 0; i < n; i++) {
            size_t src = c->bucket_idx[b][key[i] & 0xFFFFFFFFu];
            L->start[i] = c->start[src];
            L->end[i] = c->end[src];
            L->sid[i] = c->sid[src];
        }
        free(key);
        free(tmp);
        L->njt = orc_build_jump_table(L->start, n, L->tile, NULL);
        L->jt = (uint64_t*)malloc(L->njt * sizeof(uint64_t));
        orc_build_jump_table(L->start, n, L->tile, L->jt);
    }
    return NULL;
}

orc_layered_db* orc_layered_build(size_t n, const uint32_t* shard, const uint32_t* start,
                                  const uint32_t* end, const uint32_t* sid, uint32_t nshards,
                                  uint32_t nsources, uint32_t min_size, uint32_t gf,
                                  int nthreads) {
    orc_layered_db* db = (orc_layered_db*)calloc(1, sizeof *db);
    db->nshards = nshards;
    db->nsources = nsources;
    db->min_size = min_size;
    db->gf = gf;
    uint32_t maxl = 0;
    uint32_t* lay = (uint32_t*)malloc((n ? n : 1) * sizeof(uint32_t));
    for (size_t i = 0; i < n; i++) {
        lay[i] = orc_layer_for_size(min_size, gf, end[i] - start[i]);
        if (lay[i] > maxl) maxl = lay[i];
    }
    db->nlayers = n ? maxl + 1 : 0; /* meta.num_layers (build.rs:258-275) */
    size_t nb = (size_t)nshards * db->nlayers;
    db->L = (orc_layer*)calloc(nb ? nb : 1, sizeof(orc_layer));
    for (size_t i = 0; i < n; i++) db->L[(size_t)shard[i] * db->nlayers + lay[i]].n++;
    size_t** idx = (size_t**)calloc(nb ? nb : 1, sizeof(size_t*));
    size_t* fill = (size_t*)calloc(nb ? nb : 1, sizeof(size_t));
    for (size_t b = 0; b < nb; b++) {
        orc_layer* L = &db->L[b];
        uint32_t layer = (uint32_t)(b % db->nlayers);
        L->max_size = orc_layer_max_size(min_size, gf, layer);
        L->tile = orc_tile_size(min_size, gf, layer);
        size_t m = L->n ? L->n : 1;
        L->start = (uint32_t*)malloc(m * 4);
        L->end = (uint32_t*)malloc(m * 4);
        L->sid = (uint32_t*)malloc(m * 4);
        idx[b] = (size_t*)malloc(m * sizeof(size_t));
    }
    for (size_t i = 0; i < n; i++) {
        size_t b = (size_t)shard[i] * db->nlayers + lay[i];
        idx[b][fill[b]++] = i;
    }
    build_ctx c = {db, start, end, sid, idx, 0, PTHREAD_MUTEX_INITIALIZER};
    if (nthreads < 1) nthreads = 1;
    pthread_t* th = (pthread_t*)malloc(sizeof(pthread_t) * (size_t)nthreads);
    for (int t = 0; t < nthreads; t++) pthread_create(&th[t], NULL, build_worker, &c);
    for (int t = 0; t < nthreads; t++) pthread_join(th[t], NULL);
    free(th);
    for (size_t b = 0; b < nb; b++) free(idx[b]);
    free(idx);
    free(fill);
    free(lay);
    return db;
}

void orc_layered_free(orc_layered_db* db) {
    if (!db) return;
    size_t nb = (size_t)db->nshards * db->nlayers;
    for (size_t b = 0; b < nb; b++) {
        free(db->L[b].start);
        free(db->L[b].end);
        free(db->L[b].sid);
        free(db->L[b].jt);
    }
    free(db->L);
    free(db);
}

size_t orc_layered_layer_len(const orc_layered_db* db, uint32_t shard, uint32_t layer) {
    if (shard >= db->nshards || layer >= db->nlayers) return 0;
    return db->L[(size_t)shard * db->nlayers + layer].n;
}

uint32_t orc_layered_num_layers(const orc_layered_db* db) { return db->nlayers; }

/* ------------------------------------------------------- windowed sweep */

/* MappedJumpTable::lookup (src/io/jump.rs:46-51). */
static inline size_t jt_lookup(const orc_layer* L, uint32_t target) {
    size_t t = target / L->tile;
    if (t > L->njt - 1) t = L->njt - 1;
    return (size_t)L->jt[t];
}

/* query_sweep_windowed (src/sweep/query.rs:84-117).  With cnt64 set, the
 * same scan accumulates exact int64 counts and overlap bp instead of the
 * reference's u32 counts and f32 bins. */
static void sweep_windowed(const orc_layer* L, const uint32_t* qs, const uint32_t* qe,
                           const uint32_t* qsid, size_t nq, uint32_t base, uint32_t F,
                           uint32_t* cnt, float* bins, int64_t* cnt64, int64_t* bp64) {
    if (nq == 0 || L->n == 0) return;
    for (size_t k = 0; k < nq; k++) {
        uint32_t lo = qs[k] > L->max_size ? qs[k] - L->max_size : 0;
        size_t a = jt_lookup(L, lo);
        if (a > L->n) a = L->n;
        while (a < L->n && L->start[a] < lo) a++; /* partition_point */
        size_t row = (size_t)(qsid[k] - base) * F;
        for (size_t i = a; i < L->n; i++) {
            uint32_t ds = L->start[i];
            if (ds >= qe[k]) break;
            uint32_t de = L->end[i];
            if (de > qs[k]) {
                uint32_t d = L->sid[i];
                uint32_t hi = qe[k] < de ? qe[k] : de;
                uint32_t lo2 = qs[k] > ds ? qs[k] : ds;
                if (cnt64) {
                    cnt64[row + d] += 1;
                    bp64[row + d] += (int64_t)(hi - lo2);
                } else {
                    cnt[row + d] += 1;
                    bins[row + d] += (float)(hi - lo2) / (float)ORC_BIN_SIZE;
                }
            }
        }
    }
}

typedef struct {
    const orc_layered_db* db;
    uint32_t shard;
    const uint32_t *qs, *qe, *qsid; /* this shard's queries sorted by (sid, start) */
    size_t nq;
    uint32_t nrows;
    uint32_t* counts;
    float* bins;
    int64_t *counts64, *bp64; /* exact form when set */
    uint32_t nwin;
    int next;
    pthread_mutex_t mu;
} query_ctx;

static size_t lower_sid(const uint32_t* sid, size_t n, uint32_t v) {
    size_t lo = 0, hi = n;
    while (lo < hi) {
        size_t m = (lo + hi) / 2;
        if (sid[m] < v) lo = m + 1; else hi = m;
    }
    return lo;
}

#define SWEEP_WINDOW_SIZE 64u /* src/tasks/query/mod.rs:23 */

static void* query_worker(void* p) {
    query_ctx* c = (query_ctx*)p;
    uint32_t F = c->db->nsources;
    const int exact = c->counts64 != NULL;
    uint32_t* sc = (uint32_t*)malloc((size_t)SWEEP_WINDOW_SIZE * F * 4);
    float* sb = (float*)malloc((size_t)SWEEP_WINDOW_SIZE * F * 4);
    int64_t* xc = exact ? (int64_t*)malloc((size_t)SWEEP_WINDOW_SIZE * F * 8) : NULL;
    int64_t* xb = exact ? (int64_t*)malloc((size_t)SWEEP_WINDOW_SIZE * F * 8) : NULL;
    for (;;) {
        pthread_mutex_lock(&c->mu);
        int w = c->next++;
        pthread_mutex_unlock(&c->mu);
        if ((uint32_t)w >= c->nwin) break;
        uint32_t lo_sid = (uint32_t)w * SWEEP_WINDOW_SIZE;
        uint32_t hi_sid = lo_sid + SWEEP_WINDOW_SIZE < c->nrows ? lo_sid + SWEEP_WINDOW_SIZE : c->nrows;
        uint32_t rows = hi_sid - lo_sid;
        size_t a = lower_sid(c->qsid, c->nq, lo_sid), b = lower_sid(c->qsid, c->nq, hi_sid);
        memset(sc, 0, (size_t)rows * F * 4);
        memset(sb, 0, (size_t)rows * F * 4);
        if (exact) {
            memset(xc, 0, (size_t)rows * F * 8);
            memset(xb, 0, (size_t)rows * F * 8);
        }
        for (uint32_t l = 0; l < c->db->nlayers; l++) {
            const orc_layer* L = &c->db->L[(size_t)c->shard * c->db->nlayers + l];
            sweep_windowed(L, c->qs + a, c->qe + a, c->qsid + a, b - a, lo_sid, F, sc, sb, xc, xb);
        }
        /* serial merge (mod.rs:281-285): windows own disjoint row blocks */
        for (size_t i = 0; i < (size_t)rows * F; i++) {
            if (exact) {
                c->counts64[(size_t)lo_sid * F + i] += xc[i];
                c->bp64[(size_t)lo_sid * F + i] += xb[i];
            } else {
                c->counts[(size_t)lo_sid * F + i] += sc[i];
                c->bins[(size_t)lo_sid * F + i] += sb[i];
            }
        }
    }
    free(sc);
    free(sb);
    free(xc);
    free(xb);
    return NULL;
}

static int layered_query(const orc_layered_db* db, size_t nquery, const uint32_t* shard,
                         const uint32_t* start, const uint32_t* end, const uint32_t* qsid,
                         uint32_t nq, uint32_t* counts, float* bins, int64_t* counts64,
                         int64_t* bp64, int nthreads) {
    uint32_t F = db->nsources;
    if (counts64) {
        memset(counts64, 0, (size_t)nq * F * 8);
        memset(bp64, 0, (size_t)nq * F * 8);
    } else {
        memset(counts, 0, (size_t)nq * F * 4);
        memset(bins, 0, (size_t)nq * F * 4);
    }
    uint32_t nwin = (nq + SWEEP_WINDOW_SIZE - 1) / SWEEP_WINDOW_SIZE;
    if (nthreads < 1) nthreads = 1;
    int used = nthreads < (int)nwin ? nthreads : (int)nwin;
    if (used < 1) used = 1;
    /* group queries by shard; sort each group by (sid, start) (mod.rs:202) */
    size_t* cnt = (size_t*)calloc(db->nshards + 1, sizeof(size_t));
    for (size_t i = 0; i < nquery; i++)
        if (shard[i] < db->nshards) cnt[shard[i] + 1]++;
    for (uint32_t s = 0; s < db->nshards; s++) cnt[s + 1] += cnt[s];
    size_t tot = cnt[db->nshards];
    uint64_t* key = (uint64_t*)malloc((tot ? tot : 1) * sizeof(uint64_t));
    uint64_t* tmp = (uint64_t*)malloc((tot ? tot : 1) * sizeof(uint64_t));
    uint32_t* idx = (uint32_t*)malloc((tot ? tot : 1) * 4);
    size_t* fill = (size_t*)calloc(db->nshards, sizeof(size_t));
    for (size_t i = 0; i < nquery; i++)
        if (shard[i] < db->nshards) idx[cnt[shard[i]] + fill[shard[i]]++] = (uint32_t)i;
    uint32_t* qs = (uint32_t*)malloc((tot ? tot : 1) * 4);
    uint32_t* qe = (uint32_t*)malloc((tot ? tot : 1) * 4);
    uint32_t* qd = (uint32_t*)malloc((tot ? tot : 1) * 4);
    for (uint32_t s = 0; s < db->nshards; s++) {
        size_t a = cnt[s], n = cnt[s + 1] - a;
        if (n == 0) continue;
        int any_layer = 0;
        for (uint32_t l = 0; l < db->nlayers; l++)
            if (db->L[(size_t)s * db->nlayers + l].n) any_layer = 1;
        if (!any_layer) continue; /* shard absent from the DB (mod.rs:193-199, 234) */
        /* two stable passes: by start, then by sid */
        for (size_t k = 0; k < n; k++) key[k] = ((uint64_t)start[idx[a + k]] << 32) | k;
        sort_by_hi32(key, tmp, n);
        for (size_t k = 0; k < n; k++) tmp[k] = ((uint64_t)qsid[idx[a + (key[k] & 0xFFFFFFFFu)]] << 32) | (key[k] & 0xFFFFFFFFu);
        memcpy(key, tmp, n * 8);
        sort_by_hi32(key, tmp, n);
        for (size_t k = 0; k < n; k++) {
            uint32_t src = idx[a + (key[k] & 0xFFFFFFFFu)];
            qs[k] = start[src];
            qe[k] = end[src];
            qd[k] = qsid[src];
        }
        query_ctx c = {db, s, qs, qe, qd, n, nq, counts, bins, counts64, bp64, nwin, 0,
                       PTHREAD_MUTEX_INITIALIZER};
        pthread_t* th = (pthread_t*)malloc(sizeof(pthread_t) * (size_t)used);
        for (int t = 0; t < used; t++) pthread_create(&th[t], NULL, query_worker, &c);
        for (int t = 0; t < used; t++) pthread_join(th[t], NULL);
        free(th);
    }
    free(cnt); free(key); free(tmp); free(idx); free(fill); free(qs); free(qe); free(qd);
    return used;
}

int orc_layered_query(const orc_layered_db* db, size_t nquery, const uint32_t* shard,
                      const uint32_t* start, const uint32_t* end, const uint32_t* qsid,
                      uint32_t nq, uint32_t* counts, float* bins, int nthreads) {
    return layered_query(db, nquery, shard, start, end, qsid, nq, counts, bins, NULL, NULL,
                         nthreads);
}

int orc_layered_query_exact(const orc_layered_db* db, size_t nquery, const uint32_t* shard,
                            const uint32_t* start, const uint32_t* end, const uint32_t* qsid,
                            uint32_t nq, int64_t* counts, int64_t* bp, int nthreads) {
    return layered_query(db, nquery, shard, start, end, qsid, nq, NULL, NULL, counts, bp,
                         nthreads);
}

/* ------------------------------------------------------------ exact forms */

void orc_sweep_bruteforce(size_t ndb, const uint32_t* dshard, const uint32_t* dstart,
                          const uint32_t* dend, const uint32_t* dsid, uint32_t F,
                          size_t nquery, const uint32_t* qshard, const uint32_t* qstart,
                          const uint32_t* qend, const uint32_t* qsid, uint32_t nq,
                          int64_t* counts, int64_t* bp) {
    memset(counts, 0, (size_t)nq * F * 8);
    memset(bp, 0, (size_t)nq * F * 8);
    for (size_t k = 0; k < nquery; k++)
        for (size_t i = 0; i < ndb; i++) {
            if (dshard[i] != qshard[k]) continue;
            if (dstart[i] < qend[k] && dend[i] > qstart[k]) {
                uint32_t hi = qend[k] < dend[i] ? qend[k] : dend[i];
                uint32_t lo = qstart[k] > dstart[i] ? qstart[k] : dstart[i];
                counts[(size_t)qsid[k] * F + dsid[i]] += 1;
                bp[(size_t)qsid[k] * F + dsid[i]] += (int64_t)(hi - lo);
            }
        }
}

static int cmp_u32(const void* a, const void* b) {
    uint32_t x = *(const uint32_t*)a, y = *(const uint32_t*)b;
    return x < y ? -1 : x > y;
}

/* number of v[i] < x (strict) or <= x (le) in sorted v[0..n) */
static size_t rank_u32(const uint32_t* v, size_t n, uint32_t x, int le) {
    size_t lo = 0, hi = n;
    while (lo < hi) {
        size_t m = (lo + hi) / 2;
        if (le ? v[m] <= x : v[m] < x) lo = m + 1; else hi = m;
    }
    return lo;
}

typedef struct {
    size_t ndb;
    const uint32_t *dshard, *dstart, *dend, *dsid;
    uint32_t F, nq, nshards;
    const size_t* off;
    const uint32_t *S, *E;
    const int64_t *PS, *PE;
    int64_t *counts, *bp;
    size_t next;
    pthread_mutex_t mu;
} ranks_ctx;

#define RANKS_BLOCK ((size_t)1 << 16)

static void* ranks_worker(void* p) {
    ranks_ctx* c = (ranks_ctx*)p;
    const uint32_t F = c->F, nq = c->nq;
    int64_t* cnt = (int64_t*)calloc((size_t)nq * F, 8);
    int64_t* bp = (int64_t*)calloc((size_t)nq * F, 8);
    for (;;) {
        pthread_mutex_lock(&c->mu);
        size_t i0 = c->next;
        c->next += RANKS_BLOCK;
        pthread_mutex_unlock(&c->mu);
        if (i0 >= c->ndb) break;
        size_t i1 = i0 + RANKS_BLOCK < c->ndb ? i0 + RANKS_BLOCK : c->ndb;
        for (size_t i = i0; i < i1; i++) {
            if (c->dshard[i] >= c->nshards) continue;
            int64_t x0 = c->dstart[i], x1 = c->dend[i];
            for (uint32_t q = 0; q < nq; q++) {
                size_t g = (size_t)c->dshard[i] * nq + q;
                size_t a = c->off[g], n = c->off[g + 1] - a;
                if (n == 0) continue;
                const uint32_t *s = c->S + a, *e = c->E + a;
                const int64_t *ps = c->PS + a + g, *pe = c->PE + a + g;
                size_t rs1 = rank_u32(s, n, (uint32_t)x1, 0), re1 = rank_u32(e, n, (uint32_t)x1, 0);
                size_t rs0 = rank_u32(s, n, (uint32_t)x0, 0), re0 = rank_u32(e, n, (uint32_t)x0, 0);
                size_t re0le = rank_u32(e, n, (uint32_t)x0, 1);
                int64_t cq1 = x1 * (int64_t)rs1 - ps[rs1] - x1 * (int64_t)re1 + pe[re1];
                int64_t cq0 = x0 * (int64_t)rs0 - ps[rs0] - x0 * (int64_t)re0 + pe[re0];
                cnt[(size_t)q * F + c->dsid[i]] += (int64_t)rs1 - (int64_t)re0le;
                bp[(size_t)q * F + c->dsid[i]] += cq1 - cq0;
            }
        }
    }
    pthread_mutex_lock(&c->mu);
    for (size_t k = 0; k < (size_t)nq * F; k++) {
        c->counts[k] += cnt[k];
        c->bp[k] += bp[k];
    }
    pthread_mutex_unlock(&c->mu);
    free(cnt);
    free(bp);
    return NULL;
}

void orc_sweep_ranks(size_t ndb, const uint32_t* dshard, const uint32_t* dstart,
                     const uint32_t* dend, const uint32_t* dsid, uint32_t F, size_t nquery,
                     const uint32_t* qshard, const uint32_t* qstart, const uint32_t* qend,
                     const uint32_t* qsid, uint32_t nq, uint32_t nshards, int64_t* counts,
                     int64_t* bp) {
    memset(counts, 0, (size_t)nq * F * 8);
    memset(bp, 0, (size_t)nq * F * 8);
    size_t ng = (size_t)nshards * nq;
    size_t* off = (size_t*)calloc(ng + 1, sizeof(size_t));
    for (size_t k = 0; k < nquery; k++) off[(size_t)qshard[k] * nq + qsid[k] + 1]++;
    for (size_t g = 0; g < ng; g++) off[g + 1] += off[g];
    uint32_t* S = (uint32_t*)malloc((nquery ? nquery : 1) * 4);
    uint32_t* E = (uint32_t*)malloc((nquery ? nquery : 1) * 4);
    size_t* fill = (size_t*)calloc(ng ? ng : 1, sizeof(size_t));
    for (size_t k = 0; k < nquery; k++) {
        size_t g = (size_t)qshard[k] * nq + qsid[k];
        S[off[g] + fill[g]] = qstart[k];
        E[off[g] + fill[g]] = qend[k];
        fill[g]++;
    }
    int64_t* PS = (int64_t*)malloc(((nquery ? nquery : 1) + ng + 1) * 8);
    int64_t* PE = (int64_t*)malloc(((nquery ? nquery : 1) + ng + 1) * 8);
    /* prefix arrays: group g occupies PS[off[g] + g .. off[g+1] + g] */
    for (size_t g = 0; g < ng; g++) {
        size_t a = off[g], n = off[g + 1] - a;
        qsort(S + a, n, 4, cmp_u32);
        qsort(E + a, n, 4, cmp_u32);
        int64_t ps = 0, pe = 0;
        PS[a + g] = 0;
        PE[a + g] = 0;
        for (size_t k = 0; k < n; k++) {
            ps += S[a + k];
            pe += E[a + k];
            PS[a + g + k + 1] = ps;
            PE[a + g + k + 1] = pe;
        }
    }
    ranks_ctx c = {ndb, dshard, dstart, dend, dsid, F, nq, nshards, off, S, E, PS, PE, counts, bp,
                   0, PTHREAD_MUTEX_INITIALIZER};
    /* DB elements in blocks over up to 16 threads, each into a private matrix
     * merged under the lock: integer sums, so the result is independent of
     * the split */
    long nt = sysconf(_SC_NPROCESSORS_ONLN);
    int T = nt < 1 ? 1 : (nt > 16 ? 16 : (int)nt);
    if ((size_t)T > ndb / 65536 + 1) T = (int)(ndb / 65536 + 1);
    pthread_t th[16];
    for (int t = 0; t < T; t++) pthread_create(&th[t], NULL, ranks_worker, &c);
    for (int t = 0; t < T; t++) pthread_join(th[t], NULL);
    free(off); free(S); free(E); free(fill); free(PS); free(PE);
}
