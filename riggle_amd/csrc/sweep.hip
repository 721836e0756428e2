// sweep.hip -- overlap sweep (count + overlap bp, exact u64), dense and sparse.
//
// Dense path.  For one query file Q and one DB interval d on a shard (pair-free
// form of the reference's per-pair scan, src/sweep/query.rs:84-117):
//   count(Q, d) = #{q : q.s < d.e} - #{q : q.e <= d.s}
//   bp(Q, d)    = CQ(d.e) - CQ(d.s),   CQ(x) = sum_q |[q.s, q.e) n [0, x)|
// With the query's sorted event stream, a DB coordinate x of rank p (events
// ordered before it) has rS = #starts, rE = p - rS, depth = rS - rE and
// CQ(x) = x * depth - W.  So each element of the DB end stream adds
// (rS, CQ) to cell (Q, fid) and each element of the start stream subtracts
// (rE, CQ).  u64 wrap-around makes every partial sum exact.
//
// Work is cut into equal tiles of TILE merged items (DB elements + query
// events) along the merge path of each (DB stream, query segment) job, so a
// tile costs the same wherever the query's events cluster.  Persistent blocks
// take chunks of consecutive tiles from a device work queue and
// software-pipeline them: while tile t is ranked and reduced, the DB elements
// and query events of tile t+1 are already in flight into registers.  A wave
// owns 64 * ITEMS consecutive DB elements of a tile; it holds the 64 events
// following its first element in registers and ranks each 64-element row with
// two ballots (events below the row's first and last key): rows without an
// event inside share one rank, the others count the few events inside with
// cross-lane reads.  The row loop reads no LDS, so its atomics are only
// drained at the tile barrier.  Contributions are reduced per DB file in an LDS
// histogram that is flushed with one u64 atomic per touched cell when the
// block moves to another (query file, file group) or ends.
//
// Sparse path.  A small query file is swept the reference's way
// (src/sweep/query.rs:84-117): for each query interval and each size layer
// (of each DB file group), a jump-table lookup finds the first DB start that
// can overlap it and a wave scans the layer's starts up to q.end, adding
// every overlapping pair.  Its cost follows the query's pairs instead of the
// DB's size.
#include <algorithm>
#include <type_traits>

#include "engine.hpp"

namespace chk {

void query_events(Query& q, hipStream_t st);

namespace {

constexpr int NT = 256;
constexpr uint32_t QS = 256;    // events staged per tile on the fast path (nq + 1 <= QS)
constexpr int QH = QS / NT;     // staged events per thread
constexpr uint32_t MAXC = 64;   // tiles per work item at most
constexpr uint32_t QSTRIDE = 32;  // u32 words between the per-XCD queue counters (128 B)
constexpr uint32_t SPI = 256;   // query intervals per sparse work item at most (default)

__device__ inline uint32_t db_key(int32_t x) { return 2u * (uint32_t)x + 1u; }

// Block barrier that orders LDS traffic only: global loads issued for later
// tiles stay in flight across it (__syncthreads would also wait for them).
__device__ inline void lds_barrier() { asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory"); }

__device__ inline uint32_t ufl(uint32_t v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ inline int64_t ufl64(int64_t v) {
    uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
    uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)((uint64_t)v >> 32));
    return (int64_t)(((uint64_t)hi << 32) | lo);
}
__device__ inline uint32_t rdl(uint32_t v, uint32_t l) { return __builtin_amdgcn_readlane(v, l); }
__device__ inline int64_t rdl64(int64_t v, uint32_t l) {
    uint32_t lo = __builtin_amdgcn_readlane((uint32_t)v, l);
    uint32_t hi = __builtin_amdgcn_readlane((uint32_t)((uint64_t)v >> 32), l);
    return (int64_t)(((uint64_t)hi << 32) | lo);
}
__device__ inline uint32_t popc(uint64_t m) { return (uint32_t)__builtin_popcountll(m); }
// Copy of a wave-uniform value held in a vector register, so that a 64-bit
// multiply-add can take it as its addend (an instruction reads at most one
// scalar register) without a copy per use.
__device__ inline uint64_t vcopy64(uint64_t s) {
    uint32_t lo, hi;
    asm volatile("v_mov_b32 %0, %1" : "=v"(lo) : "s"((uint32_t)s));
    asm volatile("v_mov_b32 %0, %1" : "=v"(hi) : "s"((uint32_t)(s >> 32)));
    return ((uint64_t)hi << 32) | lo;
}

constexpr int BUF_DWORD3 = 0x00020000;  // raw buffer, gfx9-family descriptor word 3

__device__ inline __amdgpu_buffer_rsrc_t make_rsrc(const void* base, uint32_t bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0, (int)bytes,
                                             BUF_DWORD3);
}

// Narrows the merge-path split of diagonal d, known to lie in [a, z], with the
// stream's samples S[g] = stream[g * stride] (global stream positions; the
// job's stream starts at db_off): the split is after position p iff
// db_key(D[p]) <= Q[d - 1 - p].
__device__ inline void narrow_split(const int32_t* S, uint64_t stride, uint64_t db_off,
                                    const uint32_t* Q, uint64_t d, uint64_t& a, uint64_t& z) {
    const uint64_t g0 = (db_off + a + stride - 1) / stride, g1 = (db_off + z) / stride;
    uint64_t lo = g0, hi = g1;  // first sample where the split is at or before it
    while (lo < hi) {
        const uint64_t mid = (lo + hi) >> 1;
        const uint64_t p = mid * stride - db_off;
        if (db_key(S[mid]) <= Q[d - 1 - p]) lo = mid + 1; else hi = mid;
    }
    if (lo > g0) a = (lo - 1) * stride - db_off + 1;
    if (lo < g1) z = lo * stride - db_off;
}

// Merge-path split of every tile boundary: part[b] = # DB elements among the
// first d merged items of its job (DB element before event iff dkey <= qkey).
// The search narrows the split to PSTRIDE2 elements on the coarse samples
// (every PSTRIDE2-th stream element: small enough to stay cached), then to
// PSTRIDE on the fine ones, then finishes on the stream itself.
__global__ void partition_kernel(const Job* jobs, uint32_t njobs, uint32_t nbound, uint32_t tile,
                                 const int32_t* s_coord, const int32_t* e_coord,
                                 const int32_t* s_samp, const int32_t* e_samp,
                                 const int32_t* s_samp2, const int32_t* e_samp2,
                                 const uint32_t* qkeys, uint32_t* part) {
    uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nbound) return;
    uint32_t lo = 0, hi = njobs;  // last job with part0 <= b
    while (hi - lo > 1) {
        uint32_t mid = (lo + hi) >> 1;
        if (jobs[mid].part0 <= b) lo = mid; else hi = mid;
    }
    const Job J = jobs[lo];
    const int32_t* D = (J.is_end ? e_coord : s_coord) + J.db_off;
    const uint32_t* Q = qkeys + J.q_off;
    uint64_t total = (uint64_t)J.n_db + J.m;
    uint64_t d = std::min<uint64_t>((uint64_t)(b - J.part0) * tile, total);
    uint64_t a = d > J.m ? d - J.m : 0, z = std::min<uint64_t>(d, J.n_db);
    if (z - a > 2 * PSTRIDE2)
        narrow_split(J.is_end ? e_samp2 : s_samp2, PSTRIDE2, J.db_off, Q, d, a, z);
    if (z - a > 2 * PSTRIDE)
        narrow_split(J.is_end ? e_samp : s_samp, PSTRIDE, J.db_off, Q, d, a, z);
    while (a < z) {
        uint64_t mid = (a + z) >> 1;
        if (db_key(D[mid]) <= Q[d - 1 - mid]) a = mid + 1; else z = mid;
    }
    part[b] = (uint32_t)a;
}

struct SweepArgs {
    const Job* jobs;
    const uint32_t* chunk_job;  // nchunks + 1
    const uint32_t* part;
    uint32_t ntiles, nchunks, chunk, F;
    const int32_t* s_coord;
    const int32_t* e_coord;
    const uint16_t* s_lfid;
    const uint16_t* e_lfid;
    const uint32_t* qkey;  // 2n+1 entries (sentinel at the end)
    const uint32_t* qrs;   // 2n+1
    const int64_t* qw;     // 2n+1
    uint32_t* queue;
    unsigned long long* counts;
    unsigned long long* bp;
};

// One tile as the block sees it.
struct TMeta {
    uint64_t dbase;  // first DB element of the tile (stream index)
    uint64_t cell;   // output cell of group-local file 0
    int64_t w0;      // W at the job's first event
    uint32_t nd, nq, qb, j0, rs0, is_end, nf, pad;
};

struct Meta {
    uint64_t dbase, cell;
    int64_t w0;
    uint32_t nd, nq, qb, j0, rs0, is_end, nf;
};

// Registers holding one tile in flight.  Buffer loads bound every read by
// the tile's extent in hardware (lanes past it read 0), so loads need no
// per-lane clamping and are issued unconditionally: the compiler then counts
// outstanding loads statically and keeps the next tile's loads in flight
// while this one is reduced.
template <int ITEMS>
struct Stage {
    int32_t x[ITEMS];
    uint32_t f[ITEMS];
    uint32_t key[QH], rs[QH];
    int64_t w[QH];
};

// 8-item tiles fit 96 VGPRs (5 waves per SIMD) without spills, 12-item tiles
// 128 (4 waves); 16-item tiles need more registers than that
template <int ITEMS>
__global__ __launch_bounds__(NT, ITEMS <= 8 ? 5 : (ITEMS <= 12 ? 4 : 1)) void sweep_kernel(SweepArgs a) {
    constexpr uint32_t IPW = 64u * ITEMS;  // DB elements per wave per tile
    constexpr uint32_t TILE = NT * ITEMS;
    __shared__ unsigned long long hc[FG], hb[FG];
    // staged query events, double-buffered (register stage A uses buffer 0,
    // B buffer 1): a tile's events are written while the other buffer may
    // still be read, so one barrier per tile orders the staging
    __shared__ uint32_t lkey2[2][QS], lrs2[2][QS];
    __shared__ int64_t lw2[2][QS];
    __shared__ TMeta tm[MAXC];
    __shared__ Job lj[MAXC + 1];
    __shared__ uint32_t s_chunk;
    const uint32_t tid = threadIdx.x;
    const uint32_t lane = tid & 63u;
    const uint32_t wv = ufl(tid >> 6);  // wave index (scalar)
    const uint32_t F = a.F;
    for (uint32_t f = tid; f < FG; f += NT) hc[f] = hb[f] = 0ull;
    uint64_t cur_cell = ~0ull;
    uint32_t cur_nf = 0;

    auto flush = [&]() {
        if (cur_cell == ~0ull) return;
        for (uint32_t f = tid; f < cur_nf; f += NT) {
            const unsigned long long cv = hc[f], bv = hb[f];
            if (cv | bv) {
                atomicAdd(&a.counts[cur_cell + f], cv);
                atomicAdd(&a.bp[cur_cell + f], bv);
                hc[f] = hb[f] = 0ull;
            }
        }
    };

    auto meta = [&](uint32_t k) {
        Meta m;
        m.dbase = (uint64_t)ufl64((int64_t)tm[k].dbase);
        m.cell = (uint64_t)ufl64((int64_t)tm[k].cell);
        m.w0 = ufl64(tm[k].w0);
        m.nd = ufl(tm[k].nd);
        m.nq = ufl(tm[k].nq);
        m.qb = ufl(tm[k].qb);
        m.j0 = ufl(tm[k].j0);
        m.rs0 = ufl(tm[k].rs0);
        m.is_end = ufl(tm[k].is_end);
        m.nf = ufl(tm[k].nf);
        return m;
    };

    auto stage_load = [&](const Meta& m, Stage<ITEMS>& s) {
        // query events first: staging them into LDS then waits only for these
        // loads, leaving the DB loads in flight across the barrier
        const uint32_t nst = m.nq + 1 < QS ? m.nq + 1 : QS;  // ranks lie in [0, nq]
        const auto rk = make_rsrc(a.qkey + m.qb, nst * 4u);
        const auto rr = make_rsrc(a.qrs + m.qb, nst * 4u);
        const auto rw = make_rsrc(a.qw + m.qb, nst * 8u);
#pragma unroll
        for (int h = 0; h < QH; h++) {
            const uint32_t i = tid + h * NT;
            s.key[h] = __builtin_amdgcn_raw_buffer_load_b32(rk, i * 4u, 0, 0);
            s.rs[h] = __builtin_amdgcn_raw_buffer_load_b32(rr, i * 4u, 0, 0);
            s.w[h] = (int64_t)__builtin_amdgcn_raw_buffer_load_b64(rw, i * 8u, 0, 0);
        }
        const int32_t* coord = (m.is_end ? a.e_coord : a.s_coord) + m.dbase;
        const uint16_t* fid = (m.is_end ? a.e_lfid : a.s_lfid) + m.dbase;
        const auto rc = make_rsrc(coord, m.nd * 4u);
        const auto rf = make_rsrc(fid, m.nd * 2u);
#pragma unroll
        for (int k = 0; k < ITEMS; k++) {
            const uint32_t i = wv * IPW + k * 64u + lane;
            s.x[k] = __builtin_amdgcn_raw_buffer_load_b32(rc, i * 4u, 0, 0);
            s.f[k] = __builtin_amdgcn_raw_buffer_load_b16(rf, i * 2u, 0, 0);
        }
    };

    auto stage_store = [&](const Stage<ITEMS>& s, const Meta& m, uint32_t* lkey, uint32_t* lrs,
                           int64_t* lw) {
        if (m.cell != cur_cell) {  // uniform over the block
            lds_barrier();         // every wave's atomics into the histogram are done
            flush();
            cur_cell = m.cell;
            cur_nf = m.nf;
        }
#pragma unroll
        for (int h = 0; h < QH; h++) {
            const uint32_t i = tid + h * NT;
            if (i <= m.nq) {  // i < QS always
                lkey[i] = i < m.nq ? s.key[h] : 0xFFFFFFFFu;
                lrs[i] = s.rs[h] - m.rs0;
                lw[i] = s.w[h] - m.w0;
            }
        }
    };

    // The streams carry each element's group-local file id times 8: the byte
    // offset of its u64 slot in hc and hb.
    auto hist_add = [&](uint32_t fo, unsigned long long cv, unsigned long long bv) {
        atomicAdd(reinterpret_cast<unsigned long long*>(reinterpret_cast<char*>(hc) + fo), cv);
        atomicAdd(reinterpret_cast<unsigned long long*>(reinterpret_cast<char*>(hb) + fo), bv);
    };

    // Per-element contribution at job rank P with prefix values (rs, w) at P.
    // E stream: (+rs, x*depth - w); S stream: (rs - P, -(x*depth - w));
    // depth = 2 rs - P.
    auto add = [&](const Meta& m, int32_t x, uint32_t f, uint32_t P, uint32_t rsv, int64_t wq) {
        const int32_t depth = 2 * (int32_t)rsv - (int32_t)P;
        const int32_t dep = m.is_end ? depth : -depth;
        const int64_t wn = m.is_end ? -wq : wq;
        const unsigned long long cv =
            (unsigned long long)(m.is_end ? (int64_t)rsv : (int64_t)rsv - (int64_t)P);
        const unsigned long long bv = (unsigned long long)((int64_t)x * (int64_t)dep + wn);
        hist_add(f, cv, bv);
    };

    // Contribution of one row at ranks P (per lane) with prefix values (rsv, wq),
    // for one stream kind (see add).
    auto add_t = [&](auto end_c, int32_t x, uint32_t f, uint32_t P, uint32_t rsv, int64_t wq) {
        constexpr bool IS_END = decltype(end_c)::value;
        const int32_t depth = 2 * (int32_t)rsv - (int32_t)P;
        const int32_t dep = IS_END ? depth : -depth;
        const int64_t wn = IS_END ? -wq : wq;
        const unsigned long long cv =
            (unsigned long long)(IS_END ? (int64_t)rsv : (int64_t)rsv - (int64_t)P);
        hist_add(f, cv, (unsigned long long)((int64_t)x * (int64_t)dep + wn));
    };

    // A full wave span (ITEMS rows of 64 elements) ranked against the register
    // window, for one stream kind: no per-row bounds checks, and in a row with
    // no event inside every lane has the same rank, so its depth, count and W
    // terms are scalar.
    // nvl = valid lanes of the last row (64 for a full span; the last wave of
    // a tile with events is short by their number)
    auto rows_full = [&](auto end_c, const Stage<ITEMS>& s, uint32_t p0, uint32_t kv, uint32_t rv,
                         int64_t wvv, uint32_t nvl) {
        // event key 2y + b is below DB key 2x + 1 iff y + b <= x: compare the
        // window's y + b with the coordinates directly (no key per element);
        // past the events, 0xFFFFFFFF maps to 2^31, above every coordinate
        const uint32_t kx = (kv >> 1) + (kv & 1u);
        constexpr bool IS_END = decltype(end_c)::value;
        uint32_t ce = 0xFFFFFFFFu, rsa = 0;  // prefix values at window event ce
        int64_t wa = 0;
        // the event-free rows' terms at rank ce: depth (signed for the stream
        // kind, scalar), count and W term (vector copies)
        int32_t dep_c = 0;
        uint64_t cv_c = 0, wn_c = 0;
        // a row's base rank is the events below its first element, or (after
        // the first row) the previous row's end rank: events that fall
        // between two rows are then counted by the row's event loop, where
        // every lane is past them
        uint32_t ea = popc(__ballot(kx <= ufl((uint32_t)s.x[0])));
#pragma unroll
        for (int k = 0; k < ITEMS; k++) {
            const bool last = k == ITEMS - 1;
            const uint32_t x = (uint32_t)s.x[k];
            const uint32_t eb = popc(__ballot(kx <= rdl(x, last ? nvl - 1u : 63u)));
            if (ea != ce) {  // rows without events keep the previous row's prefix
                ce = ea;
                rsa = rdl(rv, ea);
                wa = rdl64(wvv, ea);
                const int32_t depth = 2 * (int32_t)rsa - (int32_t)(p0 + ea);
                dep_c = IS_END ? depth : -depth;
                cv_c = vcopy64(IS_END ? (uint64_t)rsa : (uint64_t)((int64_t)rsa - (int64_t)(p0 + ea)));
                wn_c = vcopy64(IS_END ? (uint64_t)-wa : (uint64_t)wa);
            }
            if (ea == eb) {
                if (!last || lane < nvl)
                    hist_add(s.f[k], cv_c, (unsigned long long)((int64_t)s.x[k] * (int64_t)dep_c + (int64_t)wn_c));
            } else {
                // a lane's rank is ea + c, c = the row's events below it; its
                // prefix values are the window's at that rank, read across
                // lanes (ea + c <= eb < 64 on this path)
                uint32_t c = 0;
                for (uint32_t t = ea; t < eb; t++) c += rdl(kx, t) <= x ? 1u : 0u;
                const int src = (int)((ea + c) << 2);
                const uint32_t rsv = (uint32_t)__builtin_amdgcn_ds_bpermute(src, (int)rv);
                const uint32_t wlo = (uint32_t)__builtin_amdgcn_ds_bpermute(src, (int)(uint32_t)wvv);
                const uint32_t whi =
                    (uint32_t)__builtin_amdgcn_ds_bpermute(src, (int)(uint32_t)((uint64_t)wvv >> 32));
                const int64_t wq = (int64_t)(((uint64_t)whi << 32) | wlo);
                if (!last || lane < nvl) add_t(end_c, s.x[k], s.f[k], p0 + ea + c, rsv, wq);
            }
            ea = eb;
        }
    };

    // rank of key dk among the staged events: two-level ballot search
    auto rank_lds = [&](uint32_t dk, uint32_t nq, const uint32_t* lkey) {
        const uint32_t si = lane * 8u;
        const uint32_t sv = si < nq ? lkey[si] : 0xFFFFFFFFu;
        const uint32_t c1 = popc(__ballot(sv < dk));
        const uint32_t base = c1 ? (c1 - 1u) * 8u : 0u;
        const uint32_t li = base + lane;
        const uint32_t lv = (lane < 8u && li < nq) ? lkey[li] : 0xFFFFFFFFu;
        return base + popc(__ballot(lv < dk));
    };

    auto compute = [&](const Stage<ITEMS>& s, const Meta& m, const uint32_t* lkey, const uint32_t* lrs,
                       const int64_t* lw) {
        const uint32_t nq = m.nq, nd = m.nd;
        const uint32_t wbase = wv * IPW;
        if (wbase >= nd) return;
        if (nq + 1 <= QS) {
            uint32_t jlo = 0, kv = 0, rv = 0;
            int64_t wvv = 0;
            auto window = [&](uint32_t dk) {
                // with fewer than 64 events the window is all of them
                jlo = nq < 64u ? 0u : ufl(rank_lds(dk, nq, lkey));
                const uint32_t wi = jlo + lane;
                kv = wi < nq ? lkey[wi] : 0xFFFFFFFFu;
                rv = wi <= nq ? lrs[wi] : 0u;
                wvv = wi <= nq ? lw[wi] : 0;
            };
            window(ufl(db_key(s.x[0])));
            // key of the wave's last element: when it is below the window's
            // end, every row of the wave ranks inside the window
            const uint32_t nvw = nd - wbase < IPW ? nd - wbase : IPW;
            const uint32_t kl = (nvw - 1u) >> 6, ll = (nvw - 1u) & 63u;
            uint32_t dkw = 0;
            if (nvw > IPW - 64u) {
                dkw = rdl(db_key(s.x[ITEMS - 1]), nvw - (IPW - 64u) - 1u);
            } else {
#pragma unroll
                for (int k = 0; k < ITEMS; k++)
                    if ((uint32_t)k == kl) dkw = rdl(db_key(s.x[k]), ll);
            }
            if (popc(__ballot(kv < dkw)) < 64u && nvw > IPW - 64u) {
                const uint32_t nvl = nvw - (IPW - 64u);
                if (m.is_end)
                    rows_full(std::true_type{}, s, m.j0 + jlo, kv, rv, wvv, nvl);
                else
                    rows_full(std::false_type{}, s, m.j0 + jlo, kv, rv, wvv, nvl);
            } else if (popc(__ballot(kv < dkw)) < 64u) {
                // Row rank offset c in [ea, eb] and the prefix values at it,
                // all from the register window: no LDS read in this loop, so
                // the rows' atomics are never waited for inside the tile.
#pragma unroll
                for (int k = 0; k < ITEMS; k++) {
                    const uint32_t base = wbase + k * 64u;
                    if (base < nd) {
                        const uint32_t nv = nd - base < 64u ? nd - base : 64u;
                        const uint32_t dk = db_key(s.x[k]);
                        const uint32_t dkf = ufl(dk), dkl = rdl(dk, nv - 1u);
                        const uint32_t ea = popc(__ballot(kv < dkf)), eb = popc(__ballot(kv < dkl));
                        const uint32_t pa = m.j0 + jlo + ea;
                        const uint32_t rsa = rdl(rv, ea);
                        const int64_t wa = rdl64(wvv, ea);
                        if (ea == eb) {
                            // no event inside the row: every lane has rank pa,
                            // so the depth and count terms are scalar
                            if (lane < nv) add(m, s.x[k], s.f[k], pa, rsa, wa);
                        } else {
                            uint32_t c = 0, rsv = rsa;
                            int64_t wq = wa;
                            for (uint32_t t = ea; t < eb; t++) {
                                const bool past = rdl(kv, t) < dk;
                                c += past ? 1u : 0u;
                                const uint32_t r1 = rdl(rv, t + 1);
                                const int64_t w1 = rdl64(wvv, t + 1);
                                rsv = past ? r1 : rsv;
                                wq = past ? w1 : wq;
                            }
                            if (lane < nv) add(m, s.x[k], s.f[k], pa + c, rsv, wq);
                        }
                    }
                }
            } else {
                // more events in the wave's span than the window holds:
                // per-lane binary search over the staged events
#pragma unroll
                for (int k = 0; k < ITEMS; k++) {
                    const uint32_t i = wbase + k * 64u + lane;
                    if (i < nd) {
                        const uint32_t dk = db_key(s.x[k]);
                        uint32_t lo = 0, hi = nq;
                        while (lo < hi) {
                            const uint32_t mid = (lo + hi) >> 1;
                            if (lkey[mid] < dk) lo = mid + 1; else hi = mid;
                        }
                        add(m, s.x[k], s.f[k], m.j0 + lo, lrs[lo], lw[lo]);
                    }
                }
            }
        } else {
            // event-dense tile: per-lane binary search over the keys in memory
#pragma unroll
            for (int k = 0; k < ITEMS; k++) {
                const uint32_t i = wbase + k * 64u + lane;
                if (i < nd) {
                    const uint32_t dk = db_key(s.x[k]);
                    uint32_t lo = 0, hi = nq;
                    while (lo < hi) {
                        const uint32_t mid = (lo + hi) >> 1;
                        if (a.qkey[m.qb + mid] < dk) lo = mid + 1; else hi = mid;
                    }
                    add(m, s.x[k], s.f[k], m.j0 + lo, a.qrs[m.qb + lo] - m.rs0,
                        a.qw[m.qb + lo] - m.w0);
                }
            }
        }
    };

    // eight queues, one per XCD (block b runs on XCD b % 8): queue x hands
    // out chunks x, x + 8, ...; a block whose queue is spent moves on to the
    // next XCD's
    uint32_t xq = blockIdx.x & 7u;
    for (;;) {
        if (tid == 0) {
            uint32_t c = ~0u;
            for (int tries = 0; tries < 8; tries++) {
                const uint64_t cc = (uint64_t)atomicAdd(a.queue + xq * QSTRIDE, 1u) * 8u + xq;
                if (cc < a.nchunks) {
                    c = (uint32_t)cc;
                    break;
                }
                xq = (xq + 1u) & 7u;
            }
            s_chunk = c;
        }
        __syncthreads();
        const uint32_t c = ufl(s_chunk);
        if (c >= a.nchunks) break;
        const uint32_t t0 = c * a.chunk;
        const uint32_t nt = min(a.chunk, a.ntiles - t0);
        const uint32_t ja = a.chunk_job[c], jn = a.chunk_job[c + 1] - ja + 1;
        for (uint32_t j = tid; j < jn; j += NT) lj[j] = a.jobs[ja + j];
        __syncthreads();
        if (tid < nt) {
            const uint32_t t = t0 + tid;
            uint32_t k = 0;
            while (k + 1 < jn && lj[k + 1].tile0 <= t) k++;
            const Job J = lj[k];
            const uint32_t kt = t - J.tile0;
            const uint32_t pi = J.part0 + kt;
            const uint32_t i0 = a.part[pi], i1 = a.part[pi + 1];
            const uint64_t total = (uint64_t)J.n_db + J.m;
            const uint64_t d0 = (uint64_t)kt * TILE;
            const uint64_t d1 = d0 + TILE < total ? d0 + TILE : total;
            TMeta T;
            T.dbase = J.db_off + i0;
            T.nd = i1 - i0;
            T.j0 = (uint32_t)(d0 - i0);
            T.nq = (uint32_t)(d1 - i1) - T.j0;
            T.qb = J.q_off + T.j0;
            T.rs0 = a.qrs[J.q_off];
            T.w0 = a.qw[J.q_off];
            T.cell = (uint64_t)J.row * F + (uint64_t)J.group * FG;
            T.nf = min(FG, F - J.group * FG);
            T.is_end = J.is_end;
            T.pad = 0;
            tm[tid] = T;
        }
        __syncthreads();
        // two register stages: the loads of one tile are issued before the
        // other tile is staged and reduced; past the chunk's end the last
        // tile is reloaded (results unused) so every iteration issues the
        // same loads
        Stage<ITEMS> A, B;
        Meta mA = meta(0);
        stage_load(mA, A);
        Meta mB = meta(min(1u, nt - 1));
        for (uint32_t k = 0;; k += 2) {
            stage_load(mB, B);
            stage_store(A, mA, lkey2[0], lrs2[0], lw2[0]);
            lds_barrier();
            compute(A, mA, lkey2[0], lrs2[0], lw2[0]);
            if (k + 1 >= nt) break;
            mA = meta(min(k + 2, nt - 1));
            stage_load(mA, A);
            stage_store(B, mB, lkey2[1], lrs2[1], lw2[1]);
            lds_barrier();
            compute(B, mB, lkey2[1], lrs2[1], lw2[1]);
            if (k + 2 >= nt) break;
            mB = meta(min(k + 3, nt - 1));
        }
    }
    flush();
}

struct SparseArgs {
    const SparseItem* items;
    uint32_t nshards;  // intervals on shard ids >= nshards are skipped
    const uint32_t* shard;
    const uint32_t* start;
    const uint32_t* end;
    const LayerSeg* lseg;
    uint32_t ngroups, nlayers, F;
    const int32_t* l_start;
    const int32_t* l_end;
    const uint32_t* l_fid;
    const uint32_t* jt;
    unsigned long long* counts;
    unsigned long long* bp;
};

// One block per work item (query intervals of one file x one DB file group).
// The item's (query interval, size layer) pairs are dealt to the block's
// waves; per pair a wave reads both ends of its candidate range from the
// jump table (the first start of the group's layer in the bucket of
// q.s - lookback + 1, and the first in the bucket after q.e - 1's) and scans
// the range 128 starts a round, the next round's (start, end, file id)
// triples in flight while the current ones are tested.  Pairs are reduced in
// an LDS histogram of the group's files, flushed once per item.
__global__ __launch_bounds__(NT) void sparse_kernel(SparseArgs a) {
    __shared__ unsigned long long hc[FG], hb[FG];
    const uint32_t tid = threadIdx.x, lane = tid & 63u;
    const uint32_t wv = ufl(tid >> 6);
    const SparseItem it = a.items[blockIdx.x];
    const uint32_t g = it.group;
    const uint32_t fbase = g * FG;
    const uint32_t nf = min(FG, a.F - fbase);
    const uint32_t NL = a.nlayers;
    for (uint32_t f = tid; f < nf; f += NT) hc[f] = hb[f] = 0ull;
    __syncthreads();
    const uint32_t nunits = (it.end - it.begin) * NL;
    for (uint32_t u = wv; u < nunits; u += NT / 64) {
        const uint32_t qi = u / NL, K = u - qi * NL;
        const uint32_t i = it.begin + qi;
        const uint32_t s = ufl(a.shard[i]);
        if (s >= a.nshards) continue;
        const LayerSeg L = a.lseg[((size_t)s * a.ngroups + g) * NL + K];
        if (L.n == 0) continue;
        const int32_t qs = (int32_t)ufl(a.start[i]), qe = (int32_t)ufl(a.end[i]);
        const int64_t lo64 = (int64_t)qs - (int64_t)L.lookback + 1;
        const int32_t lo = lo64 > 0 ? (int32_t)lo64 : 0;
        // candidates: from the first start in lo's bucket to the first start
        // in the bucket after q.e - 1's (both jump-table reads in flight at once)
        const uint32_t bkt = (uint32_t)lo >> L.shift, bkt1 = ((uint32_t)(qe - 1) >> L.shift) + 1u;
        const uint32_t p0 = bkt < L.nb ? a.jt[L.jt + bkt] : L.n;
        const uint32_t p1 = bkt1 < L.nb ? a.jt[L.jt + bkt1] : L.n;
        uint32_t p = ufl(p0);
        const uint32_t pe = ufl(p1);
        if (p >= pe) continue;
        const int32_t* ls = a.l_start + L.off;
        const int32_t* le = a.l_end + L.off;
        const uint32_t* lf = a.l_fid + L.off;
        // rounds of two 64-start batches: a lane's (start, end, file id) of
        // both come together, and the next round's are in flight while the
        // current round's are tested (lanes past the candidates hold start
        // INT_MAX)
        auto fetch = [&](uint32_t q, int32_t (&st)[2], int32_t (&en)[2], uint32_t (&fd)[2]) {
#pragma unroll
            for (int h = 0; h < 2; h++) {
                const uint32_t k = q + 64u * h + lane;
                const bool in = k < pe;
                st[h] = in ? ls[k] : 0x7FFFFFFF;
                en[h] = in ? le[k] : 0;
                fd[h] = in ? lf[k] : 0;
            }
        };
        int32_t st[2], en[2];
        uint32_t fd[2];
        fetch(p, st, en, fd);
        for (;;) {
            const uint32_t pn = p + 128;
            int32_t stn[2] = {0x7FFFFFFF, 0x7FFFFFFF}, enn[2] = {0, 0};
            uint32_t fdn[2] = {0, 0};
            if (pn < pe) fetch(pn, stn, enn, fdn);
#pragma unroll
            for (int h = 0; h < 2; h++) {
                if (st[h] < qe && st[h] >= lo && en[h] > qs) {
                    const uint32_t f = fd[h] - fbase;
                    const int64_t ov = (int64_t)min(qe, en[h]) - (int64_t)max(qs, st[h]);
                    atomicAdd(&hc[f], 1ull);
                    atomicAdd(&hb[f], (unsigned long long)ov);
                }
            }
            if (pn >= pe) break;
            p = pn;
#pragma unroll
            for (int h = 0; h < 2; h++) {
                st[h] = stn[h];
                en[h] = enn[h];
                fd[h] = fdn[h];
            }
        }
    }
    __syncthreads();
    const uint64_t cell = (uint64_t)it.row * a.F + fbase;
    for (uint32_t f = tid; f < nf; f += NT) {
        const unsigned long long cv = hc[f], bv = hb[f];
        if (cv | bv) {
            atomicAdd(&a.counts[cell + f], cv);
            atomicAdd(&a.bp[cell + f], bv);
        }
    }
}

template <int ITEMS>
void launch_sweep(const SweepArgs& args, uint32_t grid, hipStream_t st) {
    hipLaunchKernelGGL(sweep_kernel<ITEMS>, dim3(grid), dim3(NT), 0, st, args);
}

template <int ITEMS>
int sweep_occupancy() {
    int nb = 0;
    CHK_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, sweep_kernel<ITEMS>, NT, 0));
    return std::max(1, nb);
}

}  // namespace

void plan_reserve(Query& q, uint64_t n, uint32_t files) {
    Index& ix = *q.idx;
    const uint64_t S = ix.nshards, NG = ix.ngroups;
    const uint64_t jobs = (uint64_t)files * S * NG * 2;
    // every dense file's jobs cover both streams of the DB once; a job's
    // events are its (file, shard) segment, merged with every group and stream
    const uint64_t tiles = ((uint64_t)files * 2 * ix.n + 4 * NG * n) / ((uint64_t)NT * 4) + jobs;
    q.jobs.reserve(jobs);
    q.chunk_job.reserve(tiles + 1);
    q.part.reserve(tiles + jobs + 1);
    q.queue.reserve(8 * QSTRIDE);
    const uint64_t spi_min = ix.tune.sparse_items ? ix.tune.sparse_items : 4;
    q.sitems.reserve(NG * (n / spi_min + files));
}

void query_plan(Query& q) {
    Index& ix = *q.idx;
    if (q.plan_version == ix.tune_version) return;
    query_place(q);
    const uint32_t S = ix.nshards, NG = ix.ngroups;
    const Tuning& tu = ix.tune;
    q.items = tu.items ? tu.items : 8;
    q.tile = NT * q.items;
    // dense jobs: (dense file, shard, group, stream) with events and DB elements
    auto& J = q.jobs_host;
    J.clear();
    uint64_t tiles = 0, bounds = 0, ev = 0;
    const uint32_t ND = (uint32_t)q.dense_rows.size();
    for (uint32_t k = 0; k < ND; k++)
        for (uint32_t s = 0; s < S; s++) {
            const uint32_t m = 2 * q.seg_n[(size_t)k * S + s];
            const uint64_t q_off = ev;
            ev += m;
            if (!m) continue;
            for (uint32_t g = 0; g < NG; g++) {
                const uint64_t db0 = ix.str_off[(size_t)s * NG + g];
                const uint64_t nd = ix.str_off[(size_t)s * NG + g + 1] - db0;
                if (!nd) continue;
                CHK_REQUIRE(nd + m < (1ull << 32), CHK_ERR_INVALID, "shard too large for one job");
                for (uint32_t e = 0; e < 2; e++) {
                    Job j{};
                    j.db_off = db0;
                    j.n_db = (uint32_t)nd;
                    j.q_off = (uint32_t)q_off;
                    j.m = m;
                    j.row = q.dense_rows[k];
                    j.group = g;
                    j.is_end = e;
                    j.tile0 = (uint32_t)tiles;
                    j.part0 = (uint32_t)bounds;
                    const uint64_t nt = (nd + m + q.tile - 1) / q.tile;
                    tiles += nt;
                    bounds += nt + 1;
                    J.push_back(j);
                }
            }
        }
    CHK_REQUIRE(bounds < (1ull << 32), CHK_ERR_INVALID, "query batch too large for one sweep");
    q.ntiles = (uint32_t)tiles;
    q.nbounds = (uint32_t)bounds;
    int occ = 1;
    switch (q.items) {
        case 4: occ = sweep_occupancy<4>(); break;
        case 8: occ = sweep_occupancy<8>(); break;
        case 12: occ = sweep_occupancy<12>(); break;
        case 16: occ = sweep_occupancy<16>(); break;
        default: CHK_REQUIRE(false, CHK_ERR_INVALID, "sweep.items must be 4, 8, 12 or 16");
    }
    if (tu.blocks_per_cu) occ = std::min<int>(occ, (int)tu.blocks_per_cu);
    const uint64_t max_grid = (uint64_t)ix.ncu * occ;
    // about 6 chunks per block: the queues then even out the blocks' ends
    // while a chunk's pipeline start-up stays a small share of it; sizes of
    // 4 mod 8 tiles (measured on 1 M query intervals vs 145 M DB intervals:
    // 12 and 20 tiles run 2-4 % faster than 16 and 24, 28 and up slower)
    uint64_t cz = tiles / (max_grid * 6);
    if (cz >= 4) cz = (cz & ~7ull) + 4;
    q.chunk = tu.chunk ? std::min<uint32_t>(tu.chunk, MAXC)
                       : (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(MAXC, cz));
    q.nchunks = (uint32_t)((tiles + q.chunk - 1) / q.chunk);
    q.grid = (uint32_t)std::min<uint64_t>(max_grid, q.nchunks);
    auto& CJ = q.chunk_job_host;
    CJ.assign(q.nchunks + 1, 0);
    {
        uint32_t j = 0;
        for (uint32_t c = 0; c < q.nchunks; c++) {
            const uint64_t t = (uint64_t)c * q.chunk;
            while (j + 1 < J.size() && J[j + 1].tile0 <= t) j++;
            CJ[c] = j;
        }
        CJ[q.nchunks] = J.empty() ? 0 : (uint32_t)J.size() - 1;
    }
    // sparse items (file x interval chunk x DB file group): enough of them to
    // fill the device even for one small file, at most SPI intervals each
    const uint64_t n_sparse = (q.n - q.n_dense) * NG;
    const uint32_t spi =
        tu.sparse_items ? tu.sparse_items
                        : (uint32_t)std::max<uint64_t>(4, std::min<uint64_t>(SPI, (n_sparse + ix.ncu * 16 - 1) / ((uint64_t)ix.ncu * 16)));
    auto& I = q.items_host;
    I.clear();
    {
        uint64_t p = q.n_dense;
        for (uint32_t r = 0; r < q.nfiles; r++) {
            if (q.file_dense[r] || !q.file_n[r]) continue;
            for (uint32_t g = 0; g < NG; g++)
                for (uint64_t b = 0; b < q.file_n[r]; b += spi) {
                    SparseItem it{};
                    it.row = r;
                    it.begin = (uint32_t)(p + b);
                    it.end = (uint32_t)(p + std::min<uint64_t>(q.file_n[r], b + spi));
                    it.group = g;
                    I.push_back(it);
                }
            p += q.file_n[r];
        }
    }
    hipStream_t st = q.stream;
    q.jobs.reserve(J.size());
    q.chunk_job.reserve(CJ.size());
    q.part.reserve(std::max<uint32_t>(1, q.nbounds));
    q.queue.reserve(8 * QSTRIDE);
    q.sitems.reserve(I.size());
    if (!J.empty()) {
        CHK_HIP(hipMemcpyAsync(q.jobs.p, J.data(), J.size() * sizeof(Job), hipMemcpyHostToDevice, st));
        CHK_HIP(hipMemcpyAsync(q.chunk_job.p, CJ.data(), CJ.size() * 4, hipMemcpyHostToDevice, st));
    }
    if (!I.empty())
        CHK_HIP(hipMemcpyAsync(q.sitems.p, I.data(), I.size() * sizeof(SparseItem),
                               hipMemcpyHostToDevice, st));
    CHK_HIP(hipStreamSynchronize(st));
    q.plan_version = ix.tune_version;
}

void sparse_launch(Index& ix, const SparseItem* items, uint32_t nitems, const uint32_t* shard,
                   const uint32_t* start, const uint32_t* end, uint64_t* d_counts, uint64_t* d_bp,
                   hipStream_t st) {
    if (!nitems) return;
    SparseArgs sa{items, ix.nshards, shard, start, end, ix.d_lseg.p, ix.ngroups, ix.nlayers,
                  ix.nfiles, ix.l_start.p, ix.l_end.p, ix.l_fid.p, ix.jt.p,
                  reinterpret_cast<unsigned long long*>(d_counts),
                  reinterpret_cast<unsigned long long*>(d_bp)};
    hipEvent_t t0 = ix.prof.begin("sparse", st);
    hipLaunchKernelGGL(sparse_kernel, dim3(nitems), dim3(NT), 0, st, sa);
    CHK_HIP(hipGetLastError());
    ix.prof.end("sparse", t0, st);
}

void sweep_run(Query& q, uint64_t* d_counts, uint64_t* d_bp, hipStream_t st) {
    Index& ix = *q.idx;
    CHK_REQUIRE(q.plan_version == ix.tune_version, CHK_ERR_INTERNAL, "sweep plan is stale");
    const uint32_t F = ix.nfiles;
    const size_t cells = (size_t)q.nfiles * F;
    zero3_u64(d_counts, cells, d_bp, cells, reinterpret_cast<uint64_t*>(q.queue.p),
              q.jobs_host.empty() ? 0 : 8 * QSTRIDE / 2, st);
    // The sparse files' sweep and the dense files' sweep write different rows:
    // when a batch has both, the sparse kernel runs on a second stream beside
    // the dense path (forked after the zero fill, joined at the end), so its
    // latency-bound scans overlap the dense sweep's issue-bound tiles.
    const bool both = !q.jobs_host.empty() && !q.items_host.empty();
    hipStream_t sst = st;
    if (both) {
        q.aux.init();
        CHK_HIP(hipEventRecord(q.aux.fork, st));
        CHK_HIP(hipStreamWaitEvent(q.aux.s, q.aux.fork, 0));
        sst = q.aux.s;
    }
    if (!q.items_host.empty()) {
        sparse_launch(ix, q.sitems.p, (uint32_t)q.items_host.size(), q.shard.p, q.start.p, q.end.p,
                      d_counts, d_bp, sst);
    }
    if (!q.jobs_host.empty()) {
        hipEvent_t t0 = ix.prof.begin("qprep", st);
        query_events(q, st);
        ix.prof.end("qprep", t0, st);
        t0 = ix.prof.begin("partition", st);
        hipLaunchKernelGGL(partition_kernel, dim3((q.nbounds + NT - 1) / NT), dim3(NT), 0, st,
                           q.jobs.p, (uint32_t)q.jobs_host.size(), q.nbounds, q.tile, ix.s_coord.p,
                           ix.e_coord.p, ix.s_samp.p, ix.e_samp.p, ix.s_samp2.p, ix.e_samp2.p,
                           q.key32.p, q.part.p);
        CHK_HIP(hipGetLastError());
        ix.prof.end("partition", t0, st);
        SweepArgs args{q.jobs.p, q.chunk_job.p, q.part.p, q.ntiles, q.nchunks, q.chunk, F,
                       ix.s_coord.p, ix.e_coord.p, ix.s_lfid.p, ix.e_lfid.p, q.key32.p, q.rs.p,
                       q.w.p, q.queue.p,
                       reinterpret_cast<unsigned long long*>(d_counts),
                       reinterpret_cast<unsigned long long*>(d_bp)};
        t0 = ix.prof.begin("sweep", st);
        switch (q.items) {
            case 4: launch_sweep<4>(args, q.grid, st); break;
            case 8: launch_sweep<8>(args, q.grid, st); break;
            case 12: launch_sweep<12>(args, q.grid, st); break;
            default: launch_sweep<16>(args, q.grid, st); break;
        }
        CHK_HIP(hipGetLastError());
        ix.prof.end("sweep", t0, st);
    }
    if (both) {
        CHK_HIP(hipEventRecord(q.aux.join, sst));
        CHK_HIP(hipStreamWaitEvent(st, q.aux.join, 0));
    }
}

}  // namespace chk
