// sweep.hip -- merge-path overlap sweep (count + overlap bp, exact u64).
//
// For one query file Q and one DB interval d on a shard (pair-free form of the
// reference's per-pair scan, src/sweep/query.rs:148-157):
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
// own contiguous tile ranges and software-pipeline them: while tile t is
// ranked and reduced, the DB elements and query events of tile t+1 are already
// in flight into registers.  Each DB element is ranked by binary search among
// the tile's events staged in LDS, and reduced per DB file in an LDS histogram
// that is flushed with one u64 atomic per touched cell when the block moves to
// another query file or ends.
#include "engine.hpp"

#include <algorithm>

namespace chk {
namespace {

constexpr int NT = 256;
constexpr int ITEMS = 8;
constexpr uint32_t TILE = NT * ITEMS;

__device__ inline uint32_t db_key(int32_t x) { return 2u * (uint32_t)x + 1u; }

// Block barrier that orders LDS traffic only: global loads issued for later
// tiles stay in flight across it (__syncthreads would also wait for them).
__device__ inline void lds_barrier() { asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory"); }

// Merge-path split of every tile boundary: part[b] = # DB elements among the
// first d merged items of its job (DB element before event iff dkey <= qkey).
__global__ void partition_kernel(const Job* jobs, uint32_t njobs, uint32_t nbound,
                                 const int32_t* s_coord, const int32_t* e_coord,
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
    uint64_t d = std::min<uint64_t>((uint64_t)(b - J.part0) * TILE, total);
    uint64_t a = d > J.m ? d - J.m : 0, z = std::min<uint64_t>(d, J.n_db);
    while (a < z) {
        uint64_t mid = (a + z) >> 1;
        if (db_key(D[mid]) <= Q[d - 1 - mid]) a = mid + 1; else z = mid;
    }
    part[b] = (uint32_t)a;
}

struct SweepArgs {
    const Job* jobs;
    uint32_t njobs, ntiles, tiles_per_block, F, pwin;
    const uint32_t* part;
    const int32_t* s_coord;
    const int32_t* e_coord;
    const void* s_fid;
    const void* e_fid;
    const uint32_t* qkey;  // 2n+1 entries (sentinel at the end)
    const uint32_t* qrs;   // 2n+1
    const int64_t* qw;     // 2n+1
    unsigned long long* counts;
    unsigned long long* bp;
};

constexpr uint32_t QS = NT;     // query events staged per tile on the fast path
constexpr uint32_t MAXJ = 64;   // jobs a block may span (the host sizes blocks so)
constexpr uint32_t PWIN = 1024; // partition entries held in LDS at a time

// A job as a block sees it, with its segment's prefix bases.
struct LJob {
    Job J;
    uint32_t tile_end;
    uint32_t rs0;
    int64_t w0;
};

// Uniform description of one tile.
struct Meta {
    uint64_t dbase, dlast;  // first element of the tile / last element of the stream
    uint32_t qb, nd, nq, j0, qfile, is_end, jl;
};

// Registers holding one tile in flight.  Buffer loads bound every read by
// the tile's extent in hardware (lanes past it read 0), so loads need no
// per-lane clamping and are issued unconditionally: the compiler then counts
// outstanding loads statically and keeps the next tile's loads in flight
// while this one is reduced.
template <typename FidT>
struct Stage {
    int32_t x[ITEMS];
    FidT f[ITEMS];
    uint32_t key, rs;
    int64_t w;
};

constexpr int BUF_DWORD3 = 0x00020000;  // raw buffer, gfx9-family descriptor word 3

__device__ inline __amdgpu_buffer_rsrc_t make_rsrc(const void* base, uint32_t bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0, (int)bytes,
                                             BUF_DWORD3);
}

__device__ inline uint32_t ufl(uint32_t v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ inline int64_t ufl64(int64_t v) {
    uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
    uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)((uint64_t)v >> 32));
    return (int64_t)(((uint64_t)hi << 32) | lo);
}

template <typename FidT>
__device__ inline void stage_load(const SweepArgs& a, const Meta& m, Stage<FidT>& s) {
    // query events first: staging them into LDS then waits only for these
    // three loads, leaving the DB loads in flight across the barrier
    const uint32_t nst = m.nq + 1 < QS ? m.nq + 1 : QS;  // ranks lie in [0, nq]
    s.key = __builtin_amdgcn_raw_buffer_load_b32(make_rsrc(a.qkey + m.qb, nst * 4u),
                                                 threadIdx.x * 4u, 0, 0);
    s.rs = __builtin_amdgcn_raw_buffer_load_b32(make_rsrc(a.qrs + m.qb, nst * 4u),
                                                threadIdx.x * 4u, 0, 0);
    s.w = (int64_t)__builtin_amdgcn_raw_buffer_load_b64(make_rsrc(a.qw + m.qb, nst * 8u),
                                                        threadIdx.x * 8u, 0, 0);
    const int32_t* coord = (m.is_end ? a.e_coord : a.s_coord) + m.dbase;
    const FidT* fid = reinterpret_cast<const FidT*>(m.is_end ? a.e_fid : a.s_fid) + m.dbase;
    const __amdgpu_buffer_rsrc_t rc = make_rsrc(coord, m.nd * 4u);
    const __amdgpu_buffer_rsrc_t rf = make_rsrc(fid, m.nd * (uint32_t)sizeof(FidT));
#pragma unroll
    for (int k = 0; k < ITEMS; k++) {
        const uint32_t i = threadIdx.x + k * NT;
        s.x[k] = __builtin_amdgcn_raw_buffer_load_b32(rc, i * 4u, 0, 0);
        if constexpr (sizeof(FidT) == 2)
            s.f[k] = __builtin_amdgcn_raw_buffer_load_b16(rf, i * 2u, 0, 0);
        else
            s.f[k] = __builtin_amdgcn_raw_buffer_load_b32(rf, i * 4u, 0, 0);
    }
}

template <typename FidT>
__global__ __launch_bounds__(NT) void sweep_kernel(SweepArgs a) {
    extern __shared__ __align__(16) unsigned char smem[];
    const uint32_t F = a.F;
    unsigned long long* hc = reinterpret_cast<unsigned long long*>(smem);
    unsigned long long* hb = hc + F;
    uint32_t* lkey = reinterpret_cast<uint32_t*>(hb + F);     // QS (16-byte aligned)
    int64_t* lw = reinterpret_cast<int64_t*>(lkey + QS);      // QS
    LJob* ljob = reinterpret_cast<LJob*>(lw + QS);            // MAXJ
    uint32_t* lrs = reinterpret_cast<uint32_t*>(ljob + MAXJ); // QS
    uint32_t* lpart = lrs + QS;                               // PWIN
    const uint32_t tid = threadIdx.x;
    const uint32_t lane = tid & 63u;
    const uint32_t wv = __builtin_amdgcn_readfirstlane(tid >> 6);  // wave index (scalar)

    const uint32_t t_begin = blockIdx.x * a.tiles_per_block;
    const uint32_t t_end = min(a.ntiles, t_begin + a.tiles_per_block);
    if (t_begin >= t_end) return;
    for (uint32_t f = tid; f < F; f += NT) hc[f] = hb[f] = 0ull;

    // Jobs spanned by this block's tiles.
    uint32_t jfirst, jlast;
    {
        uint32_t lo = 0, hi = a.njobs;
        while (hi - lo > 1) {
            uint32_t mid = (lo + hi) >> 1;
            if (a.jobs[mid].tile0 <= t_begin) lo = mid; else hi = mid;
        }
        jfirst = lo;
        hi = a.njobs;
        while (hi - lo > 1) {
            uint32_t mid = (lo + hi) >> 1;
            if (a.jobs[mid].tile0 <= t_end - 1) lo = mid; else hi = mid;
        }
        jlast = lo;
    }
    const uint32_t nj = jlast - jfirst + 1;  // <= MAXJ
    for (uint32_t j = tid; j < nj; j += NT) {
        LJob L;
        L.J = a.jobs[jfirst + j];
        L.tile_end = jfirst + j + 1 < a.njobs ? a.jobs[jfirst + j + 1].tile0 : a.ntiles;
        L.rs0 = a.qrs[L.J.q_off];
        L.w0 = a.qw[L.J.q_off];
        ljob[j] = L;
    }
    uint32_t pbase = 0xFFFFFFFFu;
    uint32_t jl = 0;
    uint32_t cur_q = 0xFFFFFFFFu;
    lds_barrier();

    auto refill = [&](uint32_t base) {
        lds_barrier();
        const uint32_t bound = a.ntiles + a.njobs;  // total partition entries
        for (uint32_t i = tid; i < a.pwin; i += NT)
            if (base + i < bound) lpart[i] = a.part[base + i];
        lds_barrier();
        pbase = base;
    };

    // Uniform description of tile t (values read from LDS are made scalar).
    auto meta = [&](uint32_t t) {
        while (t >= ufl(ljob[jl].tile_end)) jl++;
        const uint32_t tile0 = ufl(ljob[jl].J.tile0), part0 = ufl(ljob[jl].J.part0);
        const uint32_t n_db = ufl(ljob[jl].J.n_db), mq = ufl(ljob[jl].J.m);
        const uint32_t q_off = ufl(ljob[jl].J.q_off);
        const uint64_t db_off = (uint64_t)ufl64((int64_t)ljob[jl].J.db_off);
        const uint32_t k = t - tile0;
        const uint32_t pi = part0 + k;
        if (pbase == 0xFFFFFFFFu || pi < pbase || pi + 1 >= pbase + a.pwin) refill(pi);
        const uint32_t i0 = ufl(lpart[pi - pbase]), i1 = ufl(lpart[pi + 1 - pbase]);
        const uint64_t total = (uint64_t)n_db + mq;
        const uint64_t d0 = (uint64_t)k * TILE;
        const uint64_t d1 = d0 + TILE < total ? d0 + TILE : total;
        Meta m;
        m.dbase = db_off + i0;
        m.dlast = db_off + n_db - 1;
        m.nd = i1 - i0;
        m.j0 = (uint32_t)(d0 - i0);
        m.nq = (uint32_t)(d1 - i1) - m.j0;
        m.qb = q_off + m.j0;
        m.qfile = ufl(ljob[jl].J.qfile);
        m.is_end = ufl(ljob[jl].J.is_end);
        m.jl = jl;
        return m;
    };

    auto flush = [&](uint32_t q) {
        for (uint32_t f = tid; f < F; f += NT) {
            unsigned long long cv = hc[f], bv = hb[f];
            if (cv | bv) {
                atomicAdd(&a.counts[(size_t)q * F + f], cv);
                atomicAdd(&a.bp[(size_t)q * F + f], bv);
                hc[f] = hb[f] = 0ull;
            }
        }
    };

    auto stage_store = [&](const Stage<FidT>& s, const Meta& m) {
        if (m.qfile != cur_q) {
            if (cur_q != 0xFFFFFFFFu) flush(cur_q);
            cur_q = m.qfile;
        }
        // every slot is written: keys past nq hold a sentinel above any DB key
        lkey[tid] = tid < m.nq ? s.key : 0xFFFFFFFFu;
        if (tid <= m.nq) {
            lrs[tid] = s.rs - ljob[m.jl].rs0;
            lw[tid] = s.w - ljob[m.jl].w0;
        }
    };

    // Per-element contribution: rank p (events before the element), prefix
    // values (rs, w) at p.  E stream: (+rs, x*depth - w); S stream:
    // (-(p - rs), -(x*depth - w)); depth = 2 rs - p.
    auto add = [&](const Meta& m, int32_t x, uint32_t f, uint32_t p, uint32_t rsv, int64_t wvv) {
        const int32_t depth = 2 * (int32_t)rsv - (int32_t)(m.j0 + p);
        const int64_t cq = (int64_t)x * (int64_t)depth - wvv;
        unsigned long long cv, bv;
        if (m.is_end) {
            cv = (unsigned long long)(int64_t)rsv;
            bv = (unsigned long long)cq;
        } else {
            cv = (unsigned long long)(-(int64_t)(m.j0 + p - rsv));
            bv = (unsigned long long)(-cq);
        }
        atomicAdd(&hc[f], cv);
        atomicAdd(&hb[f], bv);
    };

    auto compute = [&](const Stage<FidT>& s, const Meta& m) {
        const uint32_t nq = m.nq, nd = m.nd;
        if (nq < QS) {
            // Wave wv owns sub-rows of 64 consecutive elements (k*NT + 64 wv +
            // lane).  A sub-row with no event key inside its key span has one
            // rank for all lanes, found with a uniform cursor over the sorted
            // keys; otherwise lanes count the few keys inside the span.
            uint32_t j = 0;
#pragma unroll
            for (int k = 0; k < ITEMS; k++) {
                const uint32_t base = k * NT + 64u * wv;
                if (base < nd) {
                    const uint32_t nv = nd - base < 64u ? nd - base : 64u;
                    const uint32_t dk = db_key(s.x[k]);
                    const uint32_t dkf = ufl(dk);
                    const uint32_t dkl = __builtin_amdgcn_readlane(dk, nv - 1);
                    uint32_t kj = ufl(lkey[j]);
                    while (kj < dkf) kj = ufl(lkey[++j]);
                    if (kj >= dkl) {
                        const uint32_t rsv = ufl(lrs[j]);
                        const int64_t wq = ufl64(lw[j]);
                        if (lane < nv) add(m, s.x[k], (uint32_t)s.f[k], j, rsv, wq);
                    } else {
                        uint32_t p = j, jj = j, kk = kj;
                        while (kk < dkl) {
                            p += kk < dk ? 1u : 0u;
                            kk = ufl(lkey[++jj]);
                        }
                        if (lane < nv) add(m, s.x[k], (uint32_t)s.f[k], p, lrs[p], lw[p]);
                    }
                }
            }
        } else {
            // event-dense tile: per-lane binary search over the keys in memory
            const uint32_t rs_base = ljob[m.jl].rs0;
            const int64_t w_base = ljob[m.jl].w0;
#pragma unroll
            for (int k = 0; k < ITEMS; k++) {
                const uint32_t i = tid + k * NT;
                if (i < nd) {
                    const uint32_t dk = db_key(s.x[k]);
                    uint32_t lo = 0, hi = nq;
                    while (lo < hi) {
                        const uint32_t mid = (lo + hi) >> 1;
                        if (a.qkey[m.qb + mid] < dk) lo = mid + 1; else hi = mid;
                    }
                    add(m, s.x[k], (uint32_t)s.f[k], lo, a.qrs[m.qb + lo] - rs_base,
                        a.qw[m.qb + lo] - w_base);
                }
            }
        }
    };

    // Two register stages: the loads of one tile are issued before the other
    // tile is staged and reduced.  Past the range end the last tile is
    // reloaded (results unused) so every iteration issues the same loads.
    Stage<FidT> A, B;
    Meta mA = meta(t_begin);
    stage_load(a, mA, A);
    Meta mB = meta(min(t_begin + 1, t_end - 1));
    for (uint32_t t = t_begin;; t += 2) {
        stage_load(a, mB, B);
        stage_store(A, mA);
        lds_barrier();
        compute(A, mA);
        lds_barrier();
        if (t + 1 >= t_end) break;
        mA = meta(min(t + 2, t_end - 1));
        stage_load(a, mA, A);
        stage_store(B, mB);
        lds_barrier();
        compute(B, mB);
        lds_barrier();
        if (t + 2 >= t_end) break;
        mB = meta(min(t + 3, t_end - 1));
    }
    flush(cur_q);
}

}  // namespace

void sweep_run(Query& q, uint64_t* d_counts, uint64_t* d_bp) {
    Index& ix = *q.idx;
    hipStream_t st = ix.stream;
    const uint32_t F = ix.nfiles, S = ix.nshards, nseg = q.nfiles * S;
    const size_t out_bytes = (size_t)q.nfiles * F * sizeof(uint64_t);
    CHK_HIP(hipMemsetAsync(d_counts, 0, out_bytes, st));
    CHK_HIP(hipMemsetAsync(d_bp, 0, out_bytes, st));
    // Host job list from the segment table read back by query_prepare.
    auto& J = q.jobs_host;
    J.clear();
    uint32_t tiles = 0, bounds = 0;
    for (uint32_t r = 0; r < q.nfiles; r++)
        for (uint32_t s = 0; s < S; s++) {
            uint32_t a = q.seg_host[(size_t)r * S + s], z = q.seg_host[nseg + (size_t)r * S + s];
            uint64_t nd = ix.shard_off[s + 1] - ix.shard_off[s];
            if (z <= a || nd == 0) continue;  // shard absent from the query or the DB
            CHK_REQUIRE(nd + (z - a) < (1ull << 32), CHK_ERR_INVALID, "shard too large for one job");
            for (uint32_t e = 0; e < 2; e++) {
                Job j{};
                j.db_off = ix.shard_off[s];
                j.n_db = (uint32_t)nd;
                j.q_off = a;
                j.m = z - a;
                j.qfile = r;
                j.is_end = e;
                j.tile0 = tiles;
                j.part0 = bounds;
                uint32_t nt = (uint32_t)((nd + j.m + TILE - 1) / TILE);
                tiles += nt;
                bounds += nt + 1;
                J.push_back(j);
            }
        }
    if (J.empty()) return;
    q.jobs.reserve(J.size());
    q.part.reserve(bounds);
    CHK_HIP(hipMemcpyAsync(q.jobs.p, J.data(), J.size() * sizeof(Job), hipMemcpyHostToDevice, st));
    hipEvent_t t0 = ix.prof.begin(st);
    hipLaunchKernelGGL(partition_kernel, dim3((bounds + NT - 1) / NT), dim3(NT), 0, st, q.jobs.p,
                       (uint32_t)J.size(), bounds, ix.s_coord.p, ix.e_coord.p, q.key32.p, q.part.p);
    CHK_HIP(hipGetLastError());
    ix.prof.end("partition", t0, st);
    size_t lds = 16ull * F + QS * 16ull + MAXJ * sizeof(LJob) + PWIN * 4ull;
    CHK_REQUIRE(lds <= 64 * 1024, CHK_ERR_INVALID,
                "too many DB files for the LDS histogram of one block (limit about 3000)");
    uint32_t occ = std::max<uint32_t>(1, std::min<uint32_t>(8, (uint32_t)((160 * 1024) / (lds + 256))));
    if (ix.tune.blocks_per_cu) occ = std::min(occ, ix.tune.blocks_per_cu);
    uint32_t max_blocks = (uint32_t)ix.ncu * occ;
    uint32_t tpb = std::max<uint32_t>(1, (tiles + max_blocks - 1) / max_blocks);
    if (ix.tune.tiles_per_block) tpb = ix.tune.tiles_per_block;
    const uint32_t pwin = ix.tune.part_window ? std::max<uint32_t>(2, std::min(PWIN, ix.tune.part_window)) : PWIN;
    // a block may span at most MAXJ jobs
    auto max_span = [&](uint32_t tp) {
        uint32_t worst = 0, j = 0;
        for (uint32_t t0 = 0; t0 < tiles; t0 += tp) {
            uint32_t t1 = std::min(tiles, t0 + tp) - 1;
            while (j + 1 < J.size() && J[j + 1].tile0 <= t0) j++;
            uint32_t j2 = j;
            while (j2 + 1 < J.size() && J[j2 + 1].tile0 <= t1) j2++;
            worst = std::max(worst, j2 - j + 1);
        }
        return worst;
    };
    while (tpb > 1 && max_span(tpb) > MAXJ) tpb = (tpb + 1) / 2;
    uint32_t nblocks = (tiles + tpb - 1) / tpb;
    SweepArgs args{q.jobs.p, (uint32_t)J.size(), tiles, tpb, F, pwin, q.part.p, ix.s_coord.p,
                   ix.e_coord.p, nullptr, nullptr, q.key32.p, q.rs.p, q.w.p,
                   reinterpret_cast<unsigned long long*>(d_counts),
                   reinterpret_cast<unsigned long long*>(d_bp)};
    t0 = ix.prof.begin(st);
    if (ix.fid16) {
        args.s_fid = ix.s_fid16.p;
        args.e_fid = ix.e_fid16.p;
        hipLaunchKernelGGL(sweep_kernel<uint16_t>, dim3(nblocks), dim3(NT), lds, st, args);
    } else {
        args.s_fid = ix.s_fid.p;
        args.e_fid = ix.e_fid.p;
        hipLaunchKernelGGL(sweep_kernel<uint32_t>, dim3(nblocks), dim3(NT), lds, st, args);
    }
    CHK_HIP(hipGetLastError());
    ix.prof.end("sweep", t0, st);
}

}  // namespace chk
