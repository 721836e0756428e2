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
// tile costs the same wherever the query's events cluster.  A block stages the
// tile's query events and their prefix values in LDS, ranks each DB element by
// binary search in LDS, and reduces per DB file in an LDS histogram that is
// flushed with one u64 atomic per touched cell when the block moves to another
// query file or ends.
#include "engine.hpp"

#include <algorithm>

namespace chk {
namespace {

constexpr int NT = 256;
constexpr uint32_t TILE = 2048;

__device__ inline uint32_t db_key(int32_t x) { return 2u * (uint32_t)x + 1u; }

// Merge-path split of every tile boundary: part[b] = # DB elements among the
// first d merged items of its job (DB element before event iff dkey <= qkey).
__global__ void partition_kernel(const Job* jobs, uint32_t njobs, uint32_t nbound,
                                 const int32_t* s_coord, const int32_t* e_coord,
                                 const uint64_t* qkeys, uint32_t* part) {
    uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nbound) return;
    uint32_t lo = 0, hi = njobs;  // last job with part0 <= b
    while (hi - lo > 1) {
        uint32_t mid = (lo + hi) >> 1;
        if (jobs[mid].part0 <= b) lo = mid; else hi = mid;
    }
    const Job J = jobs[lo];
    const int32_t* D = (J.is_end ? e_coord : s_coord) + J.db_off;
    const uint64_t* Q = qkeys + J.q_off;
    uint64_t total = (uint64_t)J.n_db + J.m;
    uint64_t d = std::min<uint64_t>((uint64_t)(b - J.part0) * TILE, total);
    uint64_t a = d > J.m ? d - J.m : 0, z = std::min<uint64_t>(d, J.n_db);
    while (a < z) {
        uint64_t mid = (a + z) >> 1;
        if (db_key(D[mid]) <= (uint32_t)Q[d - 1 - mid]) a = mid + 1; else z = mid;
    }
    part[b] = (uint32_t)a;
}

__global__ __launch_bounds__(NT) void sweep_kernel(
    const Job* __restrict__ jobs, uint32_t njobs, uint32_t ntiles, uint32_t tiles_per_block,
    const uint32_t* __restrict__ part, const int32_t* __restrict__ s_coord,
    const uint32_t* __restrict__ s_fid, const int32_t* __restrict__ e_coord,
    const uint32_t* __restrict__ e_fid, const uint64_t* __restrict__ qkeys,
    const uint32_t* __restrict__ qrs, const int64_t* __restrict__ qw,
    unsigned long long* __restrict__ counts, unsigned long long* __restrict__ bp, uint32_t F) {
    extern __shared__ __align__(16) unsigned char smem[];
    unsigned long long* hc = reinterpret_cast<unsigned long long*>(smem);
    unsigned long long* hb = hc + F;
    int64_t* lw = reinterpret_cast<int64_t*>(hb + F);      // TILE+1
    uint32_t* lrs = reinterpret_cast<uint32_t*>(lw + TILE + 1);  // TILE+1
    uint32_t* lkey = lrs + TILE + 1;                          // TILE
    __shared__ Job sj;

    const int tid = threadIdx.x;
    for (uint32_t f = tid; f < F; f += NT) hc[f] = hb[f] = 0ull;

    uint32_t t_begin = blockIdx.x * tiles_per_block;
    uint32_t t_end = std::min(ntiles, t_begin + tiles_per_block);
    uint32_t cur_q = 0xFFFFFFFFu;
    for (uint32_t t = t_begin; t < t_end; t++) {
        if (tid == 0) {
            uint32_t lo = 0, hi = njobs;
            while (hi - lo > 1) {
                uint32_t mid = (lo + hi) >> 1;
                if (jobs[mid].tile0 <= t) lo = mid; else hi = mid;
            }
            sj = jobs[lo];
        }
        __syncthreads();
        const Job J = sj;
        if (J.qfile != cur_q) {
            if (cur_q != 0xFFFFFFFFu) {
                for (uint32_t f = tid; f < F; f += NT) {
                    unsigned long long c = hc[f], b = hb[f];
                    if (c | b) {
                        atomicAdd(&counts[(size_t)cur_q * F + f], c);
                        atomicAdd(&bp[(size_t)cur_q * F + f], b);
                        hc[f] = hb[f] = 0ull;
                    }
                }
            }
            cur_q = J.qfile;
        }
        const uint32_t k = t - J.tile0;
        const uint64_t total = (uint64_t)J.n_db + J.m;
        const uint64_t d0 = (uint64_t)k * TILE;
        const uint64_t d1 = std::min<uint64_t>(d0 + TILE, total);
        const uint32_t i0 = part[J.part0 + k], i1 = part[J.part0 + k + 1];
        const uint32_t j0 = (uint32_t)(d0 - i0), j1 = (uint32_t)(d1 - i1);
        const uint32_t nq = j1 - j0;
        const uint32_t rs0 = qrs[J.q_off];
        const int64_t w0 = qw[J.q_off];
        for (uint32_t x = tid; x <= nq; x += NT) {
            uint32_t g = J.q_off + j0 + x;
            lrs[x] = qrs[g] - rs0;
            lw[x] = qw[g] - w0;
            if (x < nq) lkey[x] = (uint32_t)qkeys[g];
        }
        __syncthreads();
        const int32_t* coord = (J.is_end ? e_coord : s_coord) + J.db_off;
        const uint32_t* fid = (J.is_end ? e_fid : s_fid) + J.db_off;
        for (uint32_t i = i0 + tid; i < i1; i += NT) {
            const int32_t xv = coord[i];
            const uint32_t f = fid[i];
            const uint32_t dk = db_key(xv);
            uint32_t lo = 0, hi = nq;
            while (lo < hi) {
                uint32_t mid = (lo + hi) >> 1;
                if (lkey[mid] < dk) lo = mid + 1; else hi = mid;
            }
            const int64_t rsv = lrs[lo];
            const int64_t re = (int64_t)(j0 + lo) - rsv;
            const int64_t cq = (int64_t)xv * (rsv - re) - lw[lo];
            unsigned long long c, b;
            if (J.is_end) {
                c = (unsigned long long)rsv;
                b = (unsigned long long)cq;
            } else {
                c = (unsigned long long)(-re);
                b = (unsigned long long)(-cq);
            }
            atomicAdd(&hc[f], c);
            atomicAdd(&hb[f], b);
        }
        __syncthreads();
    }
    if (cur_q != 0xFFFFFFFFu) {
        for (uint32_t f = tid; f < F; f += NT) {
            unsigned long long c = hc[f], b = hb[f];
            if (c | b) {
                atomicAdd(&counts[(size_t)cur_q * F + f], c);
                atomicAdd(&bp[(size_t)cur_q * F + f], b);
            }
        }
    }
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
                       (uint32_t)J.size(), bounds, ix.s_coord.p, ix.e_coord.p, q.keys.p, q.part.p);
    CHK_HIP(hipGetLastError());
    ix.prof.end("partition", t0, st);
    size_t lds = 16ull * F + 12ull * (TILE + 1) + 4ull * TILE + 64;
    CHK_REQUIRE(lds <= 160 * 1024, CHK_ERR_INVALID,
                "too many DB files for the LDS histogram (limit about 7000)");
    uint32_t occ = lds <= 40 * 1024 ? 4 : lds <= 53 * 1024 ? 3 : lds <= 80 * 1024 ? 2 : 1;
    uint32_t max_blocks = (uint32_t)ix.ncu * occ;
    uint32_t tpb = std::max<uint32_t>(1, (tiles + max_blocks - 1) / max_blocks);
    uint32_t nblocks = (tiles + tpb - 1) / tpb;
    t0 = ix.prof.begin(st);
    hipLaunchKernelGGL(sweep_kernel, dim3(nblocks), dim3(NT), lds, st, q.jobs.p, (uint32_t)J.size(),
                       tiles, tpb, q.part.p, ix.s_coord.p, ix.s_fid.p, ix.e_coord.p, ix.e_fid.p,
                       q.keys.p, q.rs.p, q.w.p, reinterpret_cast<unsigned long long*>(d_counts),
                       reinterpret_cast<unsigned long long*>(d_bp), F);
    CHK_HIP(hipGetLastError());
    ix.prof.end("sweep", t0, st);
}

}  // namespace chk
