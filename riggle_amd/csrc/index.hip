// index.hip -- HBM index build and query-batch preparation.
//
// Index: every DB interval appears once in the start stream of its shard
// (keyed by start) and once in the end stream (keyed by end); each stream is
// sorted by (shard, coordinate) with the file id carried along.  This replaces
// the reference's size-class layers and jump tables (src/io/meta.rs:33-81,
// src/io/jump.rs:66-113): the sweep ranks DB coordinates against the query's
// event stream instead of scanning a look-back window per query.
//
// Query preparation: each query interval yields a start event (key 2*s+1) and
// an end event (key 2*e); keys are prefixed by (query file, shard) and sorted.
// An end at x sorts before a DB coordinate x and a start at x after it, which
// is exactly the half-open overlap rule d.start < q.end && d.end > q.start of
// src/sweep/query.rs:148-151.  Exclusive prefix sums over the sorted events
// give, at any rank p, rS = # starts, rE = p - rS and W = sum(starts) -
// sum(ends) of the events ranked before p.
#include "engine.hpp"

namespace chk {
namespace {

constexpr int NT = 256;

__global__ void make_index_keys(const uint32_t* shard, const uint32_t* coord, size_t n,
                                uint64_t* keys) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) keys[i] = ((uint64_t)shard[i] << 31) | (uint64_t)coord[i];
}

__global__ void split_index_keys(const uint64_t* keys, size_t n, int32_t* coord) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) coord[i] = (int32_t)(keys[i] & 0x7FFFFFFFull);
}

__global__ void narrow_u16(const uint32_t* src, size_t n, uint16_t* dst) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = (uint16_t)src[i];
}

// Two events per query interval.  Dropped intervals (shard not in the DB)
// get the all-ones key and sort after every valid event.
__global__ void make_event_keys(const uint32_t* shard, const uint32_t* start, const uint32_t* end,
                                const uint32_t* qfile, size_t n, uint32_t nshards, int shard_bits,
                                uint64_t* keys) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t s = shard[i];
    uint64_t k0 = ~0ull, k1 = ~0ull;
    if (s < nshards) {
        uint64_t seg = ((uint64_t)(qfile ? qfile[i] : 0u) << shard_bits) | s;
        k0 = (seg << 32) | (2ull * start[i] + 1ull);
        k1 = (seg << 32) | (2ull * end[i]);
    }
    keys[2 * i] = k0;
    keys[2 * i + 1] = k1;
}

// Segment table: seg[0][g] = first event of segment g, seg[1][g] = one past
// its last event; g = qfile * nshards + shard.  Also the per-event values to
// be prefix-summed.
__global__ void segment_and_values(const uint64_t* keys, size_t m, uint32_t nshards,
                                   int shard_bits, uint32_t nseg, uint32_t* seg, uint32_t* rs,
                                   int64_t* w, uint32_t* key32) {
    size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= m) return;
    uint64_t key = keys[k];
    bool valid = key != ~0ull;
    uint32_t lo = (uint32_t)key;
    key32[k] = lo;
    rs[k] = valid ? (lo & 1u) : 0u;
    int64_t c = (int64_t)(lo >> 1);
    w[k] = valid ? ((lo & 1u) ? c : -c) : 0;
    if (!valid) return;
    uint64_t hi = key >> 32;
    uint32_t g = (uint32_t)(hi >> shard_bits) * nshards + (uint32_t)(hi & ((1ull << shard_bits) - 1));
    if (k == 0 || (keys[k - 1] >> 32) != hi) seg[g] = (uint32_t)k;
    if (k + 1 == m || (keys[k + 1] >> 32) != hi) seg[nseg + g] = (uint32_t)(k + 1);
}

inline unsigned nblk(size_t n) { return (unsigned)((n + NT - 1) / NT); }

}  // namespace

void index_build(Index& ix, const uint32_t* shard, const uint32_t* start, const uint32_t* end,
                 const uint32_t* fid) {
    const uint64_t n = ix.n;
    ix.shard_off.assign(ix.nshards + 1, 0);
    ix.fs_count.assign((size_t)ix.nfiles * ix.nshards, 0);
    for (uint64_t i = 0; i < n; i++) {
        CHK_REQUIRE(shard[i] < ix.nshards, CHK_ERR_INVALID, "interval shard id out of range");
        CHK_REQUIRE(fid[i] < ix.nfiles, CHK_ERR_INVALID, "interval file id out of range");
        CHK_REQUIRE(start[i] < end[i], CHK_ERR_INVALID, "interval start must be less than end");
        CHK_REQUIRE(end[i] <= 0x7FFFFFFFu, CHK_ERR_INVALID,
                    "interval coordinate exceeds 2^31-1 (device coordinates are int32)");
        ix.shard_off[shard[i] + 1]++;
        ix.fs_count[(size_t)fid[i] * ix.nshards + shard[i]]++;
    }
    for (uint32_t s = 0; s < ix.nshards; s++) ix.shard_off[s + 1] += ix.shard_off[s];
    ix.s_coord.reserve(n);
    ix.e_coord.reserve(n);
    ix.s_fid.reserve(n);
    ix.e_fid.reserve(n);
    if (n == 0) return;
    hipStream_t st = ix.stream;
    DevBuf<uint32_t> d_shard, d_coord;
    DevBuf<uint64_t> keys;
    d_shard.reserve(n);
    d_coord.reserve(n);
    keys.reserve(n);
    CHK_HIP(hipMemcpyAsync(d_shard.p, shard, n * 4, hipMemcpyHostToDevice, st));
    uint32_t nbits = 31 + bits_for(ix.nshards ? ix.nshards - 1 : 0);
    for (int pass = 0; pass < 2; pass++) {
        const uint32_t* coord = pass == 0 ? start : end;
        int32_t* out_coord = pass == 0 ? ix.s_coord.p : ix.e_coord.p;
        uint32_t* out_fid = pass == 0 ? ix.s_fid.p : ix.e_fid.p;
        CHK_HIP(hipMemcpyAsync(d_coord.p, coord, n * 4, hipMemcpyHostToDevice, st));
        CHK_HIP(hipMemcpyAsync(out_fid, fid, n * 4, hipMemcpyHostToDevice, st));
        hipLaunchKernelGGL(make_index_keys, dim3(nblk(n)), dim3(NT), 0, st, d_shard.p, d_coord.p, n,
                           keys.p);
        radix_sort_u64(keys.p, out_fid, n, nbits, ix.scratch, st);
        hipLaunchKernelGGL(split_index_keys, dim3(nblk(n)), dim3(NT), 0, st, keys.p, n, out_coord);
        CHK_HIP(hipGetLastError());
    }
    if (ix.nfiles <= 65536) {  // 2-byte file ids: 12 bytes per interval over both streams
        ix.fid16 = true;
        ix.s_fid16.reserve(n);
        ix.e_fid16.reserve(n);
        hipLaunchKernelGGL(narrow_u16, dim3(nblk(n)), dim3(NT), 0, st, ix.s_fid.p, n, ix.s_fid16.p);
        hipLaunchKernelGGL(narrow_u16, dim3(nblk(n)), dim3(NT), 0, st, ix.e_fid.p, n, ix.e_fid16.p);
        CHK_HIP(hipGetLastError());
        CHK_HIP(hipStreamSynchronize(st));
        ix.s_fid.release();
        ix.e_fid.release();
    }
    CHK_HIP(hipStreamSynchronize(st));
    ix.scratch.keys_alt.release();
    ix.scratch.vals_alt.release();
}

void query_upload(Query& q, const uint32_t* shard, const uint32_t* start, const uint32_t* end,
                  const uint32_t* qfile) {
    const uint64_t n = q.n;
    for (uint64_t i = 0; i < n; i++) {
        CHK_REQUIRE(start[i] < end[i], CHK_ERR_INVALID, "query interval start must be less than end");
        CHK_REQUIRE(end[i] <= 0x7FFFFFFFu, CHK_ERR_INVALID,
                    "query coordinate exceeds 2^31-1 (device coordinates are int32)");
        if (qfile) CHK_REQUIRE(qfile[i] < q.nfiles, CHK_ERR_INVALID, "query file id out of range");
    }
    CHK_REQUIRE(2 * n < (1ull << 32) - 1, CHK_ERR_INVALID, "query batch too large (2n must fit u32)");
    hipStream_t st = q.idx->stream;
    q.shard.reserve(n);
    q.start.reserve(n);
    q.end.reserve(n);
    if (n) {
        CHK_HIP(hipMemcpyAsync(q.shard.p, shard, n * 4, hipMemcpyHostToDevice, st));
        CHK_HIP(hipMemcpyAsync(q.start.p, start, n * 4, hipMemcpyHostToDevice, st));
        CHK_HIP(hipMemcpyAsync(q.end.p, end, n * 4, hipMemcpyHostToDevice, st));
    }
    if (qfile && q.nfiles > 1 && n) {
        q.qfile.reserve(n);
        CHK_HIP(hipMemcpyAsync(q.qfile.p, qfile, n * 4, hipMemcpyHostToDevice, st));
    }
    CHK_HIP(hipStreamSynchronize(st));
}

void query_prepare(Query& q) {
    Index& ix = *q.idx;
    hipStream_t st = ix.stream;
    const size_t m = 2 * q.n;
    const int shard_bits = (int)bits_for(ix.nshards ? ix.nshards - 1 : 0) + 0;
    const int sb = shard_bits ? shard_bits : 1;
    const uint32_t nseg = q.nfiles * ix.nshards;
    q.keys.reserve(m);
    q.key32.reserve(m + 1);
    q.rs.reserve(m + 1);
    q.w.reserve(m + 1);
    q.seg.reserve(2 * (size_t)nseg);
    hipEvent_t t0 = ix.prof.begin(st);
    CHK_HIP(hipMemsetAsync(q.seg.p, 0, 2 * (size_t)nseg * 4, st));
    CHK_HIP(hipMemsetAsync(q.key32.p + m, 0xFF, sizeof(uint32_t), st));  // sentinel
    if (m) {
        hipLaunchKernelGGL(make_event_keys, dim3(nblk(q.n)), dim3(NT), 0, st, q.shard.p, q.start.p,
                           q.end.p, q.nfiles > 1 ? q.qfile.p : nullptr, (size_t)q.n, ix.nshards, sb,
                           q.keys.p);
        uint32_t nbits = 32 + sb + bits_for(q.nfiles);
        radix_sort_u64(q.keys.p, nullptr, m, nbits, ix.scratch, st);
        hipLaunchKernelGGL(segment_and_values, dim3(nblk(m)), dim3(NT), 0, st, q.keys.p, m,
                           ix.nshards, sb, nseg, q.seg.p, q.rs.p, q.w.p, q.key32.p);
        CHK_HIP(hipGetLastError());
    }
    exclusive_scan_u32(q.rs.p, m, ix.scratch, st);
    exclusive_scan_i64(q.w.p, m, q.w_sums, st);
    ix.prof.end("qprep", t0, st);
    q.seg_host.resize(2 * (size_t)nseg);
    CHK_HIP(hipMemcpyAsync(q.seg_host.data(), q.seg.p, 2 * (size_t)nseg * 4, hipMemcpyDeviceToHost, st));
    CHK_HIP(hipStreamSynchronize(st));
}

}  // namespace chk
