// index.hip -- HBM index build and query-batch upload / planning.
//
// Dense form: every DB interval appears once in the start stream of its
// (shard, file group) and once in the end stream; each stream is sorted by
// coordinate and carries the group-local file id.  The sweep ranks DB
// coordinates against a query's sorted event stream, so it needs neither the
// reference's size classes nor its look-back windows (src/io/meta.rs:33-81,
// src/io/jump.rs:66-113, src/sweep/query.rs:84-117).
//
// Sparse form: per (shard, file group, size layer) the intervals sorted by
// start with their ends and file ids, plus a jump table -- the reference's layer files
// (src/io/layer.rs:34-128, src/tasks/build.rs:196-247) held in HBM, used to
// enumerate the overlapping pairs of small query files directly.
//
// Query events: each interval of a dense query file yields a start event
// (key 2*s+1) and an end event (key 2*e) prefixed by (dense file, shard) and
// sorted.  An end at x sorts before a DB coordinate x (key 2x+1) and a start
// at x after it, which is exactly the half-open rule d.start < q.end &&
// d.end > q.start of src/sweep/query.rs:105-108.  Exclusive prefix sums over
// the sorted events give, at any rank p, rS = # starts, rE = p - rS and
// W = sum(starts) - sum(ends) of the events ranked before p.
#include <algorithm>
#include <thread>

#include "dbfile.hpp"

#include "engine.hpp"

namespace chk {
namespace {

constexpr int NT = 256;
constexpr uint32_t SPARSE_UNIT = 32;  // fixed visit charge per (interval, group, layer) lookup
// Sparse visit cost in dense-element units, measured: with it the rule
// switches to the dense sweep where the two methods take the same time for
// single query files of the 1M-interval query's laws (about 75k intervals
// against the 145M-interval DB, 100k against the 2.1G-interval one).
constexpr double SPARSE_WEIGHT = 1.5;

inline unsigned nblk(size_t n) { return (unsigned)((n + NT - 1) / NT); }

// key = (segment << 31) | coordinate; segment = shard * ngroups + group
__global__ void make_stream_keys(const uint32_t* shard, const uint32_t* coord, const uint32_t* fid,
                                 size_t n, uint32_t ngroups, uint64_t* keys, uint32_t* vals) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t f = fid[i];
    const uint64_t seg = (uint64_t)shard[i] * ngroups + f / FG;
    keys[i] = (seg << 31) | (uint64_t)coord[i];
    vals[i] = f % FG;
}

__global__ void split_stream_keys(const uint64_t* keys, const uint32_t* vals, size_t n,
                                  int32_t* coord, uint16_t* lfid) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    coord[i] = (int32_t)(keys[i] & 0x7FFFFFFFull);
    lfid[i] = (uint16_t)(vals[i] * 8u);  // byte offset of the file's u64 histogram slot
}

// Size class of a length (src/io/meta.rs:33-61 for min 128, factor 4).
__host__ __device__ inline uint32_t layer_of(uint32_t len) {
    if (len < LAYER_MIN) return 0;
    const uint32_t l2 = 31u - (uint32_t)__builtin_clz(len);
    return (l2 - 7u) / 2u + 1u;
}

// key = (((shard * ngroups + group) * nlayers + layer) << 31) | start;
// payload = interval index
__global__ void make_layer_keys(const uint32_t* shard, const uint32_t* start, const uint32_t* end,
                                const uint32_t* fid, size_t n, uint32_t ngroups, uint32_t nlayers,
                                uint64_t* keys, uint32_t* vals) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint64_t seg = ((uint64_t)shard[i] * ngroups + fid[i] / FG) * nlayers + layer_of(end[i] - start[i]);
    keys[i] = (seg << 31) | (uint64_t)start[i];
    vals[i] = (uint32_t)i;
}

__global__ void gather_layer(const uint64_t* keys, const uint32_t* vals, size_t n,
                             const uint32_t* end, const uint32_t* fid, int32_t* l_start,
                             int32_t* l_end, uint32_t* l_fid) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t k = vals[i];
    l_start[i] = (int32_t)(keys[i] & 0x7FFFFFFFull);
    l_end[i] = (int32_t)end[k];
    l_fid[i] = fid[k];
}

// Jump table of every (shard, layer) segment: element i of a segment writes
// jt[t] = i for every bucket t in (bucket(i-1), bucket(i)]; the last element
// also fills the buckets after it with the segment length.
__global__ void build_jump(const LayerSeg* segs, uint32_t nseg, const int32_t* l_start,
                           uint64_t n, uint32_t* jt) {
    for (uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; g < n;
         g += (uint64_t)gridDim.x * blockDim.x) {
        uint32_t lo = 0, hi = nseg;  // last segment with off <= g
        while (hi - lo > 1) {
            uint32_t mid = (lo + hi) >> 1;
            if (segs[mid].off <= g) lo = mid; else hi = mid;
        }
        while (lo + 1 < nseg && segs[lo].n == 0) lo++;  // empty segments share offsets
        const LayerSeg S = segs[lo];
        const uint32_t i = (uint32_t)(g - S.off);
        const uint32_t b = (uint32_t)l_start[g] >> S.shift;
        const uint32_t bp = i ? ((uint32_t)l_start[g - 1] >> S.shift) + 1u : 0u;
        for (uint32_t t = bp; t <= b && t < S.nb; t++) jt[S.jt + t] = i;
        if (i + 1 == S.n)
            for (uint32_t t = b + 1; t < S.nb; t++) jt[S.jt + t] = S.n;
    }
}

// samp[g] = coord[g * stride]
__global__ void sample_stream(const int32_t* coord, size_t ns, uint32_t stride, int32_t* samp) {
    const size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g < ns) samp[g] = coord[g * stride];
}

__global__ void fill_u64(uint64_t* p, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        p[i] = 0ull;
}

// Three zero fills in one launch.
__global__ void fill3_u64(uint64_t* p0, size_t n0, uint64_t* p1, size_t n1, uint64_t* p2, size_t n2) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n0 + n1 + n2;
         i += (size_t)gridDim.x * blockDim.x) {
        if (i < n0) p0[i] = 0ull;
        else if (i < n0 + n1) p1[i - n0] = 0ull;
        else p2[i - n0 - n1] = 0ull;
    }
}

}  // namespace

void zero3_u64(uint64_t* p0, size_t n0, uint64_t* p1, size_t n1, uint64_t* p2, size_t n2,
               hipStream_t st) {
    const size_t n = n0 + n1 + n2;
    if (!n) return;
    const unsigned g = (unsigned)std::min<size_t>(nblk(n), 4096);
    hipLaunchKernelGGL(fill3_u64, dim3(g), dim3(NT), 0, st, p0, n0, p1, n1, p2, n2);
    CHK_HIP(hipGetLastError());
}

void zero_u64(uint64_t* p, size_t n, hipStream_t st) {
    if (!n) return;
    const unsigned g = (unsigned)std::min<size_t>(nblk(n), 4096);
    hipLaunchKernelGGL(fill_u64, dim3(g), dim3(NT), 0, st, p, n);
    CHK_HIP(hipGetLastError());
}

void index_init(Index& ix, int device, uint32_t nshards, uint32_t nfiles, uint64_t n) {
    CHK_REQUIRE(nfiles > 0, CHK_ERR_INVALID, "nfiles must be positive");
    CHK_HIP(hipSetDevice(device));
    ix.device = device;
    ix.nshards = nshards;
    ix.nfiles = nfiles;
    ix.n = n;
    ix.ncu = num_cus(device);
    CHK_HIP(hipStreamCreateWithFlags(&ix.stream, hipStreamNonBlocking));
}

void set_genome(Index& ix, const uint64_t* shard_size) {
    Moments& M = ix.mom;
    M.nbins.resize(ix.nshards);
    for (uint32_t s = 0; s < ix.nshards; s++) {
        CHK_REQUIRE(shard_size[s] <= 0xFFFFFFFFull, CHK_ERR_INVALID, "shard size exceeds u32");
        M.nbins[s] = (uint32_t)((shard_size[s] + 99) / 100);
    }
    M.d_nbins.reserve(std::max<uint32_t>(1, ix.nshards));
    if (ix.nshards)
        CHK_HIP(hipMemcpy(M.d_nbins.p, M.nbins.data(), ix.nshards * 4, hipMemcpyHostToDevice));
    M.genome = true;
}

void index_build(Index& ix, const uint32_t* shard, const uint32_t* start, const uint32_t* end,
                 const uint32_t* fid) {
    const uint64_t n = ix.n;
    const uint32_t S = ix.nshards;
    ix.ngroups = std::max<uint32_t>(1, (ix.nfiles + FG - 1) / FG);
    const uint32_t NG = ix.ngroups;
    ix.fs_count.assign((size_t)ix.nfiles * S, 0);
    ix.shard_n.assign(S, 0);
    ix.str_off.assign((size_t)S * NG + 1, 0);
    uint32_t maxlayer = 0;
    for (uint64_t i = 0; i < n; i++) {
        CHK_REQUIRE(shard[i] < S, CHK_ERR_INVALID, "interval shard id out of range");
        CHK_REQUIRE(fid[i] < ix.nfiles, CHK_ERR_INVALID, "interval file id out of range");
        CHK_REQUIRE(start[i] < end[i], CHK_ERR_INVALID, "interval start must be less than end");
        CHK_REQUIRE(end[i] <= 0x7FFFFFFFu, CHK_ERR_INVALID,
                    "interval coordinate exceeds 2^31-1 (device coordinates are int32)");
        ix.shard_n[shard[i]]++;
        ix.str_off[(size_t)shard[i] * NG + fid[i] / FG + 1]++;
        ix.fs_count[(size_t)fid[i] * S + shard[i]]++;
        maxlayer = std::max(maxlayer, layer_of(end[i] - start[i]));
    }
    for (size_t g = 0; g < (size_t)S * NG; g++) ix.str_off[g + 1] += ix.str_off[g];
    ix.nlayers = n ? maxlayer + 1 : 1;
    const uint32_t NL = ix.nlayers;
    ix.s_coord.reserve(n);
    ix.e_coord.reserve(n);
    ix.s_lfid.reserve(n);
    ix.e_lfid.reserve(n);
    ix.l_start.reserve(n);
    ix.l_end.reserve(n);
    ix.l_fid.reserve(n);
    hipStream_t st = ix.stream;
    // layer segment sizes on the host (for the jump-table geometry)
    // layered segments per (shard, file group, layer)
    const size_t NSEG = (size_t)S * NG * NL;
    std::vector<uint64_t> lcount(NSEG, 0);
    std::vector<uint32_t> lmax(NSEG, 0);
    for (uint64_t i = 0; i < n; i++) {
        const size_t b = ((size_t)shard[i] * NG + fid[i] / FG) * NL + layer_of(end[i] - start[i]);
        lcount[b]++;
        lmax[b] = std::max(lmax[b], start[i]);
    }
    ix.lseg.assign(NSEG, LayerSeg{});
    ix.layer_n.assign(NL, 0);
    ix.span = 0;
    for (uint32_t sh = 0; sh < S; sh++) {
        uint32_t mx = 0;
        for (size_t b = (size_t)sh * NG * NL; b < (size_t)(sh + 1) * NG * NL; b++) {
            ix.layer_n[b % NL] += lcount[b];
            mx = std::max(mx, lmax[b]);
        }
        ix.span += mx;
    }
    uint64_t off = 0, jto = 0;
    for (size_t b = 0; b < NSEG; b++) {
        LayerSeg& L = ix.lseg[b];
        const uint32_t layer = (uint32_t)(b % NL);
        L.off = off;
        L.n = (uint32_t)lcount[b];
        CHK_REQUIRE(lcount[b] < (1ull << 32), CHK_ERR_INVALID, "layer segment exceeds 2^32 intervals");
        uint64_t lb = LAYER_MIN;
        for (uint32_t k = 0; k < layer; k++) lb = std::min<uint64_t>(lb * LAYER_GF, 0xFFFFFFFFull);
        L.lookback = (uint32_t)lb;
        // power-of-two buckets: at most max(256, n/4) of them, none narrower
        // than half the layer's largest length
        const uint64_t target = std::max<uint64_t>(256, lcount[b] / 4);
        uint32_t sh = 0;
        while ((1ull << sh) < lb / 2) sh++;
        while (((uint64_t)lmax[b] >> sh) + 1 > target) sh++;
        L.shift = sh;
        L.nb = L.n ? (uint32_t)(((uint64_t)lmax[b] >> sh) + 1) : 0;
        L.jt = jto;
        off += L.n;
        jto += L.nb;
    }
    ix.jt.reserve(jto);
    ix.d_lseg.reserve(ix.lseg.size());
    CHK_HIP(hipMemcpyAsync(ix.d_lseg.p, ix.lseg.data(), ix.lseg.size() * sizeof(LayerSeg),
                           hipMemcpyHostToDevice, st));
    if (n == 0) {
        CHK_HIP(hipStreamSynchronize(st));
        return;
    }
    DevBuf<uint32_t> d_shard, d_start, d_end, d_fid, vals;
    DevBuf<uint64_t> keys;
    d_shard.reserve(n);
    d_start.reserve(n);
    d_end.reserve(n);
    d_fid.reserve(n);
    vals.reserve(n);
    keys.reserve(n);
    CHK_HIP(hipMemcpyAsync(d_shard.p, shard, n * 4, hipMemcpyHostToDevice, st));
    CHK_HIP(hipMemcpyAsync(d_start.p, start, n * 4, hipMemcpyHostToDevice, st));
    CHK_HIP(hipMemcpyAsync(d_end.p, end, n * 4, hipMemcpyHostToDevice, st));
    CHK_HIP(hipMemcpyAsync(d_fid.p, fid, n * 4, hipMemcpyHostToDevice, st));
    // dense streams
    const uint32_t sbits = 31 + bits_for((uint64_t)S * NG - 1);
    for (int pass = 0; pass < 2; pass++) {
        hipLaunchKernelGGL(make_stream_keys, dim3(nblk(n)), dim3(NT), 0, st, d_shard.p,
                           pass == 0 ? d_start.p : d_end.p, d_fid.p, (size_t)n, NG, keys.p, vals.p);
        radix_sort_u64(keys.p, vals.p, n, sbits, ix.scratch, st);
        hipLaunchKernelGGL(split_stream_keys, dim3(nblk(n)), dim3(NT), 0, st, keys.p, vals.p,
                           (size_t)n, pass == 0 ? ix.s_coord.p : ix.e_coord.p,
                           pass == 0 ? ix.s_lfid.p : ix.e_lfid.p);
        const size_t ns = (n + PSTRIDE - 1) / PSTRIDE, ns2 = (n + PSTRIDE2 - 1) / PSTRIDE2;
        DevBuf<int32_t>& samp = pass == 0 ? ix.s_samp : ix.e_samp;
        DevBuf<int32_t>& samp2 = pass == 0 ? ix.s_samp2 : ix.e_samp2;
        samp.reserve(ns + 1);
        samp2.reserve(ns2 + 1);
        if (ns)
            hipLaunchKernelGGL(sample_stream, dim3(nblk(ns)), dim3(NT), 0, st,
                               pass == 0 ? ix.s_coord.p : ix.e_coord.p, ns, PSTRIDE, samp.p);
        if (ns2)
            hipLaunchKernelGGL(sample_stream, dim3(nblk(ns2)), dim3(NT), 0, st,
                               pass == 0 ? ix.s_coord.p : ix.e_coord.p, ns2, PSTRIDE2, samp2.p);
        CHK_HIP(hipGetLastError());
    }
    // layered table + jump tables
    const uint32_t lbits = 31 + bits_for((uint64_t)NSEG - 1);
    hipLaunchKernelGGL(make_layer_keys, dim3(nblk(n)), dim3(NT), 0, st, d_shard.p, d_start.p,
                       d_end.p, d_fid.p, (size_t)n, NG, NL, keys.p, vals.p);
    radix_sort_u64(keys.p, vals.p, n, lbits, ix.scratch, st);
    hipLaunchKernelGGL(gather_layer, dim3(nblk(n)), dim3(NT), 0, st, keys.p, vals.p, (size_t)n,
                       d_end.p, d_fid.p, ix.l_start.p, ix.l_end.p, ix.l_fid.p);
    const unsigned gj = (unsigned)std::min<uint64_t>(nblk(n), (uint64_t)ix.ncu * 32);
    hipLaunchKernelGGL(build_jump, dim3(gj), dim3(NT), 0, st, ix.d_lseg.p, (uint32_t)ix.lseg.size(),
                       ix.l_start.p, n, ix.jt.p);
    CHK_HIP(hipGetLastError());
    CHK_HIP(hipStreamSynchronize(st));
    ix.scratch.keys_alt.release();
    ix.scratch.vals_alt.release();
}

// Query ingest.  A batch is first laid out canonically -- intervals grouped
// by query file in file order, those on shards absent from the DB dropped
// (the reference skips such shards, src/tasks/query/mod.rs:193-199) -- in
// pinned staging memory and copied to the device once.  query_place then
// orders the files for the sweep on the device (gather kernel), so a change
// of method threshold never goes back to the host.

// Canonical layout from arbitrary per-interval file ids (counting sort).
// Large batches are validated, counted and laid out over host threads: each
// thread takes a contiguous range of the input and writes its intervals of
// file r after those of the earlier ranges, so the layout is the sequential
// one (file order, input order within a file).  The first invalid interval
// in input order is the one reported.
void query_upload(Query& q, const uint32_t* shard, const uint32_t* start, const uint32_t* end,
                  const uint32_t* qfile) {
    Index& ix = *q.idx;
    const uint64_t n_in = q.n_in;
    const uint32_t S = ix.nshards, R = q.nfiles;
    // threads: one per 128k intervals, at most the host's, and few enough
    // that their per-(file, shard) counters stay small
    const unsigned T = (unsigned)std::max<uint64_t>(
        1, std::min<uint64_t>({db::host_threads(), n_in >> 17,
                               (1ull << 24) / std::max<uint64_t>(1, (uint64_t)R * S)}));
    struct Part {
        std::vector<uint64_t> fn, fs;  // per file, per (file, shard)
        std::vector<uint64_t> len;     // summed lengths per file (exact)
        uint32_t max_end = 0;
        uint64_t bad = ~0ull;          // first invalid interval of the range
        int why = 0;
    };
    std::vector<Part> parts(T);
    auto range = [&](unsigned t) {
        return std::make_pair(n_in * t / T, n_in * (t + 1) / T);
    };
    auto count = [&](unsigned t) {
        Part& P = parts[t];
        P.fn.assign(R, 0);
        P.fs.assign((size_t)R * S, 0);
        P.len.assign(R, 0);
        const auto [i0, i1] = range(t);
        for (uint64_t i = i0; i < i1; i++) {
            const uint32_t r = qfile ? qfile[i] : 0u;
            const int why = start[i] >= end[i] ? 1 : (end[i] > 0x7FFFFFFFu ? 2 : (r >= R ? 3 : 0));
            if (why) {
                P.bad = i;
                P.why = why;
                return;
            }
            if (shard[i] >= S || !ix.shard_n[shard[i]]) continue;
            P.fn[r]++;
            P.max_end = std::max(P.max_end, end[i]);
            P.len[r] += end[i] - start[i];
            P.fs[(size_t)r * S + shard[i]]++;
        }
    };
    auto run = [&](auto&& f) { db::host_run(T, f); };
    run(count);
    for (const Part& P : parts) {
        if (P.bad == ~0ull) continue;
        CHK_REQUIRE(P.why != 1, CHK_ERR_INVALID, "query interval start must be less than end");
        CHK_REQUIRE(P.why != 2, CHK_ERR_INVALID,
                    "query coordinate exceeds 2^31-1 (device coordinates are int32)");
        CHK_REQUIRE(false, CHK_ERR_INVALID, "query file id out of range");
    }
    q.file_n.assign(R, 0);
    q.file_len.assign(R, 0.0);
    q.fs_n.assign((size_t)R * S, 0);
    q.max_end = 0;
    std::vector<uint64_t> len(R, 0);
    for (const Part& P : parts) {
        q.max_end = std::max(q.max_end, P.max_end);
        for (uint32_t r = 0; r < R; r++) {
            q.file_n[r] += P.fn[r];
            len[r] += P.len[r];
        }
        for (size_t k = 0; k < (size_t)R * S; k++) q.fs_n[k] += P.fs[k];
    }
    for (uint32_t r = 0; r < R; r++) q.file_len[r] = (double)len[r];
    q.file_off.assign(R + 1, 0);
    for (uint32_t r = 0; r < R; r++) q.file_off[r + 1] = q.file_off[r] + q.file_n[r];
    q.n = q.file_off[R];
    CHK_REQUIRE(q.n < (1ull << 32), CHK_ERR_INVALID, "query batch exceeds 2^32 intervals");
    q.st_shard.reserve(q.n);
    q.st_start.reserve(q.n);
    q.st_end.reserve(q.n);
    // each range's first slot in every file: after the earlier ranges' intervals
    std::vector<std::vector<uint64_t>> fill(T);
    {
        std::vector<uint64_t> run_off(q.file_off.begin(), q.file_off.end() - 1);
        for (unsigned t = 0; t < T; t++) {
            fill[t] = run_off;
            for (uint32_t r = 0; r < R; r++) run_off[r] += parts[t].fn[r];
        }
    }
    run([&](unsigned t) {
        std::vector<uint64_t>& f = fill[t];
        const auto [i0, i1] = range(t);
        for (uint64_t i = i0; i < i1; i++) {
            if (shard[i] >= S || !ix.shard_n[shard[i]]) continue;
            const uint64_t k = f[qfile ? qfile[i] : 0u]++;
            q.st_shard.p[k] = shard[i];
            q.st_start.p[k] = start[i];
            q.st_end.p[k] = end[i];
        }
    });
    query_stage_commit(q);
}

// Canonical layout already in the staging buffers (file_n, file_len, fs_n
// and file_off filled by the caller): copy it to the device.
void query_stage_commit(Query& q) {
    const uint64_t n = q.n;
    q.c_shard.reserve(n);
    q.c_start.reserve(n);
    q.c_end.reserve(n);
    if (n) {
        CHK_HIP(hipMemcpyAsync(q.c_shard.p, q.st_shard.p, n * 4, hipMemcpyHostToDevice, q.stream));
        CHK_HIP(hipMemcpyAsync(q.c_start.p, q.st_start.p, n * 4, hipMemcpyHostToDevice, q.stream));
        CHK_HIP(hipMemcpyAsync(q.c_end.p, q.st_end.p, n * 4, hipMemcpyHostToDevice, q.stream));
    }
    // the staging buffers are reused by the next batch of this query object:
    // the copies must have read them before the caller refills them
    q.staged = true;
    q.placed = false;
    q.plan_version = 0xFFFFFFFFu;
    query_visits(q);
    query_plan(q);
}

namespace {

// placed[dst + k] = canonical[src + k] for every file; dense files' intervals
// also get their (dense file, shard) segment id for the event keys.
__global__ void place_kernel(const PlaceFile* files, uint32_t nfiles, uint64_t n, uint32_t S,
                             const uint32_t* cs, const uint32_t* cb, const uint32_t* ce,
                             uint32_t* shard, uint32_t* start, uint32_t* end, uint32_t* row,
                             uint32_t* segid) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x) {
        uint32_t lo = 0, hi = nfiles;  // last file with dst <= i (among non-empty ones)
        while (hi - lo > 1) {
            uint32_t mid = (lo + hi) >> 1;
            if (files[mid].dst <= i) lo = mid; else hi = mid;
        }
        const PlaceFile f = files[lo];
        const uint64_t k = f.src + (i - f.dst);
        const uint32_t sh = cs[k];
        shard[i] = sh;
        start[i] = cb[k];
        end[i] = ce[k];
        row[i] = f.row;
        if (f.drank != 0xFFFFFFFFu) segid[i] = f.drank * S + sh;
    }
}

// Estimated work of the sparse path per query file: for every (interval,
// layer) the DB starts in [q.s - lookback + 1, q.e) counted from the jump
// table at bucket granularity, plus a fixed charge per lookup.  The canonical
// layout groups intervals by file (file_off).  The estimate only picks
// between two exact methods, so it is taken over every stride-th interval of
// a file (at most VSAMP of them) and scaled by the host.
// Lanes are (sample, file group x layer) pairs: a wave takes ipw samples at a
// time (64 / nsub of them, or one with lanes looping over the sub-tables
// when nsub > 64), so a request of a few thousand intervals spreads its
// lookups over thousands of lanes.  Each block takes a contiguous range of
// samples; a lane accumulates the visits of the wave's current file in a
// register, lanes already past that file (a wave straddling a file boundary)
// add theirs directly, and the waves' sums meet in LDS: about one atomic per
// block and file.
constexpr uint64_t VSAMP = 8192;

__device__ inline uint32_t file_of(const uint64_t* off, uint32_t nfiles, uint64_t i) {
    uint32_t lo = 0, hi = nfiles;  // last file with off <= i (empty files skipped)
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (off[mid] <= i) lo = mid; else hi = mid;
    }
    return lo;
}

// fo = [file_off (R + 1) | samp_off (R + 1) | stride (R)]
__global__ __launch_bounds__(NT) void sparse_visits(const uint32_t* shard, const uint32_t* start,
                                                    const uint32_t* end, const uint64_t* fo,
                                                    uint32_t nfiles, const LayerSeg* lseg,
                                                    uint32_t nsub, const uint32_t* jt,
                                                    uint64_t per_block, uint32_t ipw,
                                                    unsigned long long* visits) {
    __shared__ uint32_t wf[NT / 64];
    __shared__ unsigned long long wsum[NT / 64];
    const uint64_t* file_off = fo;
    const uint64_t* samp_off = fo + nfiles + 1;
    const uint64_t* stride = fo + 2 * ((uint64_t)nfiles + 1);
    const uint32_t lane = threadIdx.x & 63u, wv = threadIdx.x >> 6;
    const uint32_t jl = nsub <= 64 ? lane / nsub : 0u;  // this lane's sample in a wave's chunk
    const uint32_t K0 = nsub <= 64 ? lane - jl * nsub : lane;
    const uint64_t T = samp_off[nfiles];
    const uint64_t s0 = (uint64_t)blockIdx.x * per_block;
    const uint64_t s1 = s0 + per_block < T ? s0 + per_block : T;
    uint32_t fc = 0xFFFFFFFFu;  // the wave's current file (wave-uniform) and its sample range
    uint64_t fs0 = 0, fend = 0, fbase = 0, fstride = 1;
    unsigned long long acc = 0;
    for (uint64_t sb = s0 + (uint64_t)wv * ipw; sb < s1; sb += (uint64_t)(NT / 64) * ipw) {
        if (fc == 0xFFFFFFFFu || sb >= fend) {
            if (fc != 0xFFFFFFFFu) {
                unsigned long long v = acc;
                for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
                if (lane == 0 && v) atomicAdd(&visits[fc], v);
                acc = 0;
            }
            fc = file_of(samp_off, nfiles, sb);
            fs0 = samp_off[fc];
            fend = samp_off[fc + 1];
            fbase = file_off[fc];
            fstride = stride[fc];
        }
        const uint64_t sj = sb + jl;
        const bool act = jl < ipw && sj < s1;
        uint32_t f = fc;
        unsigned long long v = 0;
        if (act) {
            uint64_t i;
            if (sj < fend) {
                i = fbase + (sj - fs0) * fstride;
            } else {
                f = file_of(samp_off, nfiles, sj);
                i = file_off[f] + (sj - samp_off[f]) * stride[f];
            }
            const uint32_t s = shard[i];
            const int64_t qs = start[i], qe = end[i];
            for (uint32_t K = K0; K < nsub; K += 64) {
                const LayerSeg L = lseg[(size_t)s * nsub + K];
                if (L.n == 0) continue;
                const int64_t a = qs - (int64_t)L.lookback + 1;
                const uint32_t b0 = (uint32_t)(a > 0 ? a : 0) >> L.shift;
                const uint32_t b1 = ((uint32_t)qe >> L.shift) + 1;
                const uint32_t j0 = b0 < L.nb ? jt[L.jt + b0] : L.n;
                const uint32_t j1 = b1 < L.nb ? jt[L.jt + b1] : L.n;
                v += (j1 > j0 ? j1 - j0 : 0) + SPARSE_UNIT;
            }
        }
        if (f != fc) {
            if (v) atomicAdd(&visits[f], v);
        } else {
            acc += v;
        }
    }
    // waves' sums: equal files merged by one thread, in wave order
    unsigned long long v = acc;
    for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
    if (lane == 0) {
        wf[wv] = fc;
        wsum[wv] = v;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t f = 0xFFFFFFFFu;
        unsigned long long sum = 0;
        for (uint32_t w = 0; w < NT / 64; w++) {
            if (wf[w] == 0xFFFFFFFFu) continue;
            if (wf[w] != f) {
                if (f != 0xFFFFFFFFu && sum) atomicAdd(&visits[f], sum);
                f = wf[w];
                sum = 0;
            }
            sum += wsum[w];
        }
        if (f != 0xFFFFFFFFu && sum) atomicAdd(&visits[f], sum);
    }
}

}  // namespace

// Method per file and device placement.  The dense path streams the whole DB
// (2 |D| elements whatever the file's size); the sparse path visits, per
// query interval and size layer K, the DB starts within lookback_K + its
// length before its end, counted per file on the device from the jump tables
// (sparse_visits) when the batch is staged.  A file goes sparse when its
// weighted visits are below the
// dense pass (the sparse path's cost per visited element against
// the dense path's), or, with an explicit sweep.sparse_max, when it has at
// most that many intervals.
// Files are ordered dense first, so each sparse file is a contiguous range
// and the dense files' events are a prefix.  Redone only when the choice
// changes.
bool file_is_dense(const Index& ix, uint64_t n, double visits) {
    if (!n) return false;
    if (ix.tune.method == 1) return true;
    if (ix.tune.method == 2) return false;
    if (ix.tune.sparse_max) return n > ix.tune.sparse_max;
    // measured cost of a sparse visit in dense-element units
    return SPARSE_WEIGHT * visits > 2.0 * (double)ix.n + 2.0 * (double)n;
}

void query_reserve(Query& q, uint64_t n, uint32_t files) {
    q.st_shard.reserve(n);
    q.st_start.reserve(n);
    q.st_end.reserve(n);
    q.c_shard.reserve(n);
    q.c_start.reserve(n);
    q.c_end.reserve(n);
    q.shard.reserve(n);
    q.start.reserve(n);
    q.end.reserve(n);
    q.row.reserve(n);
    q.segid.reserve(n);
    q.place.reserve(files);
    q.vis_off.reserve(3 * (size_t)files + 2);
    q.vis.reserve(files);
    q.keys.reserve(2 * n);
    q.key32.reserve(2 * n + 1);
    q.rs.reserve(2 * n + 1);
    q.w.reserve(2 * n + 1);
    sort_events_reserve(n, q.scratch);
    plan_reserve(q, n, files);
}

void query_visits(Query& q) {
    Index& ix = *q.idx;
    const uint32_t R = q.nfiles;
    q.file_visits.assign(R, 0.0);
    const uint32_t nsub = ix.ngroups * ix.nlayers;
    if (!q.n || ix.n == 0 || nsub == 0) return;
    // per file: every stride-th interval, at most VSAMP of them
    std::vector<uint64_t> fo(3 * (size_t)R + 2);
    uint64_t T = 0;
    for (uint32_t r = 0; r < R; r++) {
        const uint64_t n = q.file_off[r + 1] - q.file_off[r];
        const uint64_t st = std::max<uint64_t>(1, (n + VSAMP - 1) / VSAMP);
        fo[r] = q.file_off[r];
        fo[(size_t)R + 1 + r] = T;
        fo[2 * (size_t)R + 2 + r] = st;
        T += (n + st - 1) / st;
    }
    fo[R] = q.file_off[R];
    fo[2 * (size_t)R + 1] = T;
    // buffers of the query object: a free here would wait for the whole device
    DevBuf<uint64_t>& d_fo = q.vis_off;
    DevBuf<uint64_t>& d_vis = q.vis;
    d_fo.reserve(fo.size());
    d_vis.reserve(R);
    CHK_HIP(hipMemcpyAsync(d_fo.p, fo.data(), fo.size() * 8, hipMemcpyHostToDevice, q.stream));
    zero_u64(d_vis.p, R, q.stream);
    const uint32_t ipw = nsub <= 64 ? 64 / nsub : 1u;
    // contiguous sample ranges, at least one wave chunk per wave, up to eight
    // blocks per CU
    const uint64_t chunks = (T + ipw - 1) / ipw;
    const uint64_t g = std::max<uint64_t>(1, std::min<uint64_t>((chunks + NT / 64 - 1) / (NT / 64),
                                                                (uint64_t)ix.ncu * 8));
    const uint64_t per_block = (T + g - 1) / g;
    hipLaunchKernelGGL(sparse_visits, dim3((unsigned)g), dim3(NT), 0, q.stream, q.c_shard.p,
                       q.c_start.p, q.c_end.p, d_fo.p, R, ix.d_lseg.p, nsub, ix.jt.p, per_block, ipw,
                       reinterpret_cast<unsigned long long*>(d_vis.p));
    CHK_HIP(hipGetLastError());
    std::vector<uint64_t> h(R);
    CHK_HIP(hipMemcpyAsync(h.data(), d_vis.p, R * 8, hipMemcpyDeviceToHost, q.stream));
    CHK_HIP(hipStreamSynchronize(q.stream));  // fo and h are host memory of this frame
    for (uint32_t r = 0; r < R; r++) {
        const uint64_t ns = fo[(size_t)R + 2 + r] - fo[(size_t)R + 1 + r];
        if (ns) q.file_visits[r] = (double)h[r] * (double)(q.file_off[r + 1] - q.file_off[r]) / (double)ns;
    }
}

void query_place(Query& q) {
    Index& ix = *q.idx;
    const uint32_t S = ix.nshards, R = q.nfiles;
    std::vector<uint8_t> dense(R, 0);
    for (uint32_t r = 0; r < R; r++) dense[r] = file_is_dense(ix, q.file_n[r], q.file_visits[r]);
    if (q.placed && dense == q.file_dense) return;
    q.file_dense = dense;
    q.dense_rows.clear();
    std::vector<PlaceFile> pf;
    uint64_t p = 0;
    for (int pass = 0; pass < 2; pass++)
        for (uint32_t r = 0; r < R; r++) {
            if ((dense[r] != 0) != (pass == 0) || !q.file_n[r]) continue;
            PlaceFile f{};
            f.src = q.file_off[r];
            f.dst = p;
            f.n = (uint32_t)q.file_n[r];
            f.row = r;
            f.drank = pass == 0 ? (uint32_t)q.dense_rows.size() : 0xFFFFFFFFu;
            if (pass == 0) q.dense_rows.push_back(r);
            pf.push_back(f);
            p += q.file_n[r];
            if (pass == 0) q.n_dense = p;
        }
    if (q.dense_rows.empty()) q.n_dense = 0;
    CHK_REQUIRE(2 * q.n_dense < (1ull << 32) - 1, CHK_ERR_INVALID,
                "query batch too large (2 x dense intervals must fit u32)");
    const uint32_t ND = (uint32_t)q.dense_rows.size();
    q.seg_n.assign((size_t)ND * S, 0);
    for (uint32_t k = 0; k < ND; k++)
        for (uint32_t s = 0; s < S; s++) q.seg_n[(size_t)k * S + s] = q.fs_n[(size_t)q.dense_rows[k] * S + s];
    const uint64_t n = q.n;
    q.shard.reserve(n);
    q.start.reserve(n);
    q.end.reserve(n);
    q.row.reserve(n);
    q.segid.reserve(q.n_dense);
    q.place.reserve(pf.size());
    if (n) {
        CHK_HIP(hipMemcpyAsync(q.place.p, pf.data(), pf.size() * sizeof(PlaceFile),
                               hipMemcpyHostToDevice, q.stream));
        const unsigned g = (unsigned)std::min<uint64_t>(nblk(n), (uint64_t)ix.ncu * 16);
        hipLaunchKernelGGL(place_kernel, dim3(g), dim3(NT), 0, q.stream,
                           q.place.p, (uint32_t)pf.size(), n, S,
                           q.c_shard.p, q.c_start.p, q.c_end.p, q.shard.p, q.start.p, q.end.p,
                           q.row.p, q.segid.p);
        CHK_HIP(hipGetLastError());
    }
    const size_t m = 2 * q.n_dense;
    q.keys.reserve(m);
    q.key32.reserve(m + 1);
    q.rs.reserve(m + 1);
    q.w.reserve(m + 1);
    CHK_HIP(hipStreamSynchronize(q.stream));  // pf is host memory of this frame
    q.placed = true;
}

// Sort the dense files' events and build the prefix sums (enqueued on st).
void query_events(Query& q, hipStream_t st) {
    Index& ix = *q.idx;
    const uint32_t nseg = std::max<uint32_t>(1, (uint32_t)q.dense_rows.size() * ix.nshards);
    sort_events(q.segid.p, q.start.p, q.end.p, q.n_dense, bits_for(nseg - 1),
                bits_for(2ull * q.max_end), q.key32.p, q.rs.p, q.w.p, q.keys.p, q.scratch, st);
}

}  // namespace chk
