// sort.hip -- stable LSD radix sort (8-bit digits) and exclusive scans.
//
// Radix pass = upsweep (per-tile digit histogram in LDS) -> exclusive scan of
// the digit-major histogram -> scatter.  The scatter ranks keys stably inside
// a tile with wave ballots: for each 256-key round a lane finds the lanes of
// its wave holding the same digit (8 ballots), its rank among them, and the
// per-wave digit counts are prefix-summed across the 4 waves in LDS.  The
// event sort's kernels use 1024-thread blocks (see EVT).
#include "sort.hpp"

namespace chk {
namespace {

constexpr int NT = 256;            // threads per block (4 waves of 64)
constexpr int RS_ITEMS = 16;       // keys per thread per tile
constexpr int RS_TILE = NT * RS_ITEMS;
constexpr int SC_ITEMS = 8;        // scan items per thread
constexpr int SC_TILE = NT * SC_ITEMS;

template <typename T, int BT = NT>
__device__ T block_exclusive_scan(T v, T* lds, T& total) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    T x = v;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        T y = __shfl_up(x, d, 64);
        if (lane >= d) x += y;
    }
    if (lane == 63) lds[w] = x;
    __syncthreads();
    if (threadIdx.x == 0) {
        T run = 0;
        for (int i = 0; i < BT / 64; i++) {
            T t = lds[i];
            lds[i] = run;
            run += t;
        }
        lds[BT / 64] = run;
    }
    __syncthreads();
    T res = x - v + lds[w];
    total = lds[BT / 64];
    __syncthreads();
    return res;
}

// A thread's run of SC_ITEMS consecutive elements, read and written as 16-byte
// vectors when the whole run is in range (runs start 16-byte aligned), else
// element by element with zeros past n.
template <typename T>
__device__ inline void load_run(const T* p, size_t base, size_t n, T (&v)[SC_ITEMS]) {
    static_assert((sizeof(T) * SC_ITEMS) % 16 == 0, "run must be whole 16-byte vectors");
    if (base + SC_ITEMS <= n) {
        const uint4* q = reinterpret_cast<const uint4*>(p + base);
        uint4 t[sizeof(T) * SC_ITEMS / 16];
#pragma unroll
        for (int i = 0; i < (int)(sizeof(T) * SC_ITEMS / 16); i++) t[i] = q[i];
        __builtin_memcpy(v, t, sizeof(v));
    } else {
#pragma unroll
        for (int k = 0; k < SC_ITEMS; k++) v[k] = base + k < n ? p[base + k] : T(0);
    }
}

template <typename T>
__device__ inline void store_run(T* p, size_t base, size_t n, const T (&v)[SC_ITEMS]) {
    if (base + SC_ITEMS <= n) {
        uint4 t[sizeof(T) * SC_ITEMS / 16];
        __builtin_memcpy(t, v, sizeof(v));
        uint4* q = reinterpret_cast<uint4*>(p + base);
#pragma unroll
        for (int i = 0; i < (int)(sizeof(T) * SC_ITEMS / 16); i++) q[i] = t[i];
    } else {
#pragma unroll
        for (int k = 0; k < SC_ITEMS; k++)
            if (base + k < n) p[base + k] = v[k];
    }
}

// ------------------------------------------------------------------ scan

template <typename T>
__global__ __launch_bounds__(NT) void scan_tile_sums(const T* data, size_t n, T* sums) {
    __shared__ T lds[NT / 64 + 1];
    size_t base = (size_t)blockIdx.x * SC_TILE + (size_t)threadIdx.x * SC_ITEMS;
    T v[SC_ITEMS];
    load_run<T>(data, base, n, v);
    T acc = 0;
#pragma unroll
    for (int k = 0; k < SC_ITEMS; k++) acc += v[k];
    T total;
    (void)block_exclusive_scan<T>(acc, lds, total);
    if (threadIdx.x == 0) sums[blockIdx.x] = total;
}

// Single block: exclusive scan of sums[0..nb) in place (scans of many tiles).
template <typename T>
__global__ __launch_bounds__(NT) void scan_sums_single(T* sums, size_t nb) {
    __shared__ T lds[NT / 64 + 1];
    T carry = 0;
    for (size_t c0 = 0; c0 < nb; c0 += SC_TILE) {
        size_t base = c0 + (size_t)threadIdx.x * SC_ITEMS;
        T v[SC_ITEMS];
        T acc = 0;
#pragma unroll
        for (int k = 0; k < SC_ITEMS; k++) {
            v[k] = base + k < nb ? sums[base + k] : T(0);
            acc += v[k];
        }
        T total;
        T ex = block_exclusive_scan<T>(acc, lds, total) + carry;
#pragma unroll
        for (int k = 0; k < SC_ITEMS; k++) {
            if (base + k < nb) sums[base + k] = ex;
            ex += v[k];
        }
        carry += total;
    }
}

// Exclusive scan of tile t.  Its base is either sums[t] (already scanned,
// PRE = false) or, for scans of few tiles, the sum of the raw tile sums before
// it reduced by this block (PRE = true, saving a launch).  The last tile
// writes the total to data[n].
template <typename T, bool PRE>
__global__ __launch_bounds__(NT) void scan_tile_apply(T* data, size_t n, const T* sums, size_t nb) {
    __shared__ T lds[NT / 64 + 1];
    T ptot;
    if (PRE) {
        T pre = 0;
        for (size_t i = threadIdx.x; i < blockIdx.x; i += NT) pre += sums[i];
        (void)block_exclusive_scan<T>(pre, lds, ptot);
    } else {
        ptot = sums[blockIdx.x];
    }
    size_t base = (size_t)blockIdx.x * SC_TILE + (size_t)threadIdx.x * SC_ITEMS;
    T v[SC_ITEMS];
    load_run<T>(data, base, n, v);
    T acc = 0;
#pragma unroll
    for (int k = 0; k < SC_ITEMS; k++) acc += v[k];
    T total;
    T ex = block_exclusive_scan<T>(acc, lds, total) + ptot;
#pragma unroll
    for (int k = 0; k < SC_ITEMS; k++) {
        const T t = v[k];
        v[k] = ex;
        ex += t;
    }
    store_run<T>(data, base, n, v);
    if (blockIdx.x + 1 == nb && threadIdx.x == 0) data[n] = ptot + total;
}

template <typename T>
__global__ void set_zero_one(T* data) {
    if (threadIdx.x == 0) data[0] = 0;
}

template <typename T>
void exclusive_scan_impl(T* data, size_t n, T* sums, hipStream_t st) {
    size_t nb = (n + SC_TILE - 1) / SC_TILE;
    if (nb == 0) {  // a kernel, not a memset: the sequence may be graph-captured
        hipLaunchKernelGGL(set_zero_one<T>, dim3(1), dim3(64), 0, st, data);
        CHK_HIP(hipGetLastError());
        return;
    }
    hipLaunchKernelGGL(scan_tile_sums<T>, dim3((unsigned)nb), dim3(NT), 0, st, data, n, sums);
    if (nb <= 2048) {
        hipLaunchKernelGGL((scan_tile_apply<T, true>), dim3((unsigned)nb), dim3(NT), 0, st, data, n,
                           sums, nb);
    } else {
        hipLaunchKernelGGL(scan_sums_single<T>, dim3(1), dim3(NT), 0, st, sums, nb);
        hipLaunchKernelGGL((scan_tile_apply<T, false>), dim3((unsigned)nb), dim3(NT), 0, st, data,
                           n, sums, nb);
    }
    CHK_HIP(hipGetLastError());
}

// ----------------------------------------------------------- radix sort

__global__ __launch_bounds__(NT) void rs_upsweep(const uint64_t* keys, size_t n, int shift,
                                                  uint32_t* hist, uint32_t nblocks) {
    __shared__ uint32_t h[256];
    h[threadIdx.x] = 0;
    __syncthreads();
    size_t base = (size_t)blockIdx.x * RS_TILE + threadIdx.x;
#pragma unroll
    for (int i = 0; i < RS_ITEMS; i++) {
        size_t idx = base + (size_t)i * NT;
        if (idx < n) atomicAdd(&h[(uint32_t)(keys[idx] >> shift) & 255u], 1u);
    }
    __syncthreads();
    hist[(size_t)threadIdx.x * nblocks + blockIdx.x] = h[threadIdx.x];
}

__global__ __launch_bounds__(NT) void rs_scatter(const uint64_t* kin, const uint32_t* vin,
                                                  uint64_t* kout, uint32_t* vout, size_t n,
                                                  int shift, const uint32_t* offs,
                                                  uint32_t nblocks) {
    __shared__ uint32_t base[256];
    __shared__ uint32_t wcnt[NT / 64][256];
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    base[tid] = offs[(size_t)tid * nblocks + blockIdx.x];
    const uint64_t lt_mask = lane ? (~0ull >> (64 - lane)) : 0ull;
    size_t tile0 = (size_t)blockIdx.x * RS_TILE;
    for (int r = 0; r < RS_ITEMS; r++) {
        size_t idx = tile0 + (size_t)r * NT + tid;
        bool valid = idx < n;
        uint64_t k = valid ? kin[idx] : 0ull;
        uint32_t d = (uint32_t)(k >> shift) & 255u;
#pragma unroll
        for (int ww = 0; ww < NT / 64; ww++) wcnt[ww][tid] = 0;
        __syncthreads();
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; b++) {
            int bit = (d >> b) & 1;
            uint64_t bal = __ballot(bit);
            peers &= bit ? bal : ~bal;
        }
        uint32_t rank = (uint32_t)__popcll(peers & lt_mask);
        if (valid && rank == 0) wcnt[w][d] = (uint32_t)__popcll(peers);
        __syncthreads();
        {
            uint32_t run = base[tid];
#pragma unroll
            for (int ww = 0; ww < NT / 64; ww++) {
                uint32_t t = wcnt[ww][tid];
                wcnt[ww][tid] = run;
                run += t;
            }
            base[tid] = run;
        }
        __syncthreads();
        if (valid) {
            uint32_t pos = wcnt[w][d] + rank;
            kout[pos] = k;
            if (vin) vout[pos] = vin[idx];
        }
        __syncthreads();
    }
}

// ----------------------------------------------------------- event sort
//
// Sort of the 2n events of n query intervals by the key (segment, 2x + b):
// a start event has b = 1 at x = start, an end event b = 0 at x = end, so on
// equal coordinates ends come first.  LSD passes of at most EV_MAXD bits over
// the key bits the batch actually uses.  The first pass builds the keys from
// the intervals as it reads them; the last pass writes the per-event values
// (key32 = 2x + b, rs = b, w = +-x) instead of the keys.  Scatter ranking is
// wave-private: wave v ranks its 1024 keys round by round against its own
// LDS digit counters (ballots give the rank among equal digits of one round),
// and the waves' counters are prefix-summed once per tile.

// 4096-key tiles (2048-key tiles write shorter digit runs: slower) over
// 1024 threads: each of the 16 waves ranks 256 keys in 4 rounds, so a block's
// serial ranking chain is a quarter of what 256 threads would take.
constexpr int EVT = 1024;     // threads per event-sort block
constexpr int EV_ITEMS = 4;   // keys per thread per tile
constexpr int EV_TILE = EVT * EV_ITEMS;
constexpr int EV_MAXD = 9;
constexpr uint32_t EV_BINS = 1u << EV_MAXD;
constexpr int EV_WKEYS = EV_TILE / (EVT / 64);  // keys per wave and tile

struct EvSrc {
    const uint32_t* segid;
    const uint32_t* start;
    const uint32_t* end;
    uint32_t cb;  // coordinate key bits; the segment sits above them
};

// The first pass reads each interval once and takes its two events in turn:
// tile t holds the events of intervals [t * EV_TILE / 2, (t + 1) * EV_TILE / 2)
// (any fixed order of the input is fine for an LSD sort's first pass).
// Kernels load an interval's three fields for all of their intervals first
// (at clamped indices, so the loads are unconditional), then form the keys:
// every load is then in flight at once.
__device__ inline void interval_keys(uint32_t cb, uint32_t segid, uint32_t start, uint32_t end,
                                     uint64_t& ks, uint64_t& ke) {
    const uint64_t seg = (uint64_t)segid << cb;
    ks = seg | (2ull * start + 1ull);
    ke = seg | (2ull * end);
}

template <bool FIRST>
__global__ __launch_bounds__(EVT) void ev_upsweep(EvSrc src, const uint64_t* keys, size_t m,
                                                  int shift, uint32_t nbits, uint32_t* hist,
                                                  uint32_t nblocks) {
    __shared__ uint32_t h[EV_BINS];
    const uint32_t bins = 1u << nbits;
    for (uint32_t d = threadIdx.x; d < bins; d += EVT) h[d] = 0;
    __syncthreads();
    const size_t base = (size_t)blockIdx.x * EV_TILE + threadIdx.x;
    // keys past the end are read at a clamped index and counted with an
    // increment of 0, so loads and counts are unconditional and all loads are
    // issued before the first LDS atomic
    uint32_t dig[EV_ITEMS], inc[EV_ITEMS];
    if (FIRST) {
        const size_t n = m / 2, i0 = (size_t)blockIdx.x * (EV_TILE / 2) + threadIdx.x;
        uint32_t sg[EV_ITEMS / 2], s0[EV_ITEMS / 2], e0[EV_ITEMS / 2];
#pragma unroll
        for (int i = 0; i < EV_ITEMS / 2; i++) {
            const size_t iv = i0 + (size_t)i * EVT, ic = iv < n ? iv : n - 1;
            sg[i] = src.segid[ic];
            s0[i] = src.start[ic];
            e0[i] = src.end[ic];
        }
#pragma unroll
        for (int i = 0; i < EV_ITEMS / 2; i++) {
            const size_t iv = i0 + (size_t)i * EVT;
            uint64_t ks, ke;
            interval_keys(src.cb, sg[i], s0[i], e0[i], ks, ke);
            dig[2 * i] = (uint32_t)(ks >> shift) & (bins - 1);
            dig[2 * i + 1] = (uint32_t)(ke >> shift) & (bins - 1);
            inc[2 * i] = inc[2 * i + 1] = iv < n ? 1u : 0u;
        }
    } else {
#pragma unroll
        for (int i = 0; i < EV_ITEMS; i++) {
            const size_t idx = base + (size_t)i * EVT;
            dig[i] = (uint32_t)(keys[idx < m ? idx : m - 1] >> shift) & (bins - 1);
            inc[i] = idx < m ? 1u : 0u;
        }
    }
#pragma unroll
    for (int i = 0; i < EV_ITEMS; i++) atomicAdd(&h[dig[i]], inc[i]);
    __syncthreads();
    for (uint32_t d = threadIdx.x; d < bins; d += EVT) hist[(size_t)d * nblocks + blockIdx.x] = h[d];
}

template <bool FIRST, bool LAST>
__global__ __launch_bounds__(EVT) void ev_scatter(EvSrc src, const uint64_t* kin, uint64_t* kout,
                                                  size_t m, int shift, uint32_t nbits,
                                                  const uint32_t* offs, uint32_t nblocks,
                                                  uint32_t* key32, uint32_t* rs, int64_t* w) {
    __shared__ uint32_t cnt[EVT / 64][EV_BINS];  // per wave digit counts, then tile-local starts
    __shared__ uint32_t lbase[EV_BINS];           // first tile-local slot of each digit
    __shared__ uint32_t gbase[EV_BINS];           // first output position of each digit
    __shared__ uint64_t stage[EV_TILE];           // the tile in digit order
    __shared__ uint32_t lsum[EVT / 64 + 1];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const uint32_t bins = 1u << nbits;
    for (uint32_t d = tid; d < bins; d += EVT) {
#pragma unroll
        for (int v = 0; v < EVT / 64; v++) cnt[v][d] = 0;
    }
    __syncthreads();
    const uint64_t lt_mask = lane ? (~0ull >> (64 - lane)) : 0ull;
    const size_t tile0 = (size_t)blockIdx.x * EV_TILE;
    const size_t w0 = tile0 + (size_t)wv * EV_WKEYS;
    uint64_t key[EV_ITEMS];
    uint32_t loc[EV_ITEMS];
    bool val[EV_ITEMS];
    if (FIRST) {  // rounds 2q, 2q+1: start and end events of interval iv(q)
        const size_t n = m / 2;
        const size_t i0 = (size_t)blockIdx.x * (EV_TILE / 2) + (size_t)wv * (EV_WKEYS / 2) + lane;
        uint32_t sg[EV_ITEMS / 2], s0[EV_ITEMS / 2], e0[EV_ITEMS / 2];
#pragma unroll
        for (int q = 0; q < EV_ITEMS / 2; q++) {
            const size_t iv = i0 + (size_t)q * 64, ic = iv < n ? iv : n - 1;
            sg[q] = src.segid[ic];
            s0[q] = src.start[ic];
            e0[q] = src.end[ic];
            val[2 * q] = val[2 * q + 1] = iv < n;
        }
#pragma unroll
        for (int q = 0; q < EV_ITEMS / 2; q++)
            interval_keys(src.cb, sg[q], s0[q], e0[q], key[2 * q], key[2 * q + 1]);
    } else {
#pragma unroll
        for (int r = 0; r < EV_ITEMS; r++) {  // all loads issued before the ranking rounds
            const size_t idx = w0 + (size_t)r * 64 + lane;
            key[r] = kin[idx < m ? idx : m - 1];
            val[r] = idx < m;
        }
    }
#pragma unroll
    for (int r = 0; r < EV_ITEMS; r++) {
        const bool valid = val[r];
        const uint64_t k = key[r];
        const uint32_t d = (uint32_t)(k >> shift) & (bins - 1);
        uint64_t peers = __ballot(valid);
        for (uint32_t b = 0; b < nbits; b++) {
            const bool bit = (d >> b) & 1u;
            const uint64_t bal = __ballot(bit);
            peers &= bit ? bal : ~bal;
        }
        // this wave's count of digit d so far, then the group's leader adds
        // the round's group size (same wave: the read is ordered before it)
        const uint32_t pre = valid ? cnt[wv][d] : 0u;
        const uint32_t rank = (uint32_t)__popcll(peers & lt_mask);
        if (valid && rank == 0) cnt[wv][d] = pre + (uint32_t)__popcll(peers);
        loc[r] = pre + rank;
    }
    __syncthreads();
    // tile-local layout: digits in order, within a digit wave 0's keys first
    const uint32_t dpt = bins > EVT ? bins / EVT : 1u;  // digits per thread (bins is a power of 2)
    const uint32_t d0 = tid * dpt;
    uint32_t mine = 0;
    for (uint32_t j = 0; j < dpt; j++)
        if (d0 + j < bins) {
#pragma unroll
            for (int v = 0; v < EVT / 64; v++) mine += cnt[v][d0 + j];
        }
    uint32_t ttot;
    uint32_t run = block_exclusive_scan<uint32_t, EVT>(mine, lsum, ttot);
    for (uint32_t j = 0; j < dpt; j++) {
        const uint32_t d = d0 + j;
        if (d < bins) {
            lbase[d] = run;
            gbase[d] = offs[(size_t)d * nblocks + blockIdx.x];
#pragma unroll
            for (int v = 0; v < EVT / 64; v++) {
                const uint32_t t = cnt[v][d];
                cnt[v][d] = run;
                run += t;
            }
        }
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < EV_ITEMS; r++) {
        if (val[r]) {
            const uint32_t d = (uint32_t)(key[r] >> shift) & (bins - 1);
            stage[cnt[wv][d] + loc[r]] = key[r];
        }
    }
    __syncthreads();
    // write out in tile order: runs of one digit go to consecutive positions
    const uint32_t nt = (uint32_t)(m - tile0 < (size_t)EV_TILE ? m - tile0 : (size_t)EV_TILE);
    const uint64_t cmask = (1ull << src.cb) - 1ull;
    for (uint32_t j = tid; j < nt; j += EVT) {
        const uint64_t k = stage[j];
        const uint32_t d = (uint32_t)(k >> shift) & (bins - 1);
        const uint32_t pos = gbase[d] + (j - lbase[d]);
        if (LAST) {
            key32[pos] = (uint32_t)(k & cmask);
        } else {
            kout[pos] = k;
        }
    }
    if (LAST && blockIdx.x == 0 && tid == 0) key32[m] = 0xFFFFFFFFu;
}

// Events of an empty batch: the sentinel key and zero totals.
__global__ void ev_empty(uint32_t* key32, uint32_t* rs, int64_t* w) {
    if (threadIdx.x == 0) {
        key32[0] = 0xFFFFFFFFu;
        rs[0] = 0;
        w[0] = 0;
    }
}

// The per-event terms of a sorted key 2x + b: rs = b (a start), w = +x for
// a start and -x for an end.
__device__ inline void event_terms(const uint32_t (&k)[SC_ITEMS], uint32_t (&va)[SC_ITEMS],
                                   int64_t (&vb)[SC_ITEMS]) {
#pragma unroll
    for (int j = 0; j < SC_ITEMS; j++) {
        const int64_t c = (int64_t)(k[j] >> 1);
        va[j] = k[j] & 1u;
        vb[j] = (k[j] & 1u) ? c : -c;
    }
}

// Joint exclusive scan of rs (u32) and w (i64), both derived from the sorted
// keys: tile sums, then the tiles with their bases (reduced in the apply
// kernel for up to 2048 tiles).  Positions past n read as key 0 (terms 0).
__global__ __launch_bounds__(NT) void ev_scan_sums(const uint32_t* key32, size_t n,
                                                    uint32_t* srs, int64_t* sw) {
    __shared__ uint32_t l32[NT / 64 + 1];
    __shared__ int64_t l64[NT / 64 + 1];
    const size_t base = (size_t)blockIdx.x * SC_TILE + (size_t)threadIdx.x * SC_ITEMS;
    uint32_t kk[SC_ITEMS], va[SC_ITEMS];
    int64_t vb[SC_ITEMS];
    load_run<uint32_t>(key32, base, n, kk);
    event_terms(kk, va, vb);
    uint32_t a = 0;
    int64_t b = 0;
#pragma unroll
    for (int k = 0; k < SC_ITEMS; k++) {
        a += va[k];
        b += vb[k];
    }
    uint32_t ta;
    int64_t tb;
    (void)block_exclusive_scan<uint32_t>(a, l32, ta);
    (void)block_exclusive_scan<int64_t>(b, l64, tb);
    if (threadIdx.x == 0) {
        srs[blockIdx.x] = ta;
        sw[blockIdx.x] = tb;
    }
}

template <bool PRE>
__global__ __launch_bounds__(NT) void ev_scan_apply(const uint32_t* key32, uint32_t* rs, int64_t* w,
                                                     size_t n, const uint32_t* srs,
                                                     const int64_t* sw, size_t nb) {
    __shared__ uint32_t l32[NT / 64 + 1];
    __shared__ int64_t l64[NT / 64 + 1];
    uint32_t pa;
    int64_t pb;
    if (PRE) {
        uint32_t a = 0;
        int64_t b = 0;
        for (size_t i = threadIdx.x; i < blockIdx.x; i += NT) {
            a += srs[i];
            b += sw[i];
        }
        (void)block_exclusive_scan<uint32_t>(a, l32, pa);
        (void)block_exclusive_scan<int64_t>(b, l64, pb);
    } else {
        pa = srs[blockIdx.x];
        pb = sw[blockIdx.x];
    }
    const size_t base = (size_t)blockIdx.x * SC_TILE + (size_t)threadIdx.x * SC_ITEMS;
    uint32_t kk[SC_ITEMS], va[SC_ITEMS];
    int64_t vb[SC_ITEMS];
    load_run<uint32_t>(key32, base, n, kk);
    event_terms(kk, va, vb);
    uint32_t a = 0;
    int64_t b = 0;
#pragma unroll
    for (int k = 0; k < SC_ITEMS; k++) {
        a += va[k];
        b += vb[k];
    }
    uint32_t ta;
    int64_t tb;
    uint32_t ea = block_exclusive_scan<uint32_t>(a, l32, ta) + pa;
    int64_t eb = block_exclusive_scan<int64_t>(b, l64, tb) + pb;
#pragma unroll
    for (int k = 0; k < SC_ITEMS; k++) {
        const uint32_t x = va[k];
        const int64_t y = vb[k];
        va[k] = ea;
        vb[k] = eb;
        ea += x;
        eb += y;
    }
    store_run<uint32_t>(rs, base, n, va);
    store_run<int64_t>(w, base, n, vb);
    if (blockIdx.x + 1 == nb && threadIdx.x == 0) {
        rs[n] = pa + ta;
        w[n] = pb + tb;
    }
}

}  // namespace

void sort_events(const uint32_t* segid, const uint32_t* start, const uint32_t* end, size_t n,
                 uint32_t seg_bits, uint32_t coord_bits, uint32_t* key32, uint32_t* rs,
                 int64_t* w, uint64_t* keys, SortScratch& s, hipStream_t st) {
    const size_t m = 2 * n;
    if (m == 0) {
        hipLaunchKernelGGL(ev_empty, dim3(1), dim3(64), 0, st, key32, rs, w);
        CHK_HIP(hipGetLastError());
        return;
    }
    CHK_REQUIRE(m < (1ull << 32), CHK_ERR_INVALID, "event sort supports fewer than 2^32 events");
    CHK_REQUIRE(coord_bits >= 1 && coord_bits <= 32 && seg_bits <= 31, CHK_ERR_INTERNAL,
                "event key widths out of range");
    const EvSrc src{segid, start, end, coord_bits};
    const uint32_t bits = coord_bits + seg_bits;
    const uint32_t P = (bits + EV_MAXD - 1) / EV_MAXD;
    const uint32_t nblocks = (uint32_t)((m + EV_TILE - 1) / EV_TILE);
    if (P > 1) s.keys_alt.reserve(m);
    const size_t hn = (size_t)nblocks << (bits / P + (bits % P ? 1 : 0));
    s.hist.reserve(hn + 1);
    uint64_t* buf[2] = {keys, s.keys_alt.p};
    int shift = 0;
    for (uint32_t p = 0; p < P; p++) {
        const uint32_t nb = bits / P + (p < bits % P ? 1u : 0u);  // digit width, <= EV_MAXD
        const bool first = p == 0, last = p + 1 == P;
        const uint64_t* kin = first ? nullptr : buf[(p - 1) & 1];
        uint64_t* kout = last ? nullptr : buf[p & 1];
        if (first)
            hipLaunchKernelGGL(ev_upsweep<true>, dim3(nblocks), dim3(EVT), 0, st, src, kin, m, shift, nb,
                               s.hist.p, nblocks);
        else
            hipLaunchKernelGGL(ev_upsweep<false>, dim3(nblocks), dim3(EVT), 0, st, src, kin, m, shift, nb,
                               s.hist.p, nblocks);
        exclusive_scan_u32(s.hist.p, (size_t)nblocks << nb, s, st);
        auto scatter = first ? (last ? ev_scatter<true, true> : ev_scatter<true, false>)
                             : (last ? ev_scatter<false, true> : ev_scatter<false, false>);
        hipLaunchKernelGGL(scatter, dim3(nblocks), dim3(EVT), 0, st, src, kin, kout, m, shift, nb,
                           s.hist.p, nblocks, key32, rs, w);
        CHK_HIP(hipGetLastError());
        shift += (int)nb;
    }
    // exclusive prefix sums of rs and w, totals at [m]
    const size_t nbs = (m + SC_TILE - 1) / SC_TILE;
    s.sums.reserve(nbs + 1);
    s.sums64.reserve(nbs + 1);
    hipLaunchKernelGGL(ev_scan_sums, dim3((unsigned)nbs), dim3(NT), 0, st, key32, m, s.sums.p, s.sums64.p);
    if (nbs <= 2048) {
        hipLaunchKernelGGL(ev_scan_apply<true>, dim3((unsigned)nbs), dim3(NT), 0, st, key32, rs, w, m, s.sums.p,
                           s.sums64.p, nbs);
    } else {
        hipLaunchKernelGGL(scan_sums_single<uint32_t>, dim3(1), dim3(NT), 0, st, s.sums.p, nbs);
        hipLaunchKernelGGL(scan_sums_single<int64_t>, dim3(1), dim3(NT), 0, st, s.sums64.p, nbs);
        hipLaunchKernelGGL(ev_scan_apply<false>, dim3((unsigned)nbs), dim3(NT), 0, st, key32, rs, w, m, s.sums.p,
                           s.sums64.p, nbs);
    }
    CHK_HIP(hipGetLastError());
}

void sort_events_reserve(size_t n, SortScratch& s) {
    const size_t m = 2 * n;
    const size_t nblocks = (m + EV_TILE - 1) / EV_TILE;
    const size_t hn = nblocks << EV_MAXD;
    s.keys_alt.reserve(m);
    s.hist.reserve(hn + 1);
    s.sums.reserve(std::max((hn + SC_TILE - 1) / SC_TILE, (m + SC_TILE - 1) / SC_TILE) + 1);
    s.sums64.reserve((m + SC_TILE - 1) / SC_TILE + 1);
}

void exclusive_scan_u32(uint32_t* data, size_t n, SortScratch& s, hipStream_t st) {
    s.sums.reserve((n + SC_TILE - 1) / SC_TILE + 1);
    exclusive_scan_impl<uint32_t>(data, n, s.sums.p, st);
}

void exclusive_scan_i64(int64_t* data, size_t n, DevBuf<int64_t>& sums, hipStream_t st) {
    sums.reserve((n + SC_TILE - 1) / SC_TILE + 1);
    exclusive_scan_impl<int64_t>(data, n, sums.p, st);
}

void radix_sort_u64(uint64_t* keys, uint32_t* vals, size_t n, uint32_t nbits, SortScratch& s,
                    hipStream_t st) {
    if (n <= 1 || nbits == 0) return;
    CHK_REQUIRE(n < (1ull << 32), CHK_ERR_INVALID, "radix sort supports fewer than 2^32 keys");
    uint32_t nblocks = (uint32_t)((n + RS_TILE - 1) / RS_TILE);
    s.keys_alt.reserve(n);
    if (vals) s.vals_alt.reserve(n);
    size_t hn = (size_t)nblocks * 256;
    s.hist.reserve(hn + 1);
    uint64_t *ka = keys, *kb = s.keys_alt.p;
    uint32_t *va = vals, *vb = vals ? s.vals_alt.p : nullptr;
    int passes = (int)((nbits + 7) / 8);
    for (int p = 0; p < passes; p++) {
        int shift = 8 * p;
        hipLaunchKernelGGL(rs_upsweep, dim3(nblocks), dim3(NT), 0, st, ka, n, shift, s.hist.p,
                           nblocks);
        exclusive_scan_u32(s.hist.p, hn, s, st);
        hipLaunchKernelGGL(rs_scatter, dim3(nblocks), dim3(NT), 0, st, ka, va, kb, vb, n, shift,
                           s.hist.p, nblocks);
        CHK_HIP(hipGetLastError());
        std::swap(ka, kb);
        std::swap(va, vb);
    }
    if (ka != keys) {
        CHK_HIP(hipMemcpyAsync(keys, ka, n * sizeof(uint64_t), hipMemcpyDeviceToDevice, st));
        if (vals) CHK_HIP(hipMemcpyAsync(vals, va, n * sizeof(uint32_t), hipMemcpyDeviceToDevice, st));
    }
}

}  // namespace chk
