// sort.hip -- stable LSD radix sort (8-bit digits) and exclusive scans.
//
// Radix pass = upsweep (per-tile digit histogram in LDS) -> exclusive scan of
// the digit-major histogram -> scatter.  The scatter ranks keys stably inside
// a tile with wave ballots: for each 256-key round a lane finds the lanes of
// its wave holding the same digit (8 ballots), its rank among them, and the
// per-wave digit counts are prefix-summed across the 4 waves in LDS.
#include "sort.hpp"

namespace chk {
namespace {

constexpr int NT = 256;            // threads per block (4 waves of 64)
constexpr int RS_ITEMS = 16;       // keys per thread per tile
constexpr int RS_TILE = NT * RS_ITEMS;
constexpr int SC_ITEMS = 8;        // scan items per thread
constexpr int SC_TILE = NT * SC_ITEMS;

template <typename T>
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
        for (int i = 0; i < NT / 64; i++) {
            T t = lds[i];
            lds[i] = run;
            run += t;
        }
        lds[NT / 64] = run;
    }
    __syncthreads();
    T res = x - v + lds[w];
    total = lds[NT / 64];
    __syncthreads();
    return res;
}

// ------------------------------------------------------------------ scan

template <typename T>
__global__ __launch_bounds__(NT) void scan_tile_sums(const T* data, size_t n, T* sums) {
    __shared__ T lds[NT / 64 + 1];
    size_t base = (size_t)blockIdx.x * SC_TILE + (size_t)threadIdx.x * SC_ITEMS;
    T acc = 0;
#pragma unroll
    for (int k = 0; k < SC_ITEMS; k++)
        if (base + k < n) acc += data[base + k];
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
    T acc = 0;
#pragma unroll
    for (int k = 0; k < SC_ITEMS; k++) {
        v[k] = base + k < n ? data[base + k] : T(0);
        acc += v[k];
    }
    T total;
    T ex = block_exclusive_scan<T>(acc, lds, total) + ptot;
#pragma unroll
    for (int k = 0; k < SC_ITEMS; k++) {
        if (base + k < n) data[base + k] = ex;
        ex += v[k];
    }
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

}  // namespace

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
