// noalloc.hip -- an overlap sweep of one query file that allocates nothing.
//
// chk_query_create + chk_sweep_async allocate the query's device arrays and
// the sweep's work buffers.  Callers that cannot allocate device memory
// (code under HIP graph capture, tight serving loops, harnesses that forbid
// hipMalloc) use chk_sweep_noalloc instead: the query's intervals are staged
// and its work items built in one file-scope device array, `scratch`, of a
// fixed size, and the file is swept by the sparse pair kernel (the
// reference's windowed scan, exact for any query size).  Outputs are the
// caller's device buffers, cleared by a fill kernel.
//
// Scratch layout (bytes):
//   [0, 4 MAXN)           shard ids, [4 MAXN, 8 MAXN) starts, [8 MAXN, 12 MAXN) ends
//   [12 MAXN, + 16 MAXI)  sparse work items
// Each device holds its own copy of the array at its own address; host code
// never uses the array's name, only the address hipGetSymbolAddress returns
// for the current device.
#include <mutex>

#include "engine.hpp"

namespace chk {

namespace {

constexpr uint64_t MAXN = CHK_NOALLOC_MAX_INTERVALS;  // query intervals per call
constexpr uint64_t MAXI = 1u << 16;                    // sparse work items per call
constexpr uint64_t ITEMS_OFF = 12 * MAXN;
constexpr uint64_t SCRATCH_BYTES = ITEMS_OFF + MAXI * sizeof(SparseItem);
static_assert(SCRATCH_BYTES == CHK_NOALLOC_SCRATCH_BYTES, "scratch size differs from the header");
static_assert(sizeof(SparseItem) == 16, "SparseItem layout");

__device__ __align__(256) unsigned char scratch[SCRATCH_BYTES];

constexpr int MAXDEV = 64;
std::mutex g_mu;                    // guards g_ptr
void* g_ptr[MAXDEV] = {};           // scratch address per device
std::mutex g_dev_mu[MAXDEV];        // one call at a time per device: calls share its scratch

// Work item k: file row 0, DB file group k / nchunk, intervals
// [spi * (k % nchunk), min(n, spi * (k % nchunk + 1))).
__global__ void noalloc_items(SparseItem* items, uint32_t nitems, uint32_t nchunk, uint32_t spi,
                              uint32_t n) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nitems) return;
    const uint32_t g = k / nchunk, b = k - g * nchunk;
    SparseItem it;
    it.row = 0;
    it.begin = b * spi;
    it.end = min(n, (b + 1) * spi);
    it.group = g;
    items[k] = it;
}

// The device's scratch address (looked up once per device, with the device
// current).
unsigned char* scratch_ptr(int device) {
    CHK_REQUIRE(device >= 0 && device < MAXDEV, CHK_ERR_INVALID, "device id out of range");
    std::lock_guard<std::mutex> g(g_mu);
    if (!g_ptr[device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            (void)hipGetLastError();
            throw Error(CHK_ERR_HIP, std::string("hipGetSymbolAddress(scratch) failed: ") +
                                         hipGetErrorString(e));
        }
        g_ptr[device] = p;
    }
    return static_cast<unsigned char*>(g_ptr[device]);
}

}  // namespace

void sweep_noalloc(Index& ix, uint64_t n, const uint32_t* shard, const uint32_t* start,
                   const uint32_t* end, uint64_t* d_counts, uint64_t* d_bp) {
    CHK_REQUIRE(n <= MAXN, CHK_ERR_INVALID,
                "chk_sweep_noalloc takes at most " + std::to_string(MAXN) +
                    " query intervals per call (its scratch array is fixed)");
    for (uint64_t i = 0; i < n; i++) {
        CHK_REQUIRE(start[i] < end[i], CHK_ERR_INVALID, "query interval start must be less than end");
        CHK_REQUIRE(end[i] <= 0x7FFFFFFFu, CHK_ERR_INVALID,
                    "query coordinate exceeds 2^31-1 (device coordinates are int32)");
    }
    const uint64_t NG = ix.ngroups;
    CHK_REQUIRE(NG <= MAXI, CHK_ERR_INVALID, "too many DB file groups for chk_sweep_noalloc");
    // intervals per work item: the sweep's own rule (enough items to fill
    // the device, at most 256 intervals each), raised so the items fit
    uint64_t spi = std::max<uint64_t>(4, std::min<uint64_t>(256, (n * NG + ix.ncu * 16 - 1) /
                                                                     ((uint64_t)ix.ncu * 16)));
    while (NG * ((n + spi - 1) / spi) > MAXI) spi *= 2;
    const uint32_t nchunk = (uint32_t)((n + spi - 1) / spi);
    const uint32_t nitems = (uint32_t)(NG * nchunk);

    unsigned char* base = scratch_ptr(ix.device);
    std::lock_guard<std::mutex> g(g_dev_mu[ix.device]);
    hipStream_t st = ix.stream;
    uint32_t* d_shard = reinterpret_cast<uint32_t*>(base);
    uint32_t* d_start = reinterpret_cast<uint32_t*>(base + 4 * MAXN);
    uint32_t* d_end = reinterpret_cast<uint32_t*>(base + 8 * MAXN);
    SparseItem* d_items = reinterpret_cast<SparseItem*>(base + ITEMS_OFF);
    zero3_u64(d_counts, ix.nfiles, d_bp, ix.nfiles, nullptr, 0, st);
    if (n) {
        CHK_HIP(hipMemcpyAsync(d_shard, shard, n * 4, hipMemcpyHostToDevice, st));
        CHK_HIP(hipMemcpyAsync(d_start, start, n * 4, hipMemcpyHostToDevice, st));
        CHK_HIP(hipMemcpyAsync(d_end, end, n * 4, hipMemcpyHostToDevice, st));
        hipLaunchKernelGGL(noalloc_items, dim3((nitems + 255) / 256), dim3(256), 0, st, d_items,
                           nitems, nchunk, (uint32_t)spi, (uint32_t)n);
        CHK_HIP(hipGetLastError());
        sparse_launch(ix, d_items, nitems, d_shard, d_start, d_end, d_counts, d_bp, st);
    }
    // the host arrays are read and the scratch is free again once this returns
    CHK_HIP(hipStreamSynchronize(st));
}

}  // namespace chk
