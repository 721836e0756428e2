// sort.hpp -- device-wide primitives written for gfx950: stable LSD radix sort
// of 64-bit keys (optional 32-bit payload) and exclusive scans.
#pragma once

#include "common.hpp"

namespace chk {

// Scratch reused across calls on one stream.
struct SortScratch {
    DevBuf<uint64_t> keys_alt;
    DevBuf<uint32_t> vals_alt;
    DevBuf<uint32_t> hist;
    DevBuf<uint32_t> sums;
    DevBuf<int64_t> sums64;
};

// Sorts keys[0..n) (and vals, when non-null) by bits [0, nbits) of the key,
// stably.  Results end in keys / vals.
void radix_sort_u64(uint64_t* keys, uint32_t* vals, size_t n, uint32_t nbits, SortScratch& s,
                    hipStream_t st);

// Events of n query intervals sorted by (segment, 2x + b), b = 1 for the
// start event at x = start[i], b = 0 for the end event at x = end[i], with
// segid[i] < 2^seg_bits and 2 end[i] < 2^coord_bits.  Writes, for the m = 2n
// events in sorted order, key32 = 2x + b, then key32[m] = 0xFFFFFFFF, and the
// exclusive prefix sums of rs = b and w = (b ? x : -x) with the totals at [m]
// (all three arrays hold m + 1 elements).  keys (m elements) is scratch.
void sort_events(const uint32_t* segid, const uint32_t* start, const uint32_t* end, size_t n,
                 uint32_t seg_bits, uint32_t coord_bits, uint32_t* key32, uint32_t* rs,
                 int64_t* w, uint64_t* keys, SortScratch& s, hipStream_t st);

// Grows s to what sort_events needs for up to n intervals.
void sort_events_reserve(size_t n, SortScratch& s);

// In-place exclusive scans of data[0..n); data[n] receives the total, so the
// buffer must hold n+1 elements.
void exclusive_scan_u32(uint32_t* data, size_t n, SortScratch& s, hipStream_t st);
void exclusive_scan_i64(int64_t* data, size_t n, DevBuf<int64_t>& sums, hipStream_t st);

}  // namespace chk
