// common.hpp -- shared helpers of the Chuckle MI355X library: error state
// behind the C ABI, HIP call checking, device buffers.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <vector>

#include "chuckle_hip.h"

namespace chk {

// Error carried from the implementation to the C ABI boundary: every entry
// point catches it, stores the message in a thread-local and returns `code`.
struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

void set_last_error(const std::string& msg);

#define CHK_HIP(call)                                                                   \
    do {                                                                                \
        hipError_t e_ = (call);                                                         \
        if (e_ != hipSuccess)                                                           \
            throw ::chk::Error(CHK_ERR_HIP, std::string(#call " failed: ") +            \
                                                hipGetErrorString(e_) + " at " +        \
                                                __FILE__ + ":" + std::to_string(__LINE__)); \
    } while (0)

#define CHK_REQUIRE(cond, code, msg)                          \
    do {                                                      \
        if (!(cond)) throw ::chk::Error((code), (msg));       \
    } while (0)

// Owning device allocation (hipMalloc'd); grows on demand, never shrinks.
// The first allocation is exact; a buffer that has to grow again grows to at
// least 1.5x its capacity, so buffers reused across batches of varying size
// (query and serving buffers) reallocate -- a device-synchronising free --
// only a few times.
template <typename T>
struct DevBuf {
    T* p = nullptr;
    size_t n = 0;  // capacity in elements
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { release(); }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    void reserve(size_t count) {
        if (count <= n) return;
        if (n) count = std::max(count, n + n / 2);
        release();
        size_t bytes = (count ? count : 1) * sizeof(T);
        hipError_t e = hipMalloc(&p, bytes);
        if (e != hipSuccess) {
            p = nullptr;
            (void)hipGetLastError();
            throw Error(CHK_ERR_NOMEM, "hipMalloc of " + std::to_string(bytes) +
                                           " bytes failed: " + hipGetErrorString(e));
        }
        n = count;
    }
    size_t bytes() const { return n * sizeof(T); }
};

// Pinned host allocation (hipHostMalloc'd); grows on demand, never shrinks.
template <typename T>
struct PinnedBuf {
    T* p = nullptr;
    size_t n = 0;
    PinnedBuf() = default;
    PinnedBuf(const PinnedBuf&) = delete;
    PinnedBuf& operator=(const PinnedBuf&) = delete;
    void reserve(size_t count) {
        if (count <= n) return;
        if (n) count = std::max(count, n + n / 2);
        if (p) (void)hipHostFree(p);
        p = nullptr;
        n = 0;
        hipError_t e = hipHostMalloc((void**)&p, (count ? count : 1) * sizeof(T), hipHostMallocDefault);
        if (e != hipSuccess) {
            p = nullptr;
            (void)hipGetLastError();
            throw Error(CHK_ERR_NOMEM, std::string("hipHostMalloc failed: ") + hipGetErrorString(e));
        }
        n = count;
    }
    ~PinnedBuf() {
        if (p) (void)hipHostFree(p);
    }
};

inline int num_cus(int device) {
    hipDeviceProp_t prop;
    CHK_HIP(hipGetDeviceProperties(&prop, device));
    return prop.multiProcessorCount;
}

inline uint32_t bits_for(uint64_t v) {  // number of bits needed to hold values 0..v
    uint32_t b = 0;
    while (b < 64 && (v >> b)) b++;
    return b;
}

}  // namespace chk
