// util.hip -- a plain device copy used to calibrate hardware byte counters
// (FETCH_SIZE / WRITE_SIZE) against a known number of bytes.
#include "engine.hpp"

namespace chk {
namespace {
__global__ __launch_bounds__(256) void calib_copy_kernel(const uint4* __restrict__ src,
                                                         uint4* __restrict__ dst, size_t n) {
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) dst[i] = src[i];
}
}  // namespace

void calibrate_copy(Index& ix, uint64_t bytes) {
    size_t n = bytes / sizeof(uint4);
    DevBuf<uint4> a, b;
    a.reserve(n);
    b.reserve(n);
    CHK_HIP(hipMemsetAsync(a.p, 1, n * sizeof(uint4), ix.stream));
    hipEvent_t t0 = ix.prof.begin("calib_copy", ix.stream);
    hipLaunchKernelGGL(calib_copy_kernel, dim3(ix.ncu * 8), dim3(256), 0, ix.stream, a.p, b.p, n);
    CHK_HIP(hipGetLastError());
    ix.prof.end("calib_copy", t0, ix.stream);
    CHK_HIP(hipStreamSynchronize(ix.stream));
}
}  // namespace chk
