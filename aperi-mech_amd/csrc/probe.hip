// Device-to-device copy of a known size: the yardstick for HBM measurements
// (achievable bandwidth, and the scale of the FETCH_SIZE / WRITE_SIZE counters
// against bytes that are known to move).  Each thread moves 16-byte words.
#include "internal.hpp"

namespace aphip {

namespace {

__global__ __launch_bounds__(256) void k_stream_copy(const double2 *__restrict__ src, double2 *__restrict__ dst,
                                                     int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) dst[i] = src[i];
}

__global__ __launch_bounds__(256) void k_fill(double2 *p, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) p[i] = make_double2((double)i, -(double)i);
}

}  // namespace

double bandwidth_probe(int device, int64_t nbytes, int reps) {
    APHIP_REQUIRE(nbytes >= 16 && nbytes % 16 == 0, "probe size must be a positive multiple of 16 bytes");
    APHIP_REQUIRE(reps >= 1, "probe needs at least one repetition");
    APHIP_HIP(hipSetDevice(device));
    const int64_t n = nbytes / 16;
    DevBuf<double2> src, dst;
    src.alloc((size_t)n);
    dst.alloc((size_t)n);
    hipStream_t s;
    APHIP_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    hipEvent_t e0, e1;
    APHIP_HIP(hipEventCreate(&e0));
    APHIP_HIP(hipEventCreate(&e1));
    const dim3 grid((unsigned)((n + 255) / 256));
    hipLaunchKernelGGL(k_fill, grid, dim3(256), 0, s, src.p, n);
    APHIP_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_stream_copy, grid, dim3(256), 0, s, src.p, dst.p, n);  // warm-up
    APHIP_HIP(hipGetLastError());
    APHIP_HIP(hipEventRecord(e0, s));
    for (int r = 0; r < reps; ++r) hipLaunchKernelGGL(k_stream_copy, grid, dim3(256), 0, s, src.p, dst.p, n);
    APHIP_HIP(hipGetLastError());
    APHIP_HIP(hipEventRecord(e1, s));
    APHIP_HIP(hipEventSynchronize(e1));
    float ms = 0.0f;
    APHIP_HIP(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipStreamDestroy(s);
    return (double)ms / reps;
}

}  // namespace aphip
