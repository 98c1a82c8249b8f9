// Field transfer between caller-owned device arrays (caller numbering, the reference's
// field layout: node vectors n x w row-major, element tensors n_elem x 9 row-major,
// element state n_elem x w) and the context's internal arrays (Morton numbering; node
// vectors AoS, tensors SoA 9 x n_elem).  One thread per internal entity; the internal
// side is read or written contiguously, the caller side through the permutation.
#include "internal.hpp"

namespace aphip {

namespace {

constexpr int kBlock = 256;

inline dim3 blocks(int64_t n) { return dim3((unsigned)((n + kBlock - 1) / kBlock)); }

// dst[i, :] = src[iperm[i], :]  (caller -> internal, w values per entity)
__global__ __launch_bounds__(kBlock) void k_gather_rows(int64_t n, int w, const int32_t *iperm, const double *src,
                                                        double *dst) {
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const int64_t s = (int64_t)iperm[i] * w;
    for (int d = 0; d < w; ++d) dst[i * w + d] = src[s + d];
}

// dst[iperm[i], :] = src[i, :]  (internal -> caller)
__global__ __launch_bounds__(kBlock) void k_scatter_rows(int64_t n, int w, const int32_t *iperm, const double *src,
                                                         double *dst) {
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const int64_t t = (int64_t)iperm[i] * w;
    for (int d = 0; d < w; ++d) dst[t + d] = src[i * w + d];
}

// dst[iperm[e], k] = planes[k, e]  (SoA tensor -> caller row-major)
__global__ __launch_bounds__(kBlock) void k_scatter_tensor(int64_t ne, const int32_t *iperm, const double *planes,
                                                           double *dst) {
    const int64_t e = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (e >= ne) return;
    const int64_t t = (int64_t)iperm[e] * 9;
    for (int k = 0; k < 9; ++k) dst[t + k] = planes[k * ne + e];
}

// dst[i] = src[iperm[i]] and a count of the values that are not positive (nodal mass)
__global__ __launch_bounds__(kBlock) void k_gather_positive(int64_t n, const int32_t *iperm, const double *src,
                                                            double *dst, int32_t *n_bad) {
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const double x = src[iperm[i]];
    dst[i] = x;
    if (!(x > 0.0)) atomicAdd(n_bad, 1);
}

}  // namespace

void field_gather_rows(Ctx &c, int64_t n, int w, const int32_t *iperm, const double *src, double *dst) {
    if (n == 0) return;
    hipLaunchKernelGGL(k_gather_rows, blocks(n), dim3(kBlock), 0, c.stream, n, w, iperm, src, dst);
    APHIP_HIP(hipGetLastError());
}

void field_scatter_rows(Ctx &c, int64_t n, int w, const int32_t *iperm, const double *src, double *dst) {
    if (n == 0) return;
    hipLaunchKernelGGL(k_scatter_rows, blocks(n), dim3(kBlock), 0, c.stream, n, w, iperm, src, dst);
    APHIP_HIP(hipGetLastError());
}

void field_scatter_tensor(Ctx &c, const double *planes, double *dst) {
    if (c.ne == 0) return;
    hipLaunchKernelGGL(k_scatter_tensor, blocks(c.ne), dim3(kBlock), 0, c.stream, c.ne, c.d_elem_iperm.p, planes,
                       dst);
    APHIP_HIP(hipGetLastError());
}

int32_t field_gather_mass(Ctx &c, const double *src) {
    c.d_count.zero(c.stream);
    hipLaunchKernelGGL(k_gather_positive, blocks(c.nn), dim3(kBlock), 0, c.stream, c.nn, c.d_node_iperm.p, src,
                       c.mass.p, c.d_count.p);
    APHIP_HIP(hipGetLastError());
    int32_t bad = 0;
    c.d_count.download(&bad, 1, c.stream);
    APHIP_HIP(hipStreamSynchronize(c.stream));
    return bad;
}

}  // namespace aphip
