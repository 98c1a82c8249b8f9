// Kernels of the power-method stable time increment (the reference's
// PowerMethodProcessor, include/PowerMethodProcessor.h:150-470): per-node max edge
// length, the seeded start vector, the perturbation u + v l eps, the next iterate
// (F(u + v l eps) - F(u)) / (l eps m) and its L2 norm (deterministic block partial
// sums, added in a fixed order by one workgroup).  The driver is in capi.hip.
#include <cmath>

#include "internal.hpp"

namespace aphip {

namespace {

constexpr int kBlock = 256;

__device__ __forceinline__ uint64_t mix64(uint64_t x) {  // splitmix64 finalizer
    x += 0x9e3779b97f4a7c15ull;
    x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ull;
    x = (x ^ (x >> 27)) * 0x94d049bb133111ebull;
    return x ^ (x >> 31);
}

// l_i = max |x_i - x_j| over the nodes j of the elements around node i
// (src/MaxEdgeLengthProcessor.cpp:36-60, every node active).
__global__ __launch_bounds__(kBlock) void k_max_edge_any(int64_t nn, const int32_t *node_start,
                                                         const int32_t *node_inc, const int32_t *conn, int npe,
                                                         const double *X, double *l) {
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= nn) return;
    const double xi = X[3 * i], yi = X[3 * i + 1], zi = X[3 * i + 2];
    double m2 = 0.0;
    for (int32_t t = node_start[i]; t < node_start[i + 1]; ++t) {
        const int64_t e = node_inc[t] / npe;
        for (int k = 0; k < npe; ++k) {
            const int64_t j = conn[e * npe + k];
            const double dx = X[3 * j] - xi, dy = X[3 * j + 1] - yi, dz = X[3 * j + 2] - zi;
            m2 = fmax(m2, __fma_rn(dz, dz, __fma_rn(dy, dy, dx * dx)));
        }
    }
    l[i] = sqrt(m2);
}

// v_i = (2 U - 1) eps l for the DOF of caller numbering 3 id + d (InitializeEigenvector,
// :195-221, with the random start of the constructor, :156-157).
__global__ __launch_bounds__(kBlock) void k_pm_start(int64_t nn, const int32_t *caller_id, uint64_t seed,
                                                     double eps, const double *l, double *v) {
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= nn) return;
    for (int d = 0; d < 3; ++d) {
        const uint64_t key = seed ^ (uint64_t)(3 * (int64_t)caller_id[i] + d);
        const double r = (double)(mix64(key) >> 11) * 0x1.0p-53;  // [0, 1)
        v[3 * i + d] = (2.0 * r - 1.0) * (eps * l[i]);
    }
}

// u_eps = u_in + v l eps on free DOFs, u_in on fixed ones (PerturbDisplacementCoefficients, :230-260).
__global__ __launch_bounds__(kBlock) void k_pm_perturb(int64_t nn, const double *u_in, const double *v,
                                                       const double *l, const uint8_t *fixed, double eps, double *u) {
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= nn) return;
    const unsigned fx = fixed ? fixed[i] : 0u;
    for (int d = 0; d < 3; ++d) {
        const double ui = u_in[3 * i + d];
        u[3 * i + d] = (fx & (1u << d)) ? ui : __fma_rn(v[3 * i + d] * l[i], eps, ui);
    }
}

// v = (F1 - F0) / (l eps m) on free DOFs, 0 on fixed ones (ComputeNextEigenvector, :262-330);
// each block writes the sum of squares of its DOFs (tree order fixed by the block size).
__global__ __launch_bounds__(kBlock) void k_pm_next(int64_t nn, const double *f1, const double *f0,
                                                    const double *mass, const double *l, const uint8_t *fixed,
                                                    double eps, double *v, double *partial) {
    __shared__ double red[kBlock];
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    double s = 0.0;
    if (i < nn) {
        const unsigned fx = fixed ? fixed[i] : 0u;
        const double le = l[i] * eps;
        for (int d = 0; d < 3; ++d) {
            const double w = (fx & (1u << d)) ? 0.0 : (f1[3 * i + d] - f0[3 * i + d]) / (le * mass[i]);
            v[3 * i + d] = w;
            s = __fma_rn(w, w, s);
        }
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = kBlock / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

// partial[nb] = the sum of partial[0 .. nb-1] in a fixed order: lane t of one workgroup adds
// partial[t], partial[t + kSumLanes], ... in turn, then a tree over the lanes
constexpr int kSumLanes = 1024;
__global__ __launch_bounds__(kSumLanes) void k_pm_sum(int64_t nb, double *partial) {
    __shared__ double red[kSumLanes];
    double s = 0.0;
    for (int64_t b = threadIdx.x; b < nb; b += kSumLanes) s += partial[b];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = kSumLanes / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[nb] = red[0];
}

__global__ __launch_bounds__(kBlock) void k_pm_scale(int64_t n, double norm, double *v) {
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i < n) v[i] /= norm;
}

inline dim3 blocks(int64_t n) { return dim3((unsigned)((n + kBlock - 1) / kBlock)); }

}  // namespace

void pm_max_edge(Ctx &c, double *l) {
    hipLaunchKernelGGL(k_max_edge_any, blocks(c.nn), dim3(kBlock), 0, c.stream, c.nn, c.node_start.p, c.node_inc.p,
                       c.conn.p, c.npe, c.X.p, l);
    APHIP_HIP(hipGetLastError());
}

void pm_start(Ctx &c, const int32_t *caller_id, uint64_t seed, double eps, const double *l, double *v) {
    hipLaunchKernelGGL(k_pm_start, blocks(c.nn), dim3(kBlock), 0, c.stream, c.nn, caller_id, seed, eps, l, v);
    APHIP_HIP(hipGetLastError());
}

void pm_perturb(Ctx &c, const double *u_in, const double *v, const double *l, double eps) {
    hipLaunchKernelGGL(k_pm_perturb, blocks(c.nn), dim3(kBlock), 0, c.stream, c.nn, u_in, v, l,
                       c.has_fixed ? c.fixed.p : nullptr, eps, c.u.p);
    APHIP_HIP(hipGetLastError());
}

int64_t pm_partials(const Ctx &c) { return (int64_t)blocks(c.nn).x + 1; }

double pm_next(Ctx &c, const double *f0, const double *l, double eps, double *v, double *partial) {
    const dim3 g = blocks(c.nn);
    hipLaunchKernelGGL(k_pm_next, g, dim3(kBlock), 0, c.stream, c.nn, c.f.p, f0, c.mass.p, l,
                       c.has_fixed ? c.fixed.p : nullptr, eps, v, partial);
    APHIP_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_pm_sum, dim3(1), dim3(kSumLanes), 0, c.stream, (int64_t)g.x, partial);
    APHIP_HIP(hipGetLastError());
    double s = 0.0;
    APHIP_HIP(hipMemcpyAsync(&s, partial + g.x, sizeof(double), hipMemcpyDeviceToHost, c.stream));
    APHIP_HIP(hipStreamSynchronize(c.stream));
    return sqrt(s);
}

void pm_scale(Ctx &c, double norm, double *v) {
    hipLaunchKernelGGL(k_pm_scale, blocks(3 * c.nn), dim3(kBlock), 0, c.stream, 3 * c.nn, norm, v);
    APHIP_HIP(hipGetLastError());
}

}  // namespace aphip
