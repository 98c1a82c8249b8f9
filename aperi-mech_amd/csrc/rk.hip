// RKPM on gfx950: reproducing-kernel functions with element-as-cell strain smoothing
// over a Tet4 or Hex8 background mesh (one- and two-pass force), and finite-element
// strain smoothing on Tet4 (the same cells with the element's own shape functions).
//
// Set-up (once, Total Lagrangian):
//   max edge length     src/MaxEdgeLengthProcessor.cpp:36-60
//   kernel radius       src/NeighborSearchProcessor.cpp:154-185 (r = scale * l; scale 1 -> 1 + 1e-6)
//   ball search         include/NeighborSearchProcessor.h:196-317, src/NeighborSearchProcessor.cpp:185-240,
//                       include/NeighborSelectorFunctor.h:63-99 — here a uniform cell list (cell edge =
//                       largest radius) replaces STK's LBVH; each point keeps its 40 nearest accepted
//                       candidates by (tweaked d^2, id) in LDS
//   RK functions        include/FunctionCreationProcessor.h:85-130,
//                       include/ShapeFunctionsFunctorReproducingKernel.h:190-257 (4x4 moment matrix
//                       inverted by complete-pivoting LU in registers)
//   cell node lists     include/SmoothedCellDataProcessor.h:407-539 (first-seen union)
//   cell derivatives    include/SmoothedCellDataProcessor.h:542-675, include/QuadratureSmoothed.h:32-62
//   RK mass             src/MassUtils.cpp:79-112 via FunctionEvaluationProcessor::ScatterValues
//   two-pass cells      include/SmoothedCellDataProcessor.h:474-486, :637-648 (the element's own
//                       nodes, derivative (B_k V)/V)
// Per step:
//   cell kernel         include/ComputeInternalForceSmoothedCell.h:86-117, :186-224 (H = sum u (x) grad phi,
//                       stress; the stress term -V P^T of :325-326 is kept per cell)
//   node kernel         :328-349 as a node gather over the cell entries that name the node, in ascending
//                       (cell, entry) order, fused with the central-difference node update
// Nodal smoothing cells on a thex mesh (each tet split into four hexes):
//   active nodes        include/MeshLabelerProcessor.h:227-330 (each hex's node of lowest
//                       caller id; exactly one per hex), cells and subcells :690-827 (one cell
//                       per active node: the hexes around it; subdomains 0: every hex a
//                       subcell, 1: the cell one subcell); internal nodes are renumbered with
//                       the active ones first, so nodes with degrees of freedom are a prefix
//   max edge length     edges to inactive nodes doubled (src/MaxEdgeLengthProcessor.cpp:54-55)
//   search              query points: every node; candidates: the active nodes
//   subcell CSR         include/SmoothedCellDataProcessor.h:407-675 over the hexes of a subcell
//   F-bar force         include/ComputeInternalForceSmoothedCell.h:120-413 (subcell J, cell
//                       J-bar, stress of F-bar, dual pressure, un-barred stress), used when a
//                       cell has more subcells than one (:458-459)
// Per step, two-pass method (src/Solvers/Explicit.cpp:148-186):
//   gather kernel       FunctionEvaluationProcessor::ComputeValues (include/FunctionEvaluationProcessor.h:
//                       70-117): physical displacement = Phi d before the cell kernel
//   scatter kernel      ScatterValues (:130-181) as a gather over reverse neighbor lists (points that
//                       list the node, ascending), fused with the central-difference node update
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <numeric>
#include <thread>
#include <vector>

#include "element.hpp"
#include "internal.hpp"

namespace aphip {

void stable_sort_pairs(Ctx &c, const int32_t *keys_in, int32_t *keys_out, const int32_t *vals_in, int32_t *vals_out,
                       int64_t n, int key_bits);

namespace {

constexpr int kMaxNbr = 40;  // include/Constants.h:15
constexpr int kSearchBlock = 64;
constexpr int kCellBlock = 64;

__device__ __forceinline__ void ld3(const double *p, int64_t i, double *o) {
    o[0] = p[3 * i];
    o[1] = p[3 * i + 1];
    o[2] = p[3 * i + 2];
}

__device__ __forceinline__ void st3(double *p, int64_t i, const double *o) {
    p[3 * i] = o[0];
    p[3 * i + 1] = o[1];
    p[3 * i + 2] = o[2];
}

// |a - b|^2 with a fixed fused order (the oracle evaluates the same expression)
__device__ __forceinline__ double dist2(const double *a, const double *b) {
    const double dx = a[0] - b[0], dy = a[1] - b[1], dz = a[2] - b[2];
    return __fma_rn(dz, dz, __fma_rn(dy, dy, dx * dx));
}

template <int NPE>
__global__ __launch_bounds__(256) void k_max_edge(int64_t nn, const int32_t *node_start, const int32_t *node_inc,
                                                  const int32_t *conn, const double *X, double scale, int64_t n_active,
                                                  double *radius) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= nn) return;
    double xi[3];
    ld3(X, i, xi);
    double l = 0.0;
    for (int32_t t = node_start[i]; t < node_start[i + 1]; ++t) {
        const int64_t e = node_inc[t] / NPE;
#pragma unroll
        for (int k = 0; k < NPE; ++k) {
            double xj[3];
            const int32_t j = conn[NPE * e + k];
            ld3(X, j, xj);
            // an edge to a node without degrees of freedom counts twice (nodes j >= n_active)
            l = fmax(l, sqrt(dist2(xi, xj)) * (j >= n_active ? 2.0 : 1.0));
        }
    }
    radius[i] = l * scale;
}

struct Grid {
    double lo[3];
    double cell;
    int nb[3];
};

__device__ __forceinline__ int bin_coord(double x, double lo, double cell, int nb) {
    int b = (int)floor((x - lo) / cell);
    return b < 0 ? 0 : (b >= nb ? nb - 1 : b);
}

__global__ __launch_bounds__(256) void k_bin_count(int64_t nn, const double *X, Grid G, int32_t *bin_of,
                                                   int32_t *count) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= nn) return;
    const int bx = bin_coord(X[3 * i], G.lo[0], G.cell, G.nb[0]);
    const int by = bin_coord(X[3 * i + 1], G.lo[1], G.cell, G.nb[1]);
    const int bz = bin_coord(X[3 * i + 2], G.lo[2], G.cell, G.nb[2]);
    const int b = (bz * G.nb[1] + by) * G.nb[0] + bx;
    bin_of[i] = b;
    atomicAdd(&count[b], 1);
}

// order inside a bin is irrelevant: accepted candidates are ranked by (d^2, id)
__global__ __launch_bounds__(256) void k_bin_fill(int64_t nn, const int32_t *bin_of, const int32_t *bstart,
                                                  int32_t *fill, int32_t *items) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= nn) return;
    const int b = bin_of[i];
    items[bstart[b] + atomicAdd(&fill[b], 1)] = (int32_t)i;
}

struct SearchArgs {
    int64_t nn;
    const double *X;
    const double *radius;
    const int32_t *caller_id;
    Grid G;
    const int32_t *bstart;
    const int32_t *items;
    int32_t *num;
    int32_t *nbr;
    int32_t *ncand;
};

__device__ __forceinline__ bool before(double da, int ia, double db, int ib) {
    return da < db || (da == db && ia < ib);
}

__global__ __launch_bounds__(kSearchBlock) void k_search(SearchArgs A) {
    __shared__ double sd[kMaxNbr * kSearchBlock];
    __shared__ int32_t sk[kMaxNbr * kSearchBlock];  // ranking key: caller id
    __shared__ int32_t sn[kMaxNbr * kSearchBlock];  // internal id
    const int tid = threadIdx.x;
    const int64_t i = (int64_t)blockIdx.x * kSearchBlock + tid;
    if (i >= A.nn) return;
    double xi[3];
    ld3(A.X, i, xi);
    const int bx = bin_coord(xi[0], A.G.lo[0], A.G.cell, A.G.nb[0]);
    const int by = bin_coord(xi[1], A.G.lo[1], A.G.cell, A.G.nb[1]);
    const int bz = bin_coord(xi[2], A.G.lo[2], A.G.cell, A.G.nb[2]);
    int cnt = 0, seen = 0;
    for (int z = bz - 1; z <= bz + 1; ++z) {
        if (z < 0 || z >= A.G.nb[2]) continue;
        for (int y = by - 1; y <= by + 1; ++y) {
            if (y < 0 || y >= A.G.nb[1]) continue;
            for (int x = bx - 1; x <= bx + 1; ++x) {
                if (x < 0 || x >= A.G.nb[0]) continue;
                const int b = (z * A.G.nb[1] + y) * A.G.nb[0] + x;
                for (int p = A.bstart[b]; p < A.bstart[b + 1]; ++p) {
                    const int32_t j = A.items[p];
                    double xj[3];
                    ld3(A.X, j, xj);
                    const double rj = A.radius[j];
                    // the sphere of neighbor j contains point i
                    if (!(dist2(xj, xi) <= rj * rj)) continue;
                    ++seen;
                    // ranking distance with the neighbor coordinates tweaked by (1+1e-6, 1+1e-7, 1+1e-8)
                    const double tw[3] = {xj[0] * (1 + 1.0e-6), xj[1] * (1 + 1.0e-7), xj[2] * (1 + 1.0e-8)};
                    const double d2 = dist2(tw, xi);
                    const int32_t key = A.caller_id[j];
                    if (cnt == kMaxNbr && !before(d2, key, sd[(kMaxNbr - 1) * kSearchBlock + tid],
                                                  sk[(kMaxNbr - 1) * kSearchBlock + tid]))
                        continue;
                    int pos = cnt < kMaxNbr ? cnt : kMaxNbr - 1;
                    while (pos > 0 && before(d2, key, sd[(pos - 1) * kSearchBlock + tid], sk[(pos - 1) * kSearchBlock + tid])) {
                        sd[pos * kSearchBlock + tid] = sd[(pos - 1) * kSearchBlock + tid];
                        sk[pos * kSearchBlock + tid] = sk[(pos - 1) * kSearchBlock + tid];
                        sn[pos * kSearchBlock + tid] = sn[(pos - 1) * kSearchBlock + tid];
                        --pos;
                    }
                    sd[pos * kSearchBlock + tid] = d2;
                    sk[pos * kSearchBlock + tid] = key;
                    sn[pos * kSearchBlock + tid] = j;
                    if (cnt < kMaxNbr) ++cnt;
                }
            }
        }
    }
    A.num[i] = cnt;
    A.ncand[i] = seen;
    for (int k = 0; k < kMaxNbr; ++k) A.nbr[i * kMaxNbr + k] = k < cnt ? sn[k * kSearchBlock + tid] : -1;
}

// Cubic B-spline kernel, include/MathUtils.h:321-330
__device__ __forceinline__ double kernel_value(const double *d, double R) {
    const double q = sqrt(__fma_rn(d[2], d[2], __fma_rn(d[1], d[1], d[0] * d[0]))) / R;
    if (q < 0.5) return 1.0 + 6.0 * q * q * (-1.0 + q);
    if (q < 1.0) return 2.0 * (1.0 + q * (-3.0 + 3.0 * q - 1.0 * q * q));
    return 0.0;
}

__device__ __forceinline__ void swap_d(double &a, double &b) {
    const double t = a;
    a = b;
    b = t;
}

// Row 0 of the inverse of the symmetric 4x4 moment matrix by LU with complete
// pivoting (the published FullPivLU algorithm of Eigen 3.4, which InvertMatrix uses on
// CPU builds, include/MathUtils.h:307-319): first maximum in column-major scan,
// rank threshold |max pivot| * eps * 4, unknowns beyond the rank set to zero.
// All indices are compile-time after unrolling, so the matrix stays in registers.
// Returns the Frobenius norm of the full inverse.
__device__ double inverse4_row0(double (&a)[4][4], double (&row0)[4]) {
    int prow[4] = {0, 1, 2, 3}, pcol[4] = {0, 1, 2, 3};
    int nonzero = 4;
    double maxpiv = 0.0;
    bool stop = false;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        if (!stop) {
            double big = -1.0;
            int bi = k, bj = k;
#pragma unroll
            for (int j = k; j < 4; ++j)
#pragma unroll
                for (int i = k; i < 4; ++i) {
                    const double v = fabs(a[i][j]);
                    if (v > big) {
                        big = v;
                        bi = i;
                        bj = j;
                    }
                }
            if (big == 0.0) {
                nonzero = k;
                stop = true;
            } else {
                maxpiv = fmax(maxpiv, big);
#pragma unroll
                for (int r = k + 1; r < 4; ++r)
                    if (bi == r) {
#pragma unroll
                        for (int j = 0; j < 4; ++j) swap_d(a[k][j], a[r][j]);
                        const int tmp = prow[k];
                        prow[k] = prow[r];
                        prow[r] = tmp;
                    }
#pragma unroll
                for (int r = k + 1; r < 4; ++r)
                    if (bj == r) {
#pragma unroll
                        for (int i = 0; i < 4; ++i) swap_d(a[i][k], a[i][r]);
                        const int tmp = pcol[k];
                        pcol[k] = pcol[r];
                        pcol[r] = tmp;
                    }
                const double piv = a[k][k];
#pragma unroll
                for (int i = k + 1; i < 4; ++i) a[i][k] /= piv;
#pragma unroll
                for (int i = k + 1; i < 4; ++i)
#pragma unroll
                    for (int j = k + 1; j < 4; ++j) a[i][j] -= a[i][k] * a[k][j];
            }
        }
    }
    const double thr = maxpiv * (2.220446049250313e-16 * 4.0);
    int rank = 0;
#pragma unroll
    for (int i = 0; i < 4; ++i)
        if (i < nonzero && fabs(a[i][i]) > thr) ++rank;
    double fro = 0.0;
#pragma unroll
    for (int col = 0; col < 4; ++col) row0[col] = 0.0;
    if (rank == 0) return 0.0;
#pragma unroll
    for (int col = 0; col < 4; ++col) {
        double cv[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) cv[i] = (prow[i] == col) ? 1.0 : 0.0;
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j < i; ++j) cv[i] -= a[i][j] * cv[j];
#pragma unroll
        for (int i = 3; i >= 0; --i) {
            if (i < rank) {
#pragma unroll
                for (int j = i + 1; j < 4; ++j)
                    if (j < rank) cv[i] -= a[i][j] * cv[j];
                cv[i] /= a[i][i];
            }
        }
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const double x = i < rank ? cv[i] : 0.0;
            fro += x * x;
            if (pcol[i] == 0) row0[col] = x;
        }
    }
    return sqrt(fro);
}

__global__ __launch_bounds__(128) void k_rk_phi(int64_t nn, const double *X, const double *radius, const int32_t *num,
                                                const int32_t *nbr, double *phi) {
    const int64_t i = (int64_t)blockIdx.x * 128 + threadIdx.x;
    if (i >= nn) return;
    const int n = num[i];
    double *ph = phi + i * kMaxNbr;
    for (int k = 0; k < kMaxNbr; ++k) ph[k] = 0.0;
    if (n == 0) return;
    if (n == 1) {
        ph[0] = 1.0;
        return;
    }
    double xi[3];
    ld3(X, i, xi);
    double M[4][4] = {{0, 0, 0, 0}, {0, 0, 0, 0}, {0, 0, 0, 0}, {0, 0, 0, 0}};
    double wsum = 0.0;
    for (int k = 0; k < n; ++k) {
        const int32_t j = nbr[i * kMaxNbr + k];
        double xj[3];
        ld3(X, j, xj);
        const double d[3] = {xi[0] - xj[0], xi[1] - xj[1], xi[2] - xj[2]};
        const double w = kernel_value(d, radius[j]);
        if (w == 0.0) continue;
        const double Hb[4] = {1.0, d[0], d[1], d[2]};
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int b = 0; b < 4; ++b) M[a][b] += (w * Hb[a]) * Hb[b];
        wsum += w;
    }
    double r0[4];
    const double fro = inverse4_row0(M, r0);
    const bool shepard = fro < 1e-15;
    for (int k = 0; k < n; ++k) {
        const int32_t j = nbr[i * kMaxNbr + k];
        double xj[3];
        ld3(X, j, xj);
        const double d[3] = {xi[0] - xj[0], xi[1] - xj[1], xi[2] - xj[2]};
        const double w = kernel_value(d, radius[j]);
        if (shepard) {
            ph[k] = w / wsum;
        } else if (w != 0.0) {
            ph[k] = (r0[0] + r0[1] * d[0] + r0[2] * d[1] + r0[3] * d[2]) * w;
        }
    }
}

// First-seen union of the neighbor lists of a cell's NPE nodes, in LDS.
template <int NPE>
__device__ __forceinline__ int cell_union(const int32_t *en, const int32_t *num, const int32_t *nbr, int32_t *list,
                                          int stride) {
    int len = 0;
    for (int k = 0; k < NPE; ++k) {
        const int32_t a = en[k];
        const int na = num[a];
        for (int l = 0; l < na; ++l) {
            const int32_t j = nbr[(int64_t)a * kMaxNbr + l];
            bool found = false;
            for (int t = 0; t < len; ++t)
                if (list[t * stride] == j) {
                    found = true;
                    break;
                }
            if (!found) list[(len++) * stride] = j;
        }
    }
    return len;
}

// cells per block of the set-up kernels: each thread keeps its union list of up to
// NPE * kMaxNbr entries in LDS (40 KB per block either way)
template <int NPE>
constexpr int cell_block() { return NPE == 4 ? kCellBlock : kCellBlock / 2; }

template <int NPE>
__global__ __launch_bounds__(cell_block<NPE>()) void k_cell_count(int64_t ne, const int32_t *conn, const int32_t *num,
                                                                  const int32_t *nbr, int32_t *len_out) {
    constexpr int BLK = cell_block<NPE>();
    __shared__ int32_t list[NPE * kMaxNbr * BLK];
    const int64_t e = (int64_t)blockIdx.x * BLK + threadIdx.x;
    if (e >= ne) return;
    int32_t en[NPE];
#pragma unroll
    for (int k = 0; k < NPE; ++k) en[k] = conn[NPE * e + k];
    len_out[e] = cell_union<NPE>(en, num, nbr, list + threadIdx.x, BLK);
}

// Smoothed Tet4 B and volume from the faces (include/QuadratureSmoothed.h:32-62)
__host__ __device__ void smoothed_tet4(const double (&X)[4][3], double (&B)[4][3], double &V) {
    const int faces[4][3] = {{0, 1, 3}, {1, 2, 3}, {0, 3, 2}, {0, 2, 1}};
    double vol = 0.0;
#pragma unroll
    for (int k = 0; k < 4; ++k) B[k][0] = B[k][1] = B[k][2] = 0.0;
#pragma unroll
    for (int f = 0; f < 4; ++f) {
        double e1[3], e2[3], n[3];
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            e1[d] = X[faces[f][1]][d] - X[faces[f][0]][d];
            e2[d] = X[faces[f][2]][d] - X[faces[f][0]][d];
        }
        n[0] = (e1[1] * e2[2] - e1[2] * e2[1]) / 2.0;
        n[1] = (e1[2] * e2[0] - e1[0] * e2[2]) / 2.0;
        n[2] = (e1[0] * e2[1] - e1[1] * e2[0]) / 2.0;
        const double *x0 = X[faces[f][0]];
        vol += x0[0] * n[0] + x0[1] * n[1] + x0[2] * n[2];
#pragma unroll
        for (int l = 0; l < 3; ++l)
#pragma unroll
            for (int d = 0; d < 3; ++d) B[faces[f][l]][d] += n[d] / 3.0;
    }
    vol /= 3.0;
#pragma unroll
    for (int k = 0; k < 4; ++k)
#pragma unroll
        for (int d = 0; d < 3; ++d) B[k][d] /= vol;
    V = vol;
}

// Smoothed Hex8 B and volume from the faces, each split into four triangles about
// its centroid (include/QuadratureSmoothed.h:96-150)
__host__ __device__ void smoothed_hex8(const double (&X)[8][3], double (&B)[8][3], double &V) {
    const int faces[6][4] = {{0, 1, 5, 4}, {1, 2, 6, 5}, {2, 3, 7, 6}, {0, 4, 7, 3}, {0, 3, 2, 1}, {4, 5, 6, 7}};
    double vol = 0.0;
#pragma unroll
    for (int k = 0; k < 8; ++k) B[k][0] = B[k][1] = B[k][2] = 0.0;
#pragma unroll
    for (int f = 0; f < 6; ++f) {
        double mid[3] = {0.0, 0.0, 0.0};
#pragma unroll
        for (int k = 0; k < 4; ++k)
#pragma unroll
            for (int d = 0; d < 3; ++d) mid[d] += X[faces[f][k]][d];
#pragma unroll
        for (int d = 0; d < 3; ++d) mid[d] /= 4.0;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int km1 = (k + 3) & 3;
            double e1[3], e2[3], n[3];
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                e1[d] = X[faces[f][km1]][d] - mid[d];
                e2[d] = X[faces[f][k]][d] - mid[d];
            }
            n[0] = (e1[1] * e2[2] - e1[2] * e2[1]) / 2.0;
            n[1] = (e1[2] * e2[0] - e1[0] * e2[2]) / 2.0;
            n[2] = (e1[0] * e2[1] - e1[1] * e2[0]) / 2.0;
            vol += mid[0] * n[0] + mid[1] * n[1] + mid[2] * n[2];
#pragma unroll
            for (int l = 0; l < 4; ++l) {
                const double cw = (l == k || l == km1) ? 5.0 / 12.0 : 1.0 / 12.0;
#pragma unroll
                for (int d = 0; d < 3; ++d) B[faces[f][l]][d] += n[d] * cw;
            }
        }
    }
    vol /= 3.0;
#pragma unroll
    for (int k = 0; k < 8; ++k)
#pragma unroll
        for (int d = 0; d < 3; ++d) B[k][d] /= vol;
    V = vol;
}

template <int NPE>
__host__ __device__ __forceinline__ void smoothed_cell(const double (&X)[NPE][3], double (&B)[NPE][3], double &V) {
    if constexpr (NPE == 4) smoothed_tet4(X, B, V);
    else smoothed_hex8(X, B, V);
}

struct CellFillArgs {
    int64_t ne;
    const int32_t *conn;
    const double *X;
    const int32_t *num;
    const int32_t *nbr;
    const double *phi;
    const int64_t *start;
    int32_t *nodes;
    double *dphi;
    double *volume;
    int32_t *entry_cell;
};

template <int NPE>
__global__ __launch_bounds__(cell_block<NPE>()) void k_cell_fill(CellFillArgs A) {
    constexpr int BLK = cell_block<NPE>();
    __shared__ int32_t list[NPE * kMaxNbr * BLK];
    const int tid = threadIdx.x;
    const int64_t e = (int64_t)blockIdx.x * BLK + tid;
    if (e >= A.ne) return;
    int32_t en[NPE];
#pragma unroll
    for (int k = 0; k < NPE; ++k) en[k] = A.conn[NPE * e + k];
    int32_t *my = list + tid;
    const int len = cell_union<NPE>(en, A.num, A.nbr, my, BLK);
    const int64_t s0 = A.start[e];
    double X[NPE][3], B[NPE][3], V;
#pragma unroll
    for (int k = 0; k < NPE; ++k) ld3(A.X, en[k], X[k]);
    smoothed_cell<NPE>(X, B, V);
    double *dp = A.dphi + 3 * s0;
    for (int t = 0; t < len; ++t) {
        A.nodes[s0 + t] = my[t * BLK];
        A.entry_cell[s0 + t] = (int32_t)e;
        dp[3 * t] = dp[3 * t + 1] = dp[3 * t + 2] = 0.0;
    }
    // grad phi_J = sum_k B_k V phi_{k,J} / V_sub (SmoothedCellDataProcessor.h:623-662)
#pragma unroll
    for (int k = 0; k < NPE; ++k) {
        const int32_t a = en[k];
        const int na = A.num[a];
        for (int l = 0; l < na; ++l) {
            const int32_t j = A.nbr[(int64_t)a * kMaxNbr + l];
            const double fv = A.phi[(int64_t)a * kMaxNbr + l];
            int t = 0;
            while (my[t * BLK] != j) ++t;
#pragma unroll
            for (int m = 0; m < 3; ++m) dp[3 * t + m] += B[k][m] * V * fv;
        }
    }
    for (int t = 0; t < len; ++t) {
        dp[3 * t] /= V;
        dp[3 * t + 1] /= V;
        dp[3 * t + 2] /= V;
    }
    A.volume[e] = V;
}

// Two-pass cells: entry 4e + k is node k of element e with derivative (B_k V) / V.
template <int NPE>
__global__ __launch_bounds__(256) void k_cell_two_pass(int64_t ne, const int32_t *conn, const double *X,
                                                       int32_t *nodes, double *dphi, double *volume,
                                                       int32_t *entry_cell) {
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= ne) return;
    int32_t en[NPE];
#pragma unroll
    for (int k = 0; k < NPE; ++k) en[k] = conn[NPE * e + k];
    double Xe[NPE][3], B[NPE][3], V;
#pragma unroll
    for (int k = 0; k < NPE; ++k) ld3(X, en[k], Xe[k]);
    smoothed_cell<NPE>(Xe, B, V);
#pragma unroll
    for (int k = 0; k < NPE; ++k) {
        nodes[NPE * e + k] = en[k];
        entry_cell[NPE * e + k] = (int32_t)e;
#pragma unroll
        for (int m = 0; m < 3; ++m) dphi[3 * (NPE * e + k) + m] = (B[k][m] * V) / V;
    }
    volume[e] = V;
}

__global__ __launch_bounds__(256) void k_identity_functions(int64_t nn, int32_t *num, int32_t *nbr, double *phi) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= nn) return;
    num[i] = 1;
    // the whole row: the node itself with value 1, the unused entries -1 / 0
    for (int k = 0; k < kMaxNbr; ++k) {
        nbr[i * kMaxNbr + k] = k == 0 ? (int32_t)i : -1;
        phi[i * kMaxNbr + k] = k == 0 ? 1.0 : 0.0;
    }
}

__global__ __launch_bounds__(256) void k_iota(int64_t n, int32_t *v) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) v[i] = (int32_t)i;
}

// pos[perm[k]] = k
__global__ __launch_bounds__(256) void k_invert_perm(int64_t n, const int32_t *perm, int32_t *pos) {
    const int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (k < n) pos[perm[k]] = (int32_t)k;
}

__global__ __launch_bounds__(256) void k_count_keys(int64_t n, const int32_t *keys, int32_t *count) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) atomicAdd(&count[keys[i]], 1);
}

// ---- smoothed-cell force step --------------------------------------------
// Cell and node kernels of the smoothed-cell force.  Lanes in groups of 16 stride over
// a cell's (or node's) entries, accumulate their share of the sum and combine the 16
// partial sums with a xor butterfly; addition is commutative, so every lane ends with
// the same bits.  The one-pass force runs on cell tiles (below); the two-pass cells
// (four entries) keep their entry forces in node order (g) for the plain node sum.
constexpr int kCellLanes = 16;

struct CellForceArgs {
    int64_t ne, nblocks;
    const int64_t *start;
    const int32_t *nodes;
    const double *dphi;
    const double *volume;
    const double *u;
    double *H, *P;
    double *g;  // [nnz * 3] per-entry force contributions, node order
    const int32_t *pos;  // [nnz] node-order position of each entry
    const double *state_n;
    double *state_np1;
    MatParams mat;
    const double *elem_lambda, *elem_mu;
    const double *Hn, *Pn;  // rate-form materials: H and P of the previous step
    double rate_dt;
};

constexpr int kNodeBatch = 4;  // incidences per lane whose loads are issued together
// lanes per node of the two-pass gather (u = Phi d)
constexpr int kGsLanes = 8;
constexpr int kNumXcd = 8;

// Consecutive logical blocks go to the same XCD (the dispatcher places block b on
// XCD b % 8), so neighbouring cells / nodes share that XCD's L2 for their gathers.
// The grid is padded to a multiple of 8 so the map is a bijection.
__device__ __forceinline__ int64_t xcd_block(int64_t nblocks_padded) {
    const int64_t b = blockIdx.x;
    const int64_t per = nblocks_padded / kNumXcd;
    return (b % kNumXcd) * per + b / kNumXcd;
}

template <int LANES>
__device__ __forceinline__ double group_sum(double x) {
#pragma unroll
    for (int s = LANES / 2; s >= 1; s >>= 1) x += __shfl_xor(x, s, LANES);
    return x;
}

__device__ __forceinline__ double group_sum16(double x) { return group_sum<kCellLanes>(x); }

// Two-pass cells, one thread per cell: the element's four nodes, their derivative
// (B_k V) / V recomputed from the coordinates by the set-up's formula (instead of
// streaming the stored 96 B per cell), H, stress and the four entry forces
// -V P grad N_k at their node-order positions.
template <int KIND, int NPE>
__global__ __launch_bounds__(256) void k_rk_cells_two_pass(CellForceArgs A, const double *X) {
#pragma clang fp contract(off)
    const int64_t c_raw = xcd_block(A.nblocks) * 256 + threadIdx.x;
    const bool live = c_raw < A.ne;
    const int64_t c = live ? c_raw : A.ne - 1;
    int32_t en[NPE], ps[NPE];
#pragma unroll
    for (int k = 0; k < NPE; ++k) {
        en[k] = A.nodes[NPE * c + k];
        ps[k] = A.pos[NPE * c + k];
    }
    double Xe[NPE][3], uj[NPE][3];
#pragma unroll
    for (int k = 0; k < NPE; ++k) {
        ld3(X, en[k], Xe[k]);
        ld3(A.u, en[k], uj[k]);
    }
    double snp1 = 0.0, aux = 0.0;
    const double sn = load_state<KIND>(A.state_n, c), sn2 = load_state2<KIND>(A.state_n, c);
    MatParams m = A.mat;
    if (A.elem_lambda) {
        m.lambda = A.elem_lambda[c];
        m.mu = A.elem_mu[c];
    }
    double B[NPE][3], V;
    smoothed_cell<NPE>(Xe, B, V);
    double gc[NPE][3];
#pragma unroll
    for (int k = 0; k < NPE; ++k)
#pragma unroll
        for (int b = 0; b < 3; ++b) gc[k][b] = (B[k][b] * V) / V;
    double H[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0}, P[9];
#pragma unroll
    for (int k = 0; k < NPE; ++k)
#pragma unroll
        for (int a = 0; a < 3; ++a)
#pragma unroll
            for (int b = 0; b < 3; ++b) H[3 * a + b] = __fma_rn(uj[k][a], gc[k][b], H[3 * a + b]);
    RateIn rin;
    load_rate<KIND>(rin, A.Hn, A.Pn, A.ne, c, A.rate_dt);
    stress<KIND>(H, m, sn, snp1, aux, P, &rin, sn2);
    snp1 = cell_damage_marker<KIND>(snp1, m);
    if (!live) return;
    store_state<KIND>(A.state_np1, c, snp1, aux);
#pragma unroll
    for (int i = 0; i < 9; ++i) {
        store_once(A.H + i * A.ne + c, H[i]);
        store_once(A.P + i * A.ne + c, P[i]);
    }
    const double mv = -V;
#pragma unroll
    for (int k = 0; k < NPE; ++k) {
        double f[3];
#pragma unroll
        for (int a = 0; a < 3; ++a) f[a] = mv * dot3(P[3 * a], gc[k][0], P[3 * a + 1], gc[k][1], P[3 * a + 2], gc[k][2]);
        st3(A.g, ps[k], f);
    }
}

struct RkNodeArgs {
    int64_t nn, nblocks;
    const int64_t *inc_start;
    const double *g;  // node order
    const double *mass;
    const uint8_t *fixed;
    double *f, *a, *v;
    double dt;
};

// Node i's force f is complete: store it and finish the node (MODE 0: force only,
// 1: a = f/m, 2: central-difference corrector).
template <int MODE>
__device__ __forceinline__ void rk_finish(const RkNodeArgs &A, int64_t i, const double (&f)[3]) {
    A.f[3 * i] = f[0];
    A.f[3 * i + 1] = f[1];
    A.f[3 * i + 2] = f[2];
    if constexpr (MODE == 1) {
        const double m = A.mass[i];
        A.a[3 * i] = f[0] / m;
        A.a[3 * i + 1] = f[1] / m;
        A.a[3 * i + 2] = f[2] / m;
    } else if constexpr (MODE == 2) {
        double vh[3], a1[3], v1[3];
        ld3(A.v, i, vh);
        cd_correct(f, A.mass[i], A.fixed ? A.fixed[i] : 0u, A.dt, vh, a1, v1);
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            A.a[3 * i + d] = a1[d];
            A.v[3 * i + d] = v1[d];
        }
    }
}

template <int MODE>
__global__ __launch_bounds__(256) void k_rk_nodes(RkNodeArgs A) {
    const int l = threadIdx.x & (kCellLanes - 1);
    const int64_t i_raw = xcd_block(A.nblocks) * (256 / kCellLanes) + (threadIdx.x / kCellLanes);
    const bool live = i_raw < A.nn;
    const int64_t i = live ? i_raw : A.nn - 1;
    double f[3] = {0.0, 0.0, 0.0};
    const int64_t t0 = A.inc_start[i], t1 = A.inc_start[i + 1];
    // kNodeBatch incidences at a time, all loads issued before the first sum (clamped
    // indices, absent incidences add zero); ascending order.
    for (int64_t tb = t0 + l; tb < t1; tb += kNodeBatch * kCellLanes) {
        double g[kNodeBatch][3];
#pragma unroll
        for (int r = 0; r < kNodeBatch; ++r) ld3(A.g, min(tb + r * kCellLanes, t1 - 1), g[r]);
#pragma unroll
        for (int r = 0; r < kNodeBatch; ++r) {
            const bool in = tb + r * kCellLanes < t1;
#pragma unroll
            for (int d = 0; d < 3; ++d) f[d] += in ? g[r][d] : 0.0;
        }
    }
#pragma unroll
    for (int d = 0; d < 3; ++d) f[d] = group_sum16(f[d]);
    if (!live || l != 0) return;
    rk_finish<MODE>(A, i, f);
}

// ---- one-pass method on cell tiles -------------------------------------------
// A tile is a run of consecutive (Morton-ordered) cells; its node set is the sorted
// union of their node lists and every entry carries its node's index in that set
// (uint16).  One wave per tile:
//   A. stages u of the tile's nodes in LDS and clears one force row per tile node,
//   B. takes 64 / LPC cells at a time, LPC lanes per cell: each lane loads its
//      entries (lane j: j, j + LPC, ...; the first K stay in registers), sums its
//      share of H = sum_t u_node(t) (x) grad phi_t from the staged u, the group
//      combines the shares with a xor butterfly and every lane of the group
//      evaluates the stress; the lanes then add the entry forces -V P grad phi_t of
//      their registered entries into the LDS rows of the entries' nodes,
//   C. writes the tile's rows to its slots (tile-major, coalesced).
// Each entry's grad phi is read from HBM once per step: the reference's per-entry
// atomic scatter (include/ComputeInternalForceSmoothedCell.h:312-351) becomes LDS
// adds of one wave, in program order, and a node pass that sums each node's slots in
// ascending tile order and finishes the node.  Both are fixed orders, so the result
// is the same bits on every run.
struct RkTileArgs {
    int64_t ncells, ntiles, nblocks;
    int maxtn;
    const int64_t *start;
    const uint16_t *lidx;     // [nnz] entry -> index in its tile's node set
    const double *dphi;
    const double *volume;
    const double *u;
    const int32_t *t_nstart;  // [ntiles + 1] offsets of the tile node sets
    const int32_t *t_node;    // [total] internal node ids, ascending within a tile
    const int64_t *t_cstart;  // [ntiles + 1] first cell of each tile
    double *slots;            // [total * 3] tile force rows, tile-major
    double *H, *P;
    const double *state_n;
    double *state_np1;
    MatParams mat;
    const double *elem_lambda, *elem_mu;
    const double *Hn, *Pn;
    double rate_dt;
};

template <int KIND, int LPC, int K>
__global__ __launch_bounds__(64) void k_rk_tile(RkTileArgs A) {
#pragma clang fp contract(off)
    extern __shared__ __align__(16) double rk_lds[];
    double *U = rk_lds;
    double *F = rk_lds + 3 * (size_t)A.maxtn;
    const int64_t t = xcd_block(A.nblocks);
    if (t >= A.ntiles) return;
    const int lane = threadIdx.x;
    const int tb = A.t_nstart[t], tn = A.t_nstart[t + 1] - tb;
    // staging in batches of kStage rows per lane, every load of a batch issued before its
    // first LDS write (a load and its write per iteration would cost a latency each)
    constexpr int kStage = 8;
    for (int l0 = lane; l0 < tn; l0 += 64 * kStage) {
        int64_t gi[kStage];
#pragma unroll
        for (int r = 0; r < kStage; ++r) gi[r] = A.t_node[tb + min(l0 + 64 * r, tn - 1)];
        double ur[kStage][3];
#pragma unroll
        for (int r = 0; r < kStage; ++r) ld3(A.u, gi[r], ur[r]);
#pragma unroll
        for (int r = 0; r < kStage; ++r) {
            const int ln = l0 + 64 * r;
            if (ln < tn) {
#pragma unroll
                for (int d = 0; d < 3; ++d) {
                    U[3 * ln + d] = ur[r][d];
                    F[3 * ln + d] = 0.0;
                }
            }
        }
    }
    __syncthreads();
    constexpr int G = 64 / LPC;
    const int grp = lane / LPC, j = lane % LPC;
    const int64_t c0 = A.t_cstart[t], c1 = A.t_cstart[t + 1];
    for (int64_t cb = c0; cb < c1; cb += G) {
        const int64_t c_raw = cb + grp;
        const bool live = c_raw < c1;
        const int64_t c = live ? c_raw : c1 - 1;  // idle groups recompute the tile's last cell, store nothing
        const int64_t t0 = A.start[c], t1 = A.start[c + 1];
        double H[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0}, P[9];
        double gr[K][3] = {};
        int li[K] = {};
        if (t1 > t0) {
#pragma unroll
            for (int r = 0; r < K; ++r) {
                const int64_t e = min(t0 + j + r * LPC, t1 - 1);
                li[r] = A.lidx[e];
                ld3(A.dphi, e, gr[r]);
            }
#pragma unroll
            for (int r = 0; r < K; ++r) {
                const bool in = t0 + j + r * LPC < t1;
#pragma unroll
                for (int b = 0; b < 3; ++b) gr[r][b] = in ? gr[r][b] : 0.0;
                double uj[3];
                ld3(U, li[r], uj);
#pragma unroll
                for (int a = 0; a < 3; ++a)
#pragma unroll
                    for (int b = 0; b < 3; ++b) H[3 * a + b] = __fma_rn(uj[a], gr[r][b], H[3 * a + b]);
            }
        }
        for (int64_t e = t0 + j + K * LPC; e < t1; e += LPC) {  // entries past the registered ones
            double uj[3], g[3];
            ld3(U, A.lidx[e], uj);
            ld3(A.dphi, e, g);
#pragma unroll
            for (int a = 0; a < 3; ++a)
#pragma unroll
                for (int b = 0; b < 3; ++b) H[3 * a + b] = __fma_rn(uj[a], g[b], H[3 * a + b]);
        }
#pragma unroll
        for (int i = 0; i < 9; ++i) H[i] = group_sum<LPC>(H[i]);
        MatParams m = A.mat;
        if (A.elem_lambda) {
            m.lambda = A.elem_lambda[c];
            m.mu = A.elem_mu[c];
        }
        double snp1 = 0.0, aux = 0.0;
        const double sn = load_state<KIND>(A.state_n, c), sn2 = load_state2<KIND>(A.state_n, c);
        RateIn rin;
        load_rate<KIND>(rin, A.Hn, A.Pn, A.ncells, c, A.rate_dt);
        stress<KIND>(H, m, sn, snp1, aux, P, &rin, sn2);
        snp1 = cell_damage_marker<KIND>(snp1, m);
        if (live && j == 0) {
            store_state<KIND>(A.state_np1, c, snp1, aux);
            // plain stores: one lane per cell writes H and P, a sparse stream that
            // measured 2 % faster through the L2 than as non-temporal stores
#pragma unroll
            for (int i = 0; i < 9; ++i) {
                A.H[i * A.ncells + c] = H[i];
                A.P[i * A.ncells + c] = P[i];
            }
        }
        if (!live) continue;
        const double mv = -A.volume[c];
#pragma unroll
        for (int r = 0; r < K; ++r) {
            if (t0 + j + r * LPC < t1) {
#pragma unroll
                for (int a = 0; a < 3; ++a) {
                    const double fa = mv * dot3(P[3 * a], gr[r][0], P[3 * a + 1], gr[r][1], P[3 * a + 2], gr[r][2]);
                    __hip_atomic_fetch_add(F + 3 * li[r] + a, fa, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                }
            }
        }
        for (int64_t e = t0 + j + K * LPC; e < t1; e += LPC) {
            double g[3];
            ld3(A.dphi, e, g);
            const int lr = A.lidx[e];
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                const double fa = mv * dot3(P[3 * a], g[0], P[3 * a + 1], g[1], P[3 * a + 2], g[2]);
                __hip_atomic_fetch_add(F + 3 * lr + a, fa, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            }
        }
    }
    __syncthreads();
    double *dst = A.slots + 3 * (int64_t)tb;
    for (int k = lane; k < 3 * tn; k += 64) dst[k] = F[k];
}

// F-bar on cell tiles (ComputeInternalForceSmoothedCell.h:120-413): a tile holds whole
// nodal cells, so every pass of the F-bar force runs inside one wave with the subcell
// rows in LDS: (1) subcell H (LPC lanes per subcell, butterfly) and J = DetApIm1(H) + 1;
// (2) one lane per subcell: J-bar of its cell (the cell's J_k V_k summed in order over
// V_cell, :120-175), H-bar = H cbrt(J-bar / J) + (cbrt(.) - 1) I unless the ratio is
// within 3 eps of 1, the stress of H-bar (:177-213) and the subcell's term of the dual
// pressure; (3) one lane per subcell: the dual pressure (the cell's terms in order) and
// the un-barred P = P-bar s + (p - s / (3 J) P-bar : F) J F^-T (UnBarStress, :215-309);
// (4) the entry forces -V P grad phi added into the tile's LDS rows, as in k_rk_tile.
// The entries are read twice (H needs every subcell of a cell before any stress), the
// second time right after the first by the same wave.
constexpr int kFbarSubRow = 29;  // LDS doubles per subcell: H 9, J 1, P-bar 9, dual term 1, P 9

struct FbarTileArgs {
    RkTileArgs T;
    const int32_t *t_fstart;     // [ntiles + 1] first nodal cell of each tile
    const int64_t *fcell_sub;    // [n_fcells + 1] subcell range of each nodal cell
    const double *fcell_volume;  // [n_fcells]
    const int32_t *sub_fcell;    // [ncells] nodal cell of each subcell
    int maxsub;                  // most subcells in one tile
};

template <int KIND, int LPC>
__global__ __launch_bounds__(64) void k_fbar_tile(FbarTileArgs B) {
#pragma clang fp contract(off)
    constexpr double kEps = 2.2204460492503131e-16;
    const RkTileArgs &A = B.T;
    extern __shared__ __align__(16) double rk_lds[];
    double *U = rk_lds;
    double *F = U + 3 * (size_t)A.maxtn;
    double *SH = F + 3 * (size_t)A.maxtn;  // [maxsub * 9]
    double *SJ = SH + 9 * (size_t)B.maxsub;
    double *SPb = SJ + B.maxsub;           // [maxsub * 9]
    double *SD = SPb + 9 * (size_t)B.maxsub;
    double *SP = SD + B.maxsub;            // [maxsub * 9]
    const int64_t t = xcd_block(A.nblocks);
    if (t >= A.ntiles) return;
    const int lane = threadIdx.x;
    const int tb = A.t_nstart[t], tn = A.t_nstart[t + 1] - tb;
    constexpr int kStage = 8;
    for (int l0 = lane; l0 < tn; l0 += 64 * kStage) {
        int64_t gi[kStage];
#pragma unroll
        for (int r = 0; r < kStage; ++r) gi[r] = A.t_node[tb + min(l0 + 64 * r, tn - 1)];
        double ur[kStage][3];
#pragma unroll
        for (int r = 0; r < kStage; ++r) ld3(A.u, gi[r], ur[r]);
#pragma unroll
        for (int r = 0; r < kStage; ++r) {
            const int ln = l0 + 64 * r;
            if (ln < tn) {
#pragma unroll
                for (int d = 0; d < 3; ++d) {
                    U[3 * ln + d] = ur[r][d];
                    F[3 * ln + d] = 0.0;
                }
            }
        }
    }
    __syncthreads();
    constexpr int G = 64 / LPC;
    const int grp = lane / LPC, j = lane % LPC;
    const int64_t c0 = A.t_cstart[t], c1 = A.t_cstart[t + 1];
    // (1) subcell H and J
    for (int64_t cb = c0; cb < c1; cb += G) {
        const int64_t c_raw = cb + grp;
        const bool live = c_raw < c1;
        const int64_t c = live ? c_raw : c1 - 1;
        const int64_t t0 = A.start[c], t1 = A.start[c + 1];
        double H[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        for (int64_t e = t0 + j; e < t1; e += LPC) {
            double uj[3], g[3];
            ld3(U, A.lidx[e], uj);
            ld3(A.dphi, e, g);
#pragma unroll
            for (int a = 0; a < 3; ++a)
#pragma unroll
                for (int b = 0; b < 3; ++b) H[3 * a + b] = __fma_rn(uj[a], g[b], H[3 * a + b]);
        }
#pragma unroll
        for (int i = 0; i < 9; ++i) H[i] = group_sum<LPC>(H[i]);
        if (live && j == 0) {
            const int sl = (int)(c - c0);
#pragma unroll
            for (int i = 0; i < 9; ++i) {
                SH[9 * sl + i] = H[i];
                store_once(A.H + i * A.ncells + c, H[i]);
            }
            SJ[sl] = det_ap_im1(H) + 1.0;
        }
    }
    __syncthreads();
    auto jbar_of = [&](int64_t r) {
        double jb = 0.0;
        for (int64_t k = B.fcell_sub[r]; k < B.fcell_sub[r + 1]; ++k) jb = jb + SJ[k - c0] * A.volume[k];
        return jb / B.fcell_volume[r];
    };
    // (2) H-bar, its stress, the dual-pressure term
    for (int64_t s = c0 + lane; s < c1; s += 64) {
        const int sl = (int)(s - c0);
        const double jbar = jbar_of(B.sub_fcell[s]);
        double H[9], Hb[9], Pb[9];
#pragma unroll
        for (int i = 0; i < 9; ++i) H[i] = SH[9 * sl + i];
        const double ratio = jbar / SJ[sl];
        if (fabs(ratio - 1.0) < 3.0 * kEps) {
#pragma unroll
            for (int i = 0; i < 9; ++i) Hb[i] = H[i];
        } else {
            const double js = cbrt(ratio);
#pragma unroll
            for (int i = 0; i < 9; ++i) Hb[i] = H[i] * js;
            Hb[0] += js - 1.0;
            Hb[4] += js - 1.0;
            Hb[8] += js - 1.0;
        }
        double snp1 = 0.0, aux = 0.0;
        const double sn = load_state<KIND>(A.state_n, s);
        stress<KIND>(Hb, A.mat, sn, snp1, aux, Pb);
        store_state<KIND>(A.state_np1, s, snp1, aux);
#pragma unroll
        for (int i = 0; i < 9; ++i) SPb[9 * sl + i] = Pb[i];
        const double js = cbrt(jbar / SJ[sl]);
        const double djb = js / (3.0 * jbar);
        double ph = 0.0;
#pragma unroll
        for (int i = 0; i < 9; ++i) ph += Pb[i] * H[i];
        const double tr = Pb[0] + Pb[4] + Pb[8];
        SD[sl] = djb * A.volume[s] * (ph + tr);
    }
    __syncthreads();
    // (3) the dual pressure and the un-barred stress
    for (int64_t s = c0 + lane; s < c1; s += 64) {
        const int sl = (int)(s - c0);
        const int64_t r = B.sub_fcell[s];
        const double jbar = jbar_of(r);
        double dual = 0.0;
        for (int64_t k = B.fcell_sub[r]; k < B.fcell_sub[r + 1]; ++k) dual += SD[k - c0];
        dual = dual / B.fcell_volume[r];
        const double Js = SJ[sl];
        const double js = cbrt(jbar / Js);
        double Fm[9], Fi[9], Pb[9], pf = 0.0;
#pragma unroll
        for (int i = 0; i < 9; ++i) {
            Fm[i] = SH[9 * sl + i];
            Pb[i] = SPb[9 * sl + i];
        }
        Fm[0] += 1.0;
        Fm[4] += 1.0;
        Fm[8] += 1.0;
#pragma unroll
        for (int i = 0; i < 9; ++i) pf += Pb[i] * Fm[i];
        const double djj = -js / (3.0 * Js);
        const double scalar = (dual + djj * pf) * Js;
        inverse3(Fm, Fi);
#pragma unroll
        for (int a = 0; a < 3; ++a)
#pragma unroll
            for (int b = 0; b < 3; ++b) {
                const double p = Pb[3 * a + b] * js + scalar * Fi[3 * b + a];
                SP[9 * sl + 3 * a + b] = p;
                store_once(A.P + (3 * a + b) * A.ncells + s, p);
            }
    }
    __syncthreads();
    // (4) entry forces
    constexpr int BQ = 4;  // entries per lane whose loads are issued together
    for (int64_t cb = c0; cb < c1; cb += G) {
        const int64_t c = cb + grp;
        if (c >= c1) continue;
        const int sl = (int)(c - c0);
        double P[9];
#pragma unroll
        for (int i = 0; i < 9; ++i) P[i] = SP[9 * sl + i];
        const double mv = -A.volume[c];
        const int64_t t0 = A.start[c], t1 = A.start[c + 1];
        for (int64_t eb = t0 + j; eb < t1; eb += BQ * LPC) {
            int li[BQ];
            double g[BQ][3];
#pragma unroll
            for (int r = 0; r < BQ; ++r) {
                const int64_t e = min(eb + r * LPC, t1 - 1);
                li[r] = A.lidx[e];
                ld3(A.dphi, e, g[r]);
            }
#pragma unroll
            for (int r = 0; r < BQ; ++r) {
                if (eb + r * LPC < t1) {
#pragma unroll
                    for (int a = 0; a < 3; ++a) {
                        const double fa = mv * dot3(P[3 * a], g[r][0], P[3 * a + 1], g[r][1], P[3 * a + 2], g[r][2]);
                        __hip_atomic_fetch_add(F + 3 * li[r] + a, fa, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                    }
                }
            }
        }
    }
    __syncthreads();
    double *dst = A.slots + 3 * (int64_t)tb;
    for (int k = lane; k < 3 * tn; k += 64) dst[k] = F[k];
}

struct RkSlotNodeArgs {
    RkNodeArgs out;
    const int64_t *nslot_start;  // [n + 1] slot range of each node
    const int32_t *nslot;        // slots (tile-node positions) of each node, ascending tile
    const double *slots;
};

// f_i = sum of node i's tile rows in ascending tile order (lanes stride over them,
// then the butterfly), then the node update.
template <int MODE, int LANES>
__global__ __launch_bounds__(256) void k_rk_slot_nodes(RkSlotNodeArgs A) {
#pragma clang fp contract(off)
    const int l = threadIdx.x & (LANES - 1);
    const int64_t i_raw = xcd_block(A.out.nblocks) * (256 / LANES) + (threadIdx.x / LANES);
    const bool live = i_raw < A.out.nn;
    const int64_t i = live ? i_raw : A.out.nn - 1;
    double f[3] = {0.0, 0.0, 0.0};
    const int64_t s0 = A.nslot_start[i], s1 = A.nslot_start[i + 1];
    for (int64_t s = s0 + l; s < s1; s += LANES) {
        double r[3];
        ld3(A.slots, A.nslot[s], r);
        f[0] += r[0];
        f[1] += r[1];
        f[2] += r[2];
    }
#pragma unroll
    for (int d = 0; d < 3; ++d) f[d] = group_sum<LANES>(f[d]);
    if (!live || l != 0) return;
    rk_finish<MODE>(A.out, i, f);
}

// ---- two-pass method: gather before, scatter after the cell force -----------
struct RkGatherArgs {
    int64_t nn, nblocks;
    const int32_t *num, *nbr;
    const double *phi;
    const double *src;  // coefficients [nn * 3]
    double *dst;        // physical values [nn * 3]
};

// dst_i = sum_k phi_ik src_(nbr_ik): LANES lanes per node, lane l takes entries
// l, l + LANES, ... of the node's row (contiguous across the lanes), the partial
// sums are combined by the xor butterfly.  A node without neighbors copies src_i.
// (8 lanes, five entries each for a full row of 40: 3 % faster than 16 lanes.)
template <int LANES>
__global__ __launch_bounds__(256) void k_rk_gather(RkGatherArgs A) {
    constexpr int R = (kMaxNbr + LANES - 1) / LANES;
    const int l = threadIdx.x & (LANES - 1);
    const int64_t i_raw = xcd_block(A.nblocks) * (256 / LANES) + (threadIdx.x / LANES);
    const bool live = i_raw < A.nn;
    const int64_t i = live ? i_raw : A.nn - 1;
    const int n = A.num[i];
    const int64_t row = i * kMaxNbr;
    int32_t j[R];
    double w[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int k = l + r * LANES;
        const int kk = max(min(k, n - 1), 0);  // inside the row; entries past n are weighted by zero
        const int32_t jl = A.nbr[row + kk];
        const double wl = A.phi[row + kk];
        j[r] = n > 0 ? jl : (int32_t)i;
        w[r] = k < n ? wl : 0.0;
    }
    double s[3] = {0.0, 0.0, 0.0};
    double x[R][3];
#pragma unroll
    for (int r = 0; r < R; ++r) ld3(A.src, j[r], x[r]);
#pragma unroll
    for (int r = 0; r < R; ++r)
#pragma unroll
        for (int d = 0; d < 3; ++d) s[d] += x[r][d] * w[r];
#pragma unroll
    for (int d = 0; d < 3; ++d) s[d] = group_sum<LANES>(s[d]);
    if (!live || l != 0) return;
    if (n == 0) ld3(A.src, i, s);
    A.dst[3 * i] = s[0];
    A.dst[3 * i + 1] = s[1];
    A.dst[3 * i + 2] = s[2];
}

struct RkScatterArgs {
    RkNodeArgs out;          // force and node update of the coefficients
    const int32_t *num;
    const int64_t *rstart;   // [nn + 1] reverse neighbor lists
    const int32_t *rsrc;     // points i listing the node, ascending
    const double *rphi;      // phi_iJ of each reverse entry
    const double *src;       // local (physical) force [nn * 3]
};

// f_J = sum over points i that list J of phi_iJ src_i (the transpose of the
// gather), LANES lanes per node over its contiguous reverse entries, NB per lane at a
// time (16 x 3: a reverse list of about 40 entries in one round with little padding;
// 6 % faster than 16 x 4, and than 8 x 5), then the node update.  A node without
// neighbors takes src_J.
template <int MODE, int LANES, int NB>
__global__ __launch_bounds__(256) void k_rk_scatter(RkScatterArgs A) {
    const int l = threadIdx.x & (LANES - 1);
    const int64_t i_raw = xcd_block(A.out.nblocks) * (256 / LANES) + (threadIdx.x / LANES);
    const bool live = i_raw < A.out.nn;
    const int64_t i = live ? i_raw : A.out.nn - 1;
    double f[3] = {0.0, 0.0, 0.0};
    const int64_t t0 = A.rstart[i], t1 = A.rstart[i + 1];
    for (int64_t tb = t0 + l; tb < t1; tb += NB * LANES) {
        int32_t src[NB];
        double w[NB], x[NB][3];
#pragma unroll
        for (int r = 0; r < NB; ++r) {
            const int64_t t = min(tb + r * LANES, t1 - 1);
            src[r] = A.rsrc[t];
            w[r] = tb + r * LANES < t1 ? A.rphi[t] : 0.0;
        }
#pragma unroll
        for (int r = 0; r < NB; ++r) ld3(A.src, src[r], x[r]);
#pragma unroll
        for (int r = 0; r < NB; ++r)
#pragma unroll
            for (int d = 0; d < 3; ++d) f[d] += x[r][d] * w[r];
    }
#pragma unroll
    for (int d = 0; d < 3; ++d) f[d] = group_sum<LANES>(f[d]);
    if (!live || l != 0) return;
    if (A.num[i] == 0) ld3(A.src, i, f);
    rk_finish<MODE>(A.out, i, f);
}

// mass_from_elements: rho V / NPE from every cell touching the node, ascending cell
template <int NPE>
__global__ __launch_bounds__(256) void k_rk_mfe(int64_t nn, const int32_t *node_start, const int32_t *node_inc,
                                                const double *vol, double rho, double *mfe) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= nn) return;
    double m = 0.0;
    for (int32_t t = node_start[i]; t < node_start[i + 1]; ++t) m += (vol[node_inc[t] / NPE] * rho) / NPE;
    mfe[i] = m;
}

// m_J = sum over points i that list J of phi_iJ * mfe_i, ascending i (ScatterValues)
__global__ __launch_bounds__(256) void k_rk_mass(int64_t nn, const int32_t *num, const int64_t *rstart,
                                                 const int32_t *rsrc, const double *rphi, const double *mfe,
                                                 double *mass) {
    const int64_t j = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (j >= nn) return;
    if (num[j] == 0) {
        mass[j] = mfe[j];
        return;
    }
    double m = 0.0;
    for (int64_t t = rstart[j]; t < rstart[j + 1]; ++t) m += mfe[rsrc[t]] * rphi[t];
    mass[j] = m;
}

inline dim3 grid1(int64_t n, int block) { return dim3((unsigned)((n + block - 1) / block)); }

// blocks of 256 / lanes groups for n groups, padded to a multiple of the XCD count
inline int64_t padded_groups(int64_t n, int lanes = kCellLanes) {
    const int64_t per = 256 / lanes, nb = (n + per - 1) / per;
    return ((nb + kNumXcd - 1) / kNumXcd) * kNumXcd;
}

// counts -> exclusive prefix (host) -> device start array of n+1 entries
template <typename T>
void scan_counts(Ctx &c, const DevBuf<int32_t> &counts, int64_t n, DevBuf<T> &start, int64_t *total) {
    std::vector<int32_t> h(n);
    counts.download(h.data(), n, c.stream);
    APHIP_HIP(hipStreamSynchronize(c.stream));
    std::vector<T> s(n + 1, 0);
    for (int64_t i = 0; i < n; ++i) s[i + 1] = s[i] + (T)h[i];
    start.upload(s.data(), s.size(), c.stream);
    APHIP_HIP(hipStreamSynchronize(c.stream));
    if (total) *total = (int64_t)s[n];
}

}  // namespace

void fe_smoothing_setup(Ctx &c);
void rk_cells_and_order(Ctx &c, bool two_pass, bool elem_cells);
void rk_cells_order(Ctx &c, bool two_pass);

// Nodal cells of a thex mesh (MeshLabelerProcessor.h:227-330, :690-827): marks each hex's
// node of lowest caller id active, checks that every hex then holds exactly one active
// node, renumbers the internal nodes with the active ones first (c.n_dof of them) and
// records, per cell (active node, ascending caller id) its hexes in ascending caller
// element id; with subdomains 0 every hex is a subcell, with 1 the whole cell.  Returns
// the subcells as element lists (sub_start / sub_elems, internal element ids).
static void nodal_label(Ctx &c, int subdomains, std::vector<int64_t> &sub_start, std::vector<int32_t> &sub_elems) {
    CellData &D = c.cells;
    const int64_t nn = c.nn, ne = c.ne;
    APHIP_REQUIRE(c.etype == APHIP_HEX8,
                  "Nodal integration will only work correctly when being built from hexahedral elements. Use 'thex' "
                  "to divide a tetrahedral mesh into hexahedral elements.");
    APHIP_REQUIRE(subdomains == 0 || subdomains == 1, "nodal smoothing cells take subdomains 0 or 1");
    std::vector<uint8_t> act(nn, 0);
    for (int64_t e = 0; e < ne; ++e) {
        int best = 0;
        for (int k = 1; k < 8; ++k)
            if (c.node_iperm[c.conn_h[8 * e + k]] < c.node_iperm[c.conn_h[8 * e + best]]) best = k;
        act[c.conn_h[8 * e + best]] = 1;
    }
    for (int64_t e = 0; e < ne; ++e) {
        int n = 0;
        for (int k = 0; k < 8; ++k) n += act[c.conn_h[8 * e + k]];
        APHIP_REQUIRE(n == 1, "Nodal integration requires exactly one active node per element. Found " +
                                  std::to_string(n) + " active nodes in element " +
                                  std::to_string(c.elem_iperm[e]) + ".");
    }
    // active nodes first, each group in its current (Morton) order
    std::vector<int32_t> old_of_new;
    old_of_new.reserve(nn);
    for (int64_t i = 0; i < nn; ++i)
        if (act[i]) old_of_new.push_back((int32_t)i);
    const int64_t n_act = (int64_t)old_of_new.size();
    for (int64_t i = 0; i < nn; ++i)
        if (!act[i]) old_of_new.push_back((int32_t)i);
    renumber_nodes(c, old_of_new);
    c.tiles.reset();  // the tile plan's node groups are void after the renumbering
    c.n_dof = n_act;
    // cells: active nodes by caller id; hexes by caller element id
    std::vector<int32_t> act_nodes(n_act);
    std::iota(act_nodes.begin(), act_nodes.end(), 0);
    std::sort(act_nodes.begin(), act_nodes.end(),
              [&](int32_t x, int32_t y) { return c.node_iperm[x] < c.node_iperm[y]; });
    std::vector<int32_t> cell_of_node(n_act);
    for (int64_t r = 0; r < n_act; ++r) cell_of_node[act_nodes[r]] = (int32_t)r;
    std::vector<int32_t> hexes(ne);
    std::iota(hexes.begin(), hexes.end(), 0);
    std::vector<int32_t> hex_cell(ne);
    for (int64_t e = 0; e < ne; ++e)
        for (int k = 0; k < 8; ++k)
            if (c.conn_h[8 * e + k] < n_act) hex_cell[e] = cell_of_node[c.conn_h[8 * e + k]];
    std::sort(hexes.begin(), hexes.end(), [&](int32_t x, int32_t y) {
        return hex_cell[x] != hex_cell[y] ? hex_cell[x] < hex_cell[y] : c.elem_iperm[x] < c.elem_iperm[y];
    });
    D.n_fcells = n_act;
    std::vector<int64_t> cell_start(n_act + 1, 0);
    for (int64_t e = 0; e < ne; ++e) cell_start[hex_cell[e] + 1]++;
    for (int64_t r = 0; r < n_act; ++r) cell_start[r + 1] += cell_start[r];
    sub_elems = hexes;
    if (subdomains == 0) {
        sub_start.resize(ne + 1);
        std::iota(sub_start.begin(), sub_start.end(), 0);
        D.fcell_sub_h = cell_start;
    } else {
        sub_start = cell_start;
        D.fcell_sub_h.resize(n_act + 1);
        std::iota(D.fcell_sub_h.begin(), D.fcell_sub_h.end(), 0);
    }
}

// Subcell node lists (first-seen union of the neighbors of the nodes of the subcell's
// hexes, in hex order) and derivatives sum_e sum_k B_k V_e phi / V_sub
// (SmoothedCellDataProcessor.h:407-675), built on the host from the device's neighbor
// lists and functions; cell volumes are the sums of their hexes' volumes.
static void nodal_subcells(Ctx &c, const std::vector<int64_t> &sub_start, const std::vector<int32_t> &sub_elems) {
    CellData &D = c.cells;
    const int64_t nn = c.nn, ne = c.ne, ns = (int64_t)sub_start.size() - 1;
    std::vector<int32_t> hn(nn), hb(nn * kMaxNbr);
    std::vector<double> hp(nn * kMaxNbr), hX(3 * nn);
    D.num_nbr.download(hn.data(), nn, c.stream);
    D.nbr.download(hb.data(), hb.size(), c.stream);
    D.phi.download(hp.data(), hp.size(), c.stream);
    c.X.download(hX.data(), hX.size(), c.stream);
    APHIP_HIP(hipStreamSynchronize(c.stream));
    std::vector<int64_t> start(ns + 1, 0);
    std::vector<int32_t> nodes, inc_cell;
    std::vector<double> dphi, vsub(ns, 0.0), evol(ne, 0.0);
    std::vector<int64_t> stamp(nn, -1);
    std::vector<int32_t> pos(nn, 0);
    D.sub_of_elem.assign(ne, -1);
    D.sub_first_elem.assign(ns, -1);
    for (int64_t sc = 0; sc < ns; ++sc) {
        const int64_t t0 = (int64_t)nodes.size();
        start[sc] = t0;
        for (int64_t q = sub_start[sc]; q < sub_start[sc + 1]; ++q) {
            const int32_t e = sub_elems[q];
            D.sub_of_elem[e] = (int32_t)sc;
            if (q == sub_start[sc]) D.sub_first_elem[sc] = e;
            for (int k = 0; k < 8; ++k) {
                const int32_t a = c.conn_h[8 * (int64_t)e + k];
                for (int l = 0; l < hn[a]; ++l) {
                    const int32_t j = hb[(int64_t)a * kMaxNbr + l];
                    if (stamp[j] != sc) {
                        stamp[j] = sc;
                        pos[j] = (int32_t)((int64_t)nodes.size() - t0);
                        nodes.push_back(j);
                        inc_cell.push_back((int32_t)sc);
                    }
                }
            }
        }
        dphi.resize(3 * nodes.size(), 0.0);
        double vs = 0.0;
        for (int64_t q = sub_start[sc]; q < sub_start[sc + 1]; ++q) {
            const int32_t e = sub_elems[q];
            double Xe[8][3], B[8][3], V;
            for (int k = 0; k < 8; ++k)
                for (int d = 0; d < 3; ++d) Xe[k][d] = hX[3 * (int64_t)c.conn_h[8 * (int64_t)e + k] + d];
            smoothed_hex8(Xe, B, V);
            evol[e] = V;
            vs += V;
            for (int k = 0; k < 8; ++k) {
                const int32_t a = c.conn_h[8 * (int64_t)e + k];
                for (int l = 0; l < hn[a]; ++l) {
                    const int32_t j = hb[(int64_t)a * kMaxNbr + l];
                    const double fv = hp[(int64_t)a * kMaxNbr + l];
                    double *dp = &dphi[3 * (t0 + pos[j])];
                    for (int m = 0; m < 3; ++m) dp[m] += B[k][m] * V * fv;
                }
            }
        }
        for (size_t t = 3 * t0; t < dphi.size(); ++t) dphi[t] /= vs;
        vsub[sc] = vs;
    }
    start[ns] = (int64_t)nodes.size();
    D.ncells = ns;
    D.nnz = start[ns];
    APHIP_REQUIRE(D.nnz < (int64_t)INT32_MAX, "too many cell entries for int32 incidences");
    D.start.upload(start.data(), start.size(), c.stream);
    D.nodes.upload(nodes.data(), std::max<size_t>(nodes.size(), 1), c.stream);
    D.inc_cell.upload(inc_cell.data(), std::max<size_t>(inc_cell.size(), 1), c.stream);
    D.dphi.upload(dphi.data(), std::max<size_t>(dphi.size(), 1), c.stream);
    D.volume.upload(vsub.data(), std::max<int64_t>(ns, 1), c.stream);
    D.evol.upload(evol.data(), std::max<int64_t>(ne, 1), c.stream);
    // nodal cells: subcell ranges and volumes (sum of the cell's hex volumes, in order)
    std::vector<double> vcell(D.n_fcells, 0.0);
    for (int64_t r = 0; r < D.n_fcells; ++r)
        for (int64_t sc = D.fcell_sub_h[r]; sc < D.fcell_sub_h[r + 1]; ++sc)
            for (int64_t q = sub_start[sc]; q < sub_start[sc + 1]; ++q) vcell[r] += evol[sub_elems[q]];
    D.fcell_sub.upload(D.fcell_sub_h.data(), D.fcell_sub_h.size(), c.stream);
    D.fcell_volume.upload(vcell.data(), std::max<int64_t>(D.n_fcells, 1), c.stream);
    if (D.fbar) {
        std::vector<int32_t> sc_of(std::max<int64_t>(ns, 1), 0);
        for (int64_t r = 0; r < D.n_fcells; ++r)
            for (int64_t sc = D.fcell_sub_h[r]; sc < D.fcell_sub_h[r + 1]; ++sc) sc_of[sc] = (int32_t)r;
        D.sub_fcell.upload(sc_of.data(), sc_of.size(), c.stream);
    }
    APHIP_HIP(hipStreamSynchronize(c.stream));
}

void rk_setup(Ctx &c, double scale, bool two_pass, bool fe_cells, int nodal_subdomains, bool fbar) {
    CellData &D = c.cells;
    D.reset();
    c.n_dof = c.nn;
    const int64_t nn = c.nn;
    if (fe_cells) {
        fe_smoothing_setup(c);
        return;
    }
    if (scale == 1.0) scale += 1.0e-6;  // src/NeighborSearchProcessor.cpp:161-163
    const bool nodal = nodal_subdomains >= 0;
    std::vector<int64_t> sub_start;
    std::vector<int32_t> sub_elems;
    if (nodal) {
        nodal_label(c, nodal_subdomains, sub_start, sub_elems);
        D.nodal = true;
        // F-bar only where a cell holds more than one subcell (ComputeInternalForceSmoothedCell.h:458-459)
        D.fbar = fbar && (int64_t)sub_start.size() - 1 != D.n_fcells;
    }
    const int64_t n_act = c.n_dof;  // search candidates: the nodes with degrees of freedom

    // kernel radius from the max edge length
    D.radius.alloc(nn);
    const bool hex = c.etype == APHIP_HEX8;  // Hex8 background: ElementReproducingKernelHex8
    hipLaunchKernelGGL(hex ? k_max_edge<8> : k_max_edge<4>, grid1(nn, 256), dim3(256), 0, c.stream, nn, c.node_start.p,
                       c.node_inc.p, c.conn.p, c.X.p, scale, n_act, D.radius.p);
    APHIP_HIP(hipGetLastError());

    // uniform cell list of the candidates with cell edge = their largest radius
    std::vector<double> hX(3 * nn), hr(nn);
    c.X.download(hX.data(), hX.size(), c.stream);
    D.radius.download(hr.data(), hr.size(), c.stream);
    APHIP_HIP(hipStreamSynchronize(c.stream));
    Grid G;
    double hi[3];
    for (int d = 0; d < 3; ++d) {
        G.lo[d] = INFINITY;
        hi[d] = -INFINITY;
    }
    double rmax = 0.0;
    for (int64_t i = 0; i < nn; ++i) {
        for (int d = 0; d < 3; ++d) {
            G.lo[d] = std::min(G.lo[d], hX[3 * i + d]);
            hi[d] = std::max(hi[d], hX[3 * i + d]);
        }
        if (i < n_act) rmax = std::max(rmax, hr[i]);
    }
    double ext = 0.0;
    for (int d = 0; d < 3; ++d) ext = std::max(ext, hi[d] - G.lo[d]);
    G.cell = rmax > 0.0 ? rmax : (ext > 0.0 ? ext : 1.0);
    // keep the grid modest; a coarser cell only adds candidates, never drops one
    while (true) {
        int64_t tot = 1;
        for (int d = 0; d < 3; ++d) {
            G.nb[d] = (int)std::floor((hi[d] - G.lo[d]) / G.cell) + 1;
            tot *= G.nb[d];
        }
        if (tot <= (int64_t)1 << 24) break;
        G.cell *= 1.26;
    }
    const int64_t nbins = (int64_t)G.nb[0] * G.nb[1] * G.nb[2];
    DevBuf<int32_t> bin_of, bcount, bstart, bfill, items;
    bin_of.alloc(std::max<int64_t>(n_act, 1));
    bcount.alloc(nbins);
    bcount.zero(c.stream);
    hipLaunchKernelGGL(k_bin_count, grid1(n_act, 256), dim3(256), 0, c.stream, n_act, c.X.p, G, bin_of.p, bcount.p);
    APHIP_HIP(hipGetLastError());
    scan_counts(c, bcount, nbins, bstart, nullptr);
    bfill.alloc(nbins);
    bfill.zero(c.stream);
    items.alloc(std::max<int64_t>(n_act, 1));
    hipLaunchKernelGGL(k_bin_fill, grid1(n_act, 256), dim3(256), 0, c.stream, n_act, bin_of.p, bstart.p, bfill.p,
                       items.p);
    APHIP_HIP(hipGetLastError());

    // ball search of every node among the candidates, ranked by caller ids for ties
    DevBuf<int32_t> ncand;
    D.num_nbr.alloc(nn);
    D.nbr.alloc(nn * kMaxNbr);
    ncand.alloc(nn);
    SearchArgs S{nn, c.X.p, D.radius.p, c.d_node_iperm.p, G, bstart.p, items.p, D.num_nbr.p, D.nbr.p, ncand.p};
    hipLaunchKernelGGL(k_search, grid1(nn, kSearchBlock), dim3(kSearchBlock), 0, c.stream, S);
    APHIP_HIP(hipGetLastError());
    {
        std::vector<int32_t> hn(nn), hc(nn);
        D.num_nbr.download(hn.data(), nn, c.stream);
        ncand.download(hc.data(), nn, c.stream);
        APHIP_HIP(hipStreamSynchronize(c.stream));
        int64_t tr = 0;
        for (int64_t i = 0; i < nn; ++i) tr += hc[i] - hn[i];
        D.n_truncated = tr;
    }

    // RK function values
    D.phi.alloc(nn * kMaxNbr);
    hipLaunchKernelGGL(k_rk_phi, grid1(nn, 128), dim3(128), 0, c.stream, nn, c.X.p, D.radius.p, D.num_nbr.p, D.nbr.p,
                       D.phi.p);
    APHIP_HIP(hipGetLastError());

    if (nodal) {
        nodal_subcells(c, sub_start, sub_elems);
        rk_cells_order(c, false);
    } else {
        rk_cells_and_order(c, two_pass, two_pass);
    }
}

// Finite-element strain smoothing with element cells (ElementSmoothedTetrahedron4,
// include/ElementSmoothedTetrahedron4.h): the element's own shape functions, i.e. every
// node its own only "neighbor" with value 1, and the cells of the element's nodes with
// derivative (B_k V) / V (SmoothedCellDataProcessor.h:623-668 with unit function values).
// The force runs on the one-pass kernels; the mass scatter is the identity.
void fe_smoothing_setup(Ctx &c) {
    CellData &D = c.cells;
    const int64_t nn = c.nn;
    D.radius.alloc(nn);
    D.radius.zero(c.stream);
    D.num_nbr.alloc(nn);
    D.nbr.alloc(nn * kMaxNbr);
    D.phi.alloc(nn * kMaxNbr);
    hipLaunchKernelGGL(k_identity_functions, grid1(nn, 256), dim3(256), 0, c.stream, nn, D.num_nbr.p, D.nbr.p,
                       D.phi.p);
    APHIP_HIP(hipGetLastError());
    D.n_truncated = 0;
    rk_cells_and_order(c, false, true);
}

// Cells (element-node cells when elem_cells), reverse neighbor lists and the node
// order of the cell entries; two_pass selects the force method.
void rk_cells_and_order(Ctx &c, bool two_pass, bool elem_cells) {
    CellData &D = c.cells;
    const int64_t nn = c.nn, ne = c.ne;
    const bool hex = c.etype == APHIP_HEX8;
    const int npe = hex ? 8 : 4;
    // smoothed cells: counts, offsets, lists and derivatives
    D.two_pass = two_pass;
    if (elem_cells) {
        // the element's own nodes (SmoothedCellDataProcessor.h:474-486, :637-648)
        D.nnz = npe * ne;
        APHIP_REQUIRE(D.nnz < (int64_t)INT32_MAX, "too many cell entries for int32 incidences");
        std::vector<int64_t> hs(ne + 1);
        for (int64_t e = 0; e <= ne; ++e) hs[e] = npe * e;
        D.start.upload(hs.data(), hs.size(), c.stream);
        D.nodes.alloc(std::max<int64_t>(D.nnz, 1));
        D.dphi.alloc(3 * std::max<int64_t>(D.nnz, 1));
        D.volume.alloc(std::max<int64_t>(ne, 1));
        D.inc_cell.alloc(std::max<int64_t>(D.nnz, 1));
        if (ne) {
            hipLaunchKernelGGL(hex ? k_cell_two_pass<8> : k_cell_two_pass<4>, grid1(ne, 256), dim3(256), 0, c.stream, ne, c.conn.p, c.X.p,
                               D.nodes.p, D.dphi.p, D.volume.p, D.inc_cell.p);
            APHIP_HIP(hipGetLastError());
        }
        if (two_pass) {
            D.u_phys.alloc(3 * nn);
            D.u_phys.zero(c.stream);
            D.f_local.alloc(3 * nn);
            D.f_local.zero(c.stream);
        }
    } else {
        DevBuf<int32_t> clen;
        clen.alloc(std::max<int64_t>(ne, 1));
        if (ne) {
            const int blk = hex ? cell_block<8>() : cell_block<4>();
            hipLaunchKernelGGL(hex ? k_cell_count<8> : k_cell_count<4>, grid1(ne, blk), dim3(blk), 0, c.stream, ne, c.conn.p,
                               D.num_nbr.p, D.nbr.p, clen.p);
            APHIP_HIP(hipGetLastError());
        }
        scan_counts(c, clen, ne, D.start, &D.nnz);
        APHIP_REQUIRE(D.nnz < (int64_t)INT32_MAX, "too many cell entries for int32 incidences");
        D.nodes.alloc(std::max<int64_t>(D.nnz, 1));
        D.dphi.alloc(3 * std::max<int64_t>(D.nnz, 1));
        D.volume.alloc(std::max<int64_t>(ne, 1));
        D.inc_cell.alloc(std::max<int64_t>(D.nnz, 1));
        if (ne) {
            CellFillArgs F{ne, c.conn.p, c.X.p, D.num_nbr.p, D.nbr.p, D.phi.p, D.start.p, D.nodes.p, D.dphi.p,
                           D.volume.p, D.inc_cell.p};
            const int blk = hex ? cell_block<8>() : cell_block<4>();
            hipLaunchKernelGGL(hex ? k_cell_fill<8> : k_cell_fill<4>, grid1(ne, blk), dim3(blk), 0, c.stream, F);
            APHIP_HIP(hipGetLastError());
        }
    }

    D.ncells = ne;
    rk_cells_order(c, two_pass);
}

// LDS of one tile-kernel workgroup (u and force rows of the tile nodes; F-bar also the
// subcell rows) is kept to 64 KiB so that several tiles share a CU.
constexpr int64_t kTileLdsMax = 64 * 1024;

static int64_t tile_lds(int maxtn, int maxsub, bool fbar) {
    return 48 * (int64_t)std::max(maxtn, 1) + (fbar ? 8 * (int64_t)kFbarSubRow * std::max(maxsub, 1) : 0);
}

// Cell tiles of the one-pass force: runs of tile_cells consecutive cells, the sorted
// union of their node lists, every entry's index in it, and every node's slots (its
// positions in the tile-major node sets, ascending tile).  The tile size starts at 64
// cells (256 for cells of at most four entries), is halved while the mesh gives fewer
// than 2048 tiles (F-bar, whose tiles do several passes over their subcells: 8192; on a
// thex mesh of 192,000 subcells 16 instead of 64 subcells per tile took the step from
// 0.131 to 0.121 ms), and then until every tile's node set fits the LDS budget.  The lanes
// per cell follow the longest cell list (4 up to 4 entries, 8 up to 96, else 16).
// APHIP_RK_TILE_CELLS and APHIP_RK_TILE_LANES override the two choices.
static void rk_tile_plan(Ctx &c) {
    CellData &D = c.cells;
    const int64_t nc = D.ncells, nnz = D.nnz, nn = c.nn;
    std::vector<int64_t> hs(nc + 1, 0);
    std::vector<int32_t> hn(std::max<int64_t>(nnz, 1));
    D.start.download(hs.data(), nc + 1, c.stream);
    if (nnz) D.nodes.download(hn.data(), nnz, c.stream);
    APHIP_HIP(hipStreamSynchronize(c.stream));
    int64_t maxlen = 0;
    for (int64_t k = 0; k < nc; ++k) maxlen = std::max(maxlen, hs[k + 1] - hs[k]);
    D.tile_lpc = maxlen <= 4 ? 4 : (maxlen <= 96 ? 8 : 16);
    if (const char *env = std::getenv("APHIP_RK_TILE_LANES")) {
        const int v = std::atoi(env);
        if (v == 4 || v == 8 || v == 16) D.tile_lpc = v;
    }
    int tc = D.tile_lpc == 4 ? 256 : 64;
    // small meshes: at least 2048 tiles (eight waves per CU; F-bar 8192) while a tile keeps 8 cells
    const int64_t min_tiles = D.fbar ? 8192 : 2048;
    while (tc > 8 && (nc + tc - 1) / tc < min_tiles) tc /= 2;
    if (const char *env = std::getenv("APHIP_RK_TILE_CELLS")) {
        const int v = std::atoi(env);
        if (v >= 1 && v <= 4096) tc = v;
    }
    const int nthreads = (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    std::vector<uint16_t> lidx(std::max<int64_t>(nnz, 1));
    std::vector<std::vector<int32_t>> sets;
    int64_t nt = 0;
    int maxtn = 0, maxsub = 0;
    std::vector<int64_t> cst;  // first cell of each tile
    std::vector<int32_t> fst;  // F-bar: first nodal cell of each tile
    while (true) {
        cst.clear();
        fst.clear();
        if (D.fbar) {
            // whole nodal cells, as many as fit tc subcells (a larger cell gets a tile alone)
            const std::vector<int64_t> &fs = D.fcell_sub_h;
            for (int64_t r = 0; r < D.n_fcells;) {
                cst.push_back(fs[r]);
                fst.push_back((int32_t)r);
                int64_t r1 = r + 1;
                while (r1 < D.n_fcells && fs[r1 + 1] - fs[r] <= tc) ++r1;
                r = r1;
            }
            cst.push_back(nc);
            fst.push_back((int32_t)D.n_fcells);
        } else {
            for (int64_t k = 0; k < nc; k += tc) cst.push_back(k);
            cst.push_back(nc);
        }
        nt = (int64_t)cst.size() - 1;
        maxsub = 0;
        for (int64_t t = 0; t < nt; ++t) maxsub = std::max<int64_t>(maxsub, cst[t + 1] - cst[t]);
        sets.assign(nt, {});
        std::vector<int> tile_max(nthreads, 0);
        auto work = [&](int w) {
            std::vector<int32_t> loc(nn, 0);
            for (int64_t t = w; t < nt; t += nthreads) {
                const int64_t e0 = hs[cst[t]], e1 = hs[cst[t + 1]];
                std::vector<int32_t> &u = sets[t];
                u.assign(hn.begin() + e0, hn.begin() + e1);
                std::sort(u.begin(), u.end());
                u.erase(std::unique(u.begin(), u.end()), u.end());
                tile_max[w] = std::max(tile_max[w], (int)u.size());
                if (u.size() > 65535) continue;  // too large: the caller halves the tile
                for (size_t k = 0; k < u.size(); ++k) loc[u[k]] = (int32_t)k;
                for (int64_t e = e0; e < e1; ++e) lidx[e] = (uint16_t)loc[hn[e]];
            }
        };
        std::vector<std::thread> pool;
        for (int w = 1; w < nthreads; ++w) pool.emplace_back(work, w);
        work(0);
        for (auto &th : pool) th.join();
        maxtn = *std::max_element(tile_max.begin(), tile_max.end());
        if (tile_lds(maxtn, maxsub, D.fbar) <= kTileLdsMax || tc == 1) break;
        tc = std::max(1, tc / 2);
    }
    APHIP_REQUIRE(tile_lds(maxtn, maxsub, D.fbar) <= kTileLdsMax,
                  "a smoothing cell names too many nodes (or a nodal cell holds too many subcells) for the tile kernels");
    std::vector<int32_t> tstart(nt + 1, 0);
    for (int64_t t = 0; t < nt; ++t) {
        APHIP_REQUIRE((int64_t)tstart[t] + (int64_t)sets[t].size() < (int64_t)INT32_MAX, "too many tile nodes");
        tstart[t + 1] = tstart[t] + (int32_t)sets[t].size();
    }
    const int64_t total = tstart[nt];
    std::vector<int32_t> tnode(std::max<int64_t>(total, 1));
    std::vector<int64_t> sstart(nn + 1, 0);
    for (int64_t t = 0; t < nt; ++t) {
        std::copy(sets[t].begin(), sets[t].end(), tnode.begin() + tstart[t]);
        for (int32_t g : sets[t]) sstart[g + 1]++;
    }
    for (int64_t i = 0; i < nn; ++i) sstart[i + 1] += sstart[i];
    std::vector<int32_t> nslot(std::max<int64_t>(total, 1));
    {
        std::vector<int64_t> fill(sstart.begin(), sstart.end() - 1);
        for (int64_t k = 0; k < total; ++k) nslot[fill[tnode[k]]++] = (int32_t)k;  // tiles in ascending order
    }
    sets.clear();
    D.tile_cells = tc;
    D.tile_maxtn = maxtn;
    D.tile_maxsub = maxsub;
    D.ntiles = nt;
    D.tile_total = total;
    D.t_cstart.upload(cst.data(), cst.size(), c.stream);
    if (D.fbar) D.t_fstart.upload(fst.data(), fst.size(), c.stream);
    D.lidx.upload(lidx.data(), lidx.size(), c.stream);
    D.t_nstart.upload(tstart.data(), tstart.size(), c.stream);
    D.t_node.upload(tnode.data(), tnode.size(), c.stream);
    D.nslot_start.upload(sstart.data(), sstart.size(), c.stream);
    D.nslot.upload(nslot.data(), nslot.size(), c.stream);
    D.slots.alloc(3 * std::max<int64_t>(total, 1));
    D.slots.zero(c.stream);
    APHIP_HIP(hipStreamSynchronize(c.stream));
    D.tiled = true;
}

// Reverse neighbor lists and the node order of the cell entries (after the cell CSR).
void rk_cells_order(Ctx &c, bool two_pass) {
    CellData &D = c.cells;
    const int64_t nn = c.nn;
    // reverse neighbor lists: for node J the points i that list it, ascending i,
    // with phi_iJ (the mass scatter and the two-pass force scatter)
    {
        std::vector<int32_t> hn(nn), hb(nn * kMaxNbr);
        std::vector<double> hp(nn * kMaxNbr);
        D.num_nbr.download(hn.data(), nn, c.stream);
        D.nbr.download(hb.data(), hb.size(), c.stream);
        D.phi.download(hp.data(), hp.size(), c.stream);
        APHIP_HIP(hipStreamSynchronize(c.stream));
        std::vector<int64_t> rstart(nn + 1, 0);
        for (int64_t i = 0; i < nn; ++i)
            for (int l = 0; l < hn[i]; ++l) rstart[hb[i * kMaxNbr + l] + 1]++;
        for (int64_t j = 0; j < nn; ++j) rstart[j + 1] += rstart[j];
        const int64_t nrev = std::max<int64_t>(rstart[nn], 1);
        std::vector<int32_t> rsrc(nrev, 0);
        std::vector<double> rphi(nrev, 0.0);
        std::vector<int64_t> fill(rstart.begin(), rstart.end() - 1);
        for (int64_t i = 0; i < nn; ++i)
            for (int l = 0; l < hn[i]; ++l) {
                const int64_t t = fill[hb[i * kMaxNbr + l]]++;
                rsrc[t] = (int32_t)i;
                rphi[t] = hp[i * kMaxNbr + l];
            }
        D.rstart.upload(rstart.data(), rstart.size(), c.stream);
        D.rsrc.upload(rsrc.data(), rsrc.size(), c.stream);
        D.rphi.upload(rphi.data(), rphi.size(), c.stream);
        APHIP_HIP(hipStreamSynchronize(c.stream));
    }

    // the one-pass force (F-bar included) runs on cell tiles; the two-pass force keeps
    // its entry forces in node order
    const bool tiled = !two_pass;
    if (tiled) {
        rk_tile_plan(c);
        D.inc_cell.release();
    }
    // node -> entries naming it, ascending entry (stable radix sort by node)
    if (!tiled) {
        DevBuf<int32_t> vals, keys_sorted, ccount;
        vals.alloc(std::max<int64_t>(D.nnz, 1));
        keys_sorted.alloc(std::max<int64_t>(D.nnz, 1));
        D.node_inc.release();
        DevBuf<int32_t> inc;
        inc.alloc(std::max<int64_t>(D.nnz, 1));
        hipLaunchKernelGGL(k_iota, grid1(D.nnz, 256), dim3(256), 0, c.stream, D.nnz, vals.p);
        int bits = 1;
        while (((int64_t)1 << bits) < nn) ++bits;
        stable_sort_pairs(c, D.nodes.p, keys_sorted.p, vals.p, inc.p, D.nnz, bits);
        ccount.alloc(nn);
        ccount.zero(c.stream);
        hipLaunchKernelGGL(k_count_keys, grid1(D.nnz, 256), dim3(256), 0, c.stream, D.nnz, D.nodes.p, ccount.p);
        scan_counts(c, ccount, nn, D.node_start, nullptr);
        D.entry_pos.alloc(std::max<int64_t>(D.nnz, 1));
        hipLaunchKernelGGL(k_invert_perm, grid1(D.nnz, 256), dim3(256), 0, c.stream, D.nnz, inc.p, D.entry_pos.p);
        APHIP_HIP(hipGetLastError());
    }
    if (two_pass) D.g.alloc(3 * std::max<int64_t>(D.nnz, 1));
    APHIP_HIP(hipStreamSynchronize(c.stream));
    D.ready = true;
}

template <int KIND>
static void launch_rk_tile(const RkTileArgs &T, int lpc, dim3 g, size_t lds, hipStream_t s) {
    if (lpc == 4) hipLaunchKernelGGL((k_rk_tile<KIND, 4, 1>), g, dim3(64), lds, s, T);
    else if (lpc == 8) hipLaunchKernelGGL((k_rk_tile<KIND, 8, 12>), g, dim3(64), lds, s, T);
    else hipLaunchKernelGGL((k_rk_tile<KIND, 16, 8>), g, dim3(64), lds, s, T);
}

static RkTileArgs rk_tile_args(Ctx &c) {
    CellData &D = c.cells;
    const int64_t nbt = ((D.ntiles + kNumXcd - 1) / kNumXcd) * kNumXcd;
    return RkTileArgs{D.ncells, D.ntiles, nbt, D.tile_maxtn, D.start.p, D.lidx.p, D.dphi.p,
                      D.volume.p, c.u.p, D.t_nstart.p, D.t_node.p, D.t_cstart.p, D.slots.p, c.Hnp1().p, c.Pnp1().p,
                      c.state[1 - c.state_np1].p, c.state[c.state_np1].p, c.mat.params(),
                      c.has_fields ? c.elem_lambda.p : nullptr, c.has_fields ? c.elem_mu.p : nullptr,
                      c.mat.rate_form() ? c.Hn().p : nullptr, c.mat.rate_form() ? c.Pn().p : nullptr, c.rate_dt};
}

template <int KIND>
static void launch_fbar_tile(const FbarTileArgs &B, int lpc, dim3 g, size_t lds, hipStream_t s) {
    if (lpc == 4) hipLaunchKernelGGL((k_fbar_tile<KIND, 4>), g, dim3(64), lds, s, B);
    else if (lpc == 8) hipLaunchKernelGGL((k_fbar_tile<KIND, 8>), g, dim3(64), lds, s, B);
    else hipLaunchKernelGGL((k_fbar_tile<KIND, 16>), g, dim3(64), lds, s, B);
}

static void rk_fbar_tiles(Ctx &c) {
    CellData &D = c.cells;
    if (D.ntiles == 0) return;
    const FbarTileArgs B{rk_tile_args(c), D.t_fstart.p, D.fcell_sub.p, D.fcell_volume.p, D.sub_fcell.p,
                         std::max(D.tile_maxsub, 1)};
    const dim3 g((unsigned)B.T.nblocks);
    const size_t lds = (size_t)tile_lds(D.tile_maxtn, D.tile_maxsub, true);
    const int lpc = D.tile_lpc;
    switch (c.mat.kind) {
        case APHIP_LINEAR_ELASTIC: launch_fbar_tile<APHIP_LINEAR_ELASTIC>(B, lpc, g, lds, c.stream); break;
        case APHIP_NEO_HOOKEAN: launch_fbar_tile<APHIP_NEO_HOOKEAN>(B, lpc, g, lds, c.stream); break;
        case APHIP_J2: launch_fbar_tile<APHIP_J2>(B, lpc, g, lds, c.stream); break;
        case APHIP_SVK: launch_fbar_tile<APHIP_SVK>(B, lpc, g, lds, c.stream); break;
        default: throw Error(APHIP_ERR_UNSUPPORTED, "F-bar takes the linear elastic, neo-Hookean, J2 and SVK materials");
    }
    APHIP_HIP(hipGetLastError());
}

static void rk_tile_force(Ctx &c) {
    CellData &D = c.cells;
    if (D.ntiles == 0) return;
    const RkTileArgs T = rk_tile_args(c);
    const dim3 g((unsigned)T.nblocks);
    const size_t lds = 48 * (size_t)std::max(D.tile_maxtn, 1);
    const int lpc = D.tile_lpc;
    switch (c.mat.kind) {
        case APHIP_LINEAR_ELASTIC: launch_rk_tile<APHIP_LINEAR_ELASTIC>(T, lpc, g, lds, c.stream); break;
        case APHIP_NEO_HOOKEAN: launch_rk_tile<APHIP_NEO_HOOKEAN>(T, lpc, g, lds, c.stream); break;
        case APHIP_J2: launch_rk_tile<APHIP_J2>(T, lpc, g, lds, c.stream); break;
        case APHIP_SVK: launch_rk_tile<APHIP_SVK>(T, lpc, g, lds, c.stream); break;
        case APHIP_NEO_HOOKEAN_DAMAGE: launch_rk_tile<APHIP_NEO_HOOKEAN_DAMAGE>(T, lpc, g, lds, c.stream); break;
        case APHIP_POWER_LAW_CREEP: launch_rk_tile<APHIP_POWER_LAW_CREEP>(T, lpc, g, lds, c.stream); break;
        case APHIP_DRUCKER_PRAGER: launch_rk_tile<APHIP_DRUCKER_PRAGER>(T, lpc, g, lds, c.stream); break;
        default: throw Error(APHIP_ERR_INVALID, "material not set");
    }
    APHIP_HIP(hipGetLastError());
}

// The two-pass cells of a Tet4 background are the elements themselves with the
// element's own gradient (the smoothed gradient of a linear tet), so their force is the
// Gauss Tet4 element force on the physical displacement: the element tile plan computes it
// with the deterministic LDS gather and no per-entry force array.  Not for the damage
// material, whose smoothed-cell update marks new failures differently
// (ComputeInternalForceSmoothedCell.h:219-223).
static bool two_pass_on_tiles(const Ctx &c) {
    return c.cells.two_pass && c.etype == APHIP_TET4 && c.tiles.ready && c.mat.kind != APHIP_NEO_HOOKEAN_DAMAGE;
}

void rk_cell_force(Ctx &c) {
    CellData &D = c.cells;
    if (D.two_pass) {  // physical displacement = Phi d (ExplicitSolver::ComputeForce, Explicit.cpp:148-152)
        const int64_t nbg = padded_groups(c.nn, kGsLanes);
        RkGatherArgs G{c.nn, nbg, D.num_nbr.p, D.nbr.p, D.phi.p, c.u.p, D.u_phys.p};
        hipLaunchKernelGGL(k_rk_gather<kGsLanes>, dim3((unsigned)nbg), dim3(256), 0, c.stream, G);
        APHIP_HIP(hipGetLastError());
        if (two_pass_on_tiles(c)) {
            const bool prof = c.prof.enabled;  // timed as this force pass, not as tile launches
            c.prof.enabled = false;
            tile_run(c, 0.0, 0, D.u_phys.p, D.f_local.p);
            c.prof.enabled = prof;
            return;
        }
    }
    if (D.ncells == 0) return;
    if (D.tiled && !D.fbar) {
        rk_tile_force(c);
        return;
    }
    const int64_t nbc = padded_groups(D.ncells, D.two_pass ? 1 : kCellLanes);
    CellForceArgs A{D.ncells, nbc, D.start.p, D.nodes.p, D.dphi.p, D.volume.p, D.two_pass ? D.u_phys.p : c.u.p,
                    c.Hnp1().p, c.Pnp1().p, D.g.p, D.entry_pos.p,
                    c.state[1 - c.state_np1].p, c.state[c.state_np1].p,
                    c.mat.params(),
                    c.has_fields ? c.elem_lambda.p : nullptr, c.has_fields ? c.elem_mu.p : nullptr,
                    c.mat.rate_form() ? c.Hn().p : nullptr, c.mat.rate_form() ? c.Pn().p : nullptr, c.rate_dt};
    const dim3 g((unsigned)nbc);
    if (D.fbar) {
        rk_fbar_tiles(c);
        return;
    }
    if (D.two_pass) {
        const double *X = c.X.p;
        const bool hex = c.etype == APHIP_HEX8;
        auto go = [&](auto k4, auto k8) {
            hipLaunchKernelGGL(hex ? k8 : k4, g, dim3(256), 0, c.stream, A, X);
        };
        switch (c.mat.kind) {
            case APHIP_LINEAR_ELASTIC:
                go(k_rk_cells_two_pass<APHIP_LINEAR_ELASTIC, 4>, k_rk_cells_two_pass<APHIP_LINEAR_ELASTIC, 8>);
                break;
            case APHIP_NEO_HOOKEAN:
                go(k_rk_cells_two_pass<APHIP_NEO_HOOKEAN, 4>, k_rk_cells_two_pass<APHIP_NEO_HOOKEAN, 8>);
                break;
            case APHIP_J2: go(k_rk_cells_two_pass<APHIP_J2, 4>, k_rk_cells_two_pass<APHIP_J2, 8>); break;
            case APHIP_SVK: go(k_rk_cells_two_pass<APHIP_SVK, 4>, k_rk_cells_two_pass<APHIP_SVK, 8>); break;
            case APHIP_NEO_HOOKEAN_DAMAGE:
                go(k_rk_cells_two_pass<APHIP_NEO_HOOKEAN_DAMAGE, 4>, k_rk_cells_two_pass<APHIP_NEO_HOOKEAN_DAMAGE, 8>);
                break;
            case APHIP_POWER_LAW_CREEP:
                go(k_rk_cells_two_pass<APHIP_POWER_LAW_CREEP, 4>, k_rk_cells_two_pass<APHIP_POWER_LAW_CREEP, 8>);
                break;
            case APHIP_DRUCKER_PRAGER:
                go(k_rk_cells_two_pass<APHIP_DRUCKER_PRAGER, 4>, k_rk_cells_two_pass<APHIP_DRUCKER_PRAGER, 8>);
                break;
            default: throw Error(APHIP_ERR_INVALID, "material not set");
        }
    }
    APHIP_HIP(hipGetLastError());
}

void rk_node_update(Ctx &c, double dt, int mode) {
    CellData &D = c.cells;
    const int64_t nbn = padded_groups(c.nn);
    const dim3 g((unsigned)nbn);
    RkNodeArgs A{c.nn, nbn, D.node_start.p, D.g.p, c.mass.p,
                 c.has_fixed ? c.fixed.p : nullptr, c.f.p, c.a.p, c.v.p, dt};
    if (D.two_pass) {
        // cell entries -> local force (unless the element tiles assembled it), then
        // f = Phi^T f_local with the node update (ScatterValues, Explicit.cpp:181-186)
        if (!two_pass_on_tiles(c)) {
            RkNodeArgs L = A;
            L.f = D.f_local.p;
            hipLaunchKernelGGL(k_rk_nodes<0>, g, dim3(256), 0, c.stream, L);
            APHIP_HIP(hipGetLastError());
        }
        RkScatterArgs S{A, D.num_nbr.p, D.rstart.p, D.rsrc.p, D.rphi.p, D.f_local.p};
        constexpr int NB = (kMaxNbr + kCellLanes - 1) / kCellLanes;
        if (mode == 0) hipLaunchKernelGGL((k_rk_scatter<0, kCellLanes, NB>), g, dim3(256), 0, c.stream, S);
        else if (mode == 1) hipLaunchKernelGGL((k_rk_scatter<1, kCellLanes, NB>), g, dim3(256), 0, c.stream, S);
        else hipLaunchKernelGGL((k_rk_scatter<2, kCellLanes, NB>), g, dim3(256), 0, c.stream, S);
        APHIP_HIP(hipGetLastError());
        return;
    }
    // nodes with degrees of freedom only (nodal cells put them first; the others keep
    // their values, as the reference's integrator runs over the active part)
    if (D.tiled) {
        constexpr int kSlotLanes = 8;
        const int64_t nbs = padded_groups(c.n_dof, kSlotLanes);
        RkSlotNodeArgs N{A, D.nslot_start.p, D.nslot.p, D.slots.p};
        N.out.nn = c.n_dof;
        N.out.nblocks = nbs;
        const dim3 gs((unsigned)nbs);
        if (mode == 0) hipLaunchKernelGGL((k_rk_slot_nodes<0, kSlotLanes>), gs, dim3(256), 0, c.stream, N);
        else if (mode == 1) hipLaunchKernelGGL((k_rk_slot_nodes<1, kSlotLanes>), gs, dim3(256), 0, c.stream, N);
        else hipLaunchKernelGGL((k_rk_slot_nodes<2, kSlotLanes>), gs, dim3(256), 0, c.stream, N);
        APHIP_HIP(hipGetLastError());
        return;
    }
    throw Error(APHIP_ERR_INVALID, "one-pass cells without a tile plan");
}

void rk_mass(Ctx &c, double *total) {
    CellData &D = c.cells;
    const int64_t nn = c.nn;
    DevBuf<double> mfe;
    mfe.alloc(nn);
    hipLaunchKernelGGL(c.etype == APHIP_HEX8 ? k_rk_mfe<8> : k_rk_mfe<4>, grid1(nn, 256), dim3(256), 0, c.stream, nn, c.node_start.p, c.node_inc.p,
                       D.nodal ? D.evol.p : D.volume.p, c.mat.density, mfe.p);
    APHIP_HIP(hipGetLastError());
    c.mass.alloc(nn);
    hipLaunchKernelGGL(k_rk_mass, grid1(nn, 256), dim3(256), 0, c.stream, nn, D.num_nbr.p, D.rstart.p, D.rsrc.p,
                       D.rphi.p, mfe.p, c.mass.p);
    APHIP_HIP(hipGetLastError());
    std::vector<double> hm(nn);
    c.mass.download(hm.data(), nn, c.stream);
    APHIP_HIP(hipStreamSynchronize(c.stream));
    double tot = 0.0;
    for (double x : hm) tot += x;
    if (total) *total = tot;
}

}  // namespace aphip
