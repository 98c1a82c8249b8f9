/*
 * or_rk.c — RKPM setup and the strain-smoothed cell force.  TEST INFRASTRUCTURE (oracle).
 *
 * Restates:
 *   max edge length     src/MaxEdgeLengthProcessor.cpp:36-60 (longest distance to any node of a
 *                       connected element, doubled for inactive nodes)
 *   ball search         include/NeighborSearchProcessor.h:196-317, src/NeighborSearchProcessor.cpp:
 *                       154-240: j is a neighbor of point i when |x_i - x_j| <= r_j (sphere of the
 *                       neighbor contains the point); candidates sorted by (tweaked d^2, id), the
 *                       neighbor coordinates scaled by (1+1e-6, 1+1e-7, 1+1e-8); first
 *                       MAX_NODE_NUM_NEIGHBORS kept (include/NeighborSelectorFunctor.h:63-99,
 *                       include/Constants.h:15)
 *   RK function values  include/FunctionCreationProcessor.h:85-130,
 *                       include/ShapeFunctionsFunctorReproducingKernel.h:190-257 (linear basis,
 *                       kernel of the neighbor's radius, shifted = x_i - x_j, Shepard fallback)
 *   cell node lists     include/SmoothedCellDataProcessor.h:407-539 (one-pass: first-seen union of
 *                       the neighbors of the cell element's nodes)
 *   cell derivatives    include/SmoothedCellDataProcessor.h:542-675 with SmoothedQuadratureTet4
 *   cell force          include/ComputeInternalForceSmoothedCell.h:86-117, :186-224, :312-351, :416-442
 *   RK mass             src/MassUtils.cpp:79-112 with FunctionEvaluationProcessor::ScatterValues
 *                       (include/FunctionEvaluationProcessor.h:130-181)
 *   two-pass method     include/SmoothedCellDataProcessor.h:474-486, :637-648 (cells list the
 *                       element's own nodes, derivative (B_k V)/V), FunctionEvaluationProcessor::
 *                       ComputeValues (include/FunctionEvaluationProcessor.h:70-117, neighbors
 *                       taken in descending order) and ScatterValues (:130-181), called around
 *                       the internal force by ExplicitSolver::ComputeForce (src/Solvers/
 *                       Explicit.cpp:148-186)
 *
 * Pinning (tests/test_oracle_rk.py): neighbor counts of NeighborSearchProcessorTest.cpp:68-207
 * (radii 0.5, 1.01, 1.01*diagonal, variable radius 1.1 and 1.01 incl. the sqrt2/sqrt3 radius
 * counts), grid search == brute-force O(N^2) search on jittered lattices; phi: partition of
 * unity and linear reproduction sum phi_j x_j = x_i (ShapeFunctionsFunctorsTest.cpp:182-272),
 * tet-only neighbors reproduce the FE values; cell derivatives: partition of nullity and a
 * linear field u = (F-I)X giving H = F-I exactly (KinematicsTestFixture.h:338-357).
 * Two-pass: gather of a linear coefficient field d = G X + c returns G X + c at every node
 * (linear reproduction) and the cells then give H = G exactly; the two-pass force Phi^T f(Phi d)
 * equals the one-pass force of the same d (the one-pass cell derivative is the same sum taken
 * in the other order), which fails for a transposed Phi in either the gather or the scatter.
 * Inclusion ties on the sphere boundary follow the third-party STK intersection test and
 * are parity unpinned (jittered inputs make them measure-zero).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

/* |a - b|^2 with one fixed fused evaluation order, so that every build of the
 * search agrees to the bit on inclusion and ranking: fma(dz, dz, fma(dy, dy, dx*dx)). */
static double dist2(const double *a, const double *b) {
    const double dx = a[0] - b[0], dy = a[1] - b[1], dz = a[2] - b[2];
    return fma(dz, dz, fma(dy, dy, dx * dx));
}

int or_max_edge_length(int64_t n_nodes, const double *X, int64_t n_elem, const int32_t *conn, int npe,
                       const uint8_t *active, double *max_edge) {
    memset(max_edge, 0, sizeof(double) * n_nodes);
    for (int64_t e = 0; e < n_elem; ++e)
        for (int a = 0; a < npe; ++a) {
            const int64_t i = conn[e * npe + a];
            for (int b = 0; b < npe; ++b) {
                const int64_t j = conn[e * npe + b];
                const double s = (active && !active[j]) ? 2.0 : 1.0;
                const double len = sqrt(dist2(&X[3 * i], &X[3 * j])) * s;
                if (len > max_edge[i]) max_edge[i] = len;
            }
        }
    return 0;
}

typedef struct {
    double d2;
    int32_t id;
} cand_t;

static int cand_cmp(const void *pa, const void *pb) {
    const cand_t *a = (const cand_t *)pa, *b = (const cand_t *)pb;
    if (a->d2 == b->d2) return (a->id > b->id) - (a->id < b->id);
    return (a->d2 > b->d2) - (a->d2 < b->d2);
}

static double tweaked_d2(const double *xi, const double *xj) {
    const double tw[3] = {xj[0] * (1 + 1.0e-6), xj[1] * (1 + 1.0e-7), xj[2] * (1 + 1.0e-8)};
    return dist2(tw, xi);
}

static int contains(const double *xi, const double *xj, double rj) {
    return dist2(xj, xi) <= rj * rj;
}

int or_neighbor_search(int64_t n_nodes, const double *X, const double *radius, int max_neighbors,
                       int brute_force, int32_t *num_neighbors, int32_t *neighbors, int64_t *n_truncated) {
    return or_neighbor_search_masked(n_nodes, X, radius, max_neighbors, NULL, brute_force, num_neighbors, neighbors,
                                     n_truncated);
}

int or_neighbor_search_masked(int64_t n_nodes, const double *X, const double *radius, int max_neighbors,
                              const uint8_t *range_mask, int brute_force, int32_t *num_neighbors,
                              int32_t *neighbors, int64_t *n_truncated) {
    int64_t trunc = 0;
    int64_t cap = 256;
    cand_t *c = (cand_t *)malloc(sizeof(cand_t) * cap);
    if (!c) return -1;
    double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY}, rmax = 0.0;
    for (int64_t i = 0; i < n_nodes; ++i) {
        for (int d = 0; d < 3; ++d) {
            if (X[3 * i + d] < lo[d]) lo[d] = X[3 * i + d];
            if (X[3 * i + d] > hi[d]) hi[d] = X[3 * i + d];
        }
        if ((!range_mask || range_mask[i]) && radius[i] > rmax) rmax = radius[i];
    }
    int64_t nb[3] = {1, 1, 1};
    int64_t *bin_start = NULL, *bin_items = NULL;
    double cell = rmax > 0.0 ? rmax : 1.0;
    if (!brute_force) {
        for (int d = 0; d < 3; ++d) {
            nb[d] = (int64_t)floor((hi[d] - lo[d]) / cell) + 1;
            if (nb[d] < 1) nb[d] = 1;
            if (nb[d] > 1024) nb[d] = 1024;
        }
        const int64_t nbins = nb[0] * nb[1] * nb[2];
        bin_start = (int64_t *)calloc((size_t)nbins + 1, sizeof(int64_t));
        bin_items = (int64_t *)malloc(sizeof(int64_t) * (size_t)(n_nodes > 0 ? n_nodes : 1));
        int64_t *fill = (int64_t *)calloc((size_t)nbins, sizeof(int64_t));
        int64_t *binof = (int64_t *)malloc(sizeof(int64_t) * (size_t)(n_nodes > 0 ? n_nodes : 1));
        for (int64_t i = 0; i < n_nodes; ++i) {
            int64_t b[3];
            for (int d = 0; d < 3; ++d) {
                b[d] = (int64_t)floor((X[3 * i + d] - lo[d]) / cell);
                if (b[d] < 0) b[d] = 0;
                if (b[d] >= nb[d]) b[d] = nb[d] - 1;
            }
            binof[i] = (b[2] * nb[1] + b[1]) * nb[0] + b[0];
            if (!range_mask || range_mask[i]) bin_start[binof[i] + 1]++;
        }
        for (int64_t b = 0; b < nbins; ++b) bin_start[b + 1] += bin_start[b];
        for (int64_t i = 0; i < n_nodes; ++i)
            if (!range_mask || range_mask[i]) bin_items[bin_start[binof[i]] + fill[binof[i]]++] = i;
        free(fill);
        free(binof);
    }
    for (int64_t i = 0; i < n_nodes; ++i) {
        const double *xi = &X[3 * i];
        int64_t nc = 0;
        if (brute_force) {
            for (int64_t j = 0; j < n_nodes; ++j)
                if ((!range_mask || range_mask[j]) && contains(xi, &X[3 * j], radius[j])) {
                    if (nc == cap) { cap *= 2; c = (cand_t *)realloc(c, sizeof(cand_t) * cap); }
                    c[nc].d2 = tweaked_d2(xi, &X[3 * j]);
                    c[nc].id = (int32_t)j;
                    ++nc;
                }
        } else {
            int64_t b[3];
            for (int d = 0; d < 3; ++d) {
                b[d] = (int64_t)floor((xi[d] - lo[d]) / cell);
                if (b[d] < 0) b[d] = 0;
                if (b[d] >= nb[d]) b[d] = nb[d] - 1;
            }
            for (int64_t bz = b[2] - 1; bz <= b[2] + 1; ++bz)
                for (int64_t by = b[1] - 1; by <= b[1] + 1; ++by)
                    for (int64_t bx = b[0] - 1; bx <= b[0] + 1; ++bx) {
                        if (bx < 0 || by < 0 || bz < 0 || bx >= nb[0] || by >= nb[1] || bz >= nb[2]) continue;
                        const int64_t bb = (bz * nb[1] + by) * nb[0] + bx;
                        for (int64_t t = bin_start[bb]; t < bin_start[bb + 1]; ++t) {
                            const int64_t j = bin_items[t];
                            if (!contains(xi, &X[3 * j], radius[j])) continue;
                            if (nc == cap) { cap *= 2; c = (cand_t *)realloc(c, sizeof(cand_t) * cap); }
                            c[nc].d2 = tweaked_d2(xi, &X[3 * j]);
                            c[nc].id = (int32_t)j;
                            ++nc;
                        }
                    }
        }
        qsort(c, (size_t)nc, sizeof(cand_t), cand_cmp);
        const int64_t keep = nc < max_neighbors ? nc : max_neighbors;
        if (nc > keep) trunc += nc - keep;
        num_neighbors[i] = (int32_t)keep;
        for (int64_t k = 0; k < keep; ++k) neighbors[i * max_neighbors + k] = c[k].id;
        for (int64_t k = keep; k < max_neighbors; ++k) neighbors[i * max_neighbors + k] = -1;
    }
    free(c);
    free(bin_start);
    free(bin_items);
    if (n_truncated) *n_truncated = trunc;
    return 0;
}

int or_rk_function_values(int64_t n_nodes, const double *X, const double *radius, int max_neighbors,
                          const int32_t *num_neighbors, const int32_t *neighbors, double *phi) {
    for (int64_t i = 0; i < n_nodes; ++i) {
        const int n = num_neighbors[i];
        double *ph = &phi[i * max_neighbors];
        for (int k = 0; k < max_neighbors; ++k) ph[k] = 0.0;
        if (n == 0) continue;
        if (n == 1) { ph[0] = 1.0; continue; }
        double sh[64][3], w[64];
        for (int k = 0; k < n; ++k) {
            const int64_t j = neighbors[i * max_neighbors + k];
            for (int d = 0; d < 3; ++d) sh[k][d] = X[3 * i + d] - X[3 * j + d];
            w[k] = or_kernel_value(sh[k], radius[j]);
        }
        double M[16] = {0}, Mi[16], wsum = 0.0;
        for (int k = 0; k < n; ++k) {
            if (w[k] == 0.0) continue;
            const double Hb[4] = {1.0, sh[k][0], sh[k][1], sh[k][2]};
            for (int a = 0; a < 4; ++a)
                for (int b = 0; b < 4; ++b) M[4 * a + b] += (w[k] * Hb[a]) * Hb[b];
            wsum += w[k];
        }
        or_fullpivlu_inverse(4, M, Mi);
        double fro = 0.0;
        for (int a = 0; a < 16; ++a) fro += Mi[a] * Mi[a];
        if (sqrt(fro) < 1e-15) {
            for (int k = 0; k < n; ++k) ph[k] = w[k] / wsum;
            continue;
        }
        for (int k = 0; k < n; ++k) {
            if (w[k] == 0.0) continue;
            const double Hb[4] = {1.0, sh[k][0], sh[k][1], sh[k][2]};
            double s = 0.0;
            for (int a = 0; a < 4; ++a) s += Mi[a] * Hb[a];
            ph[k] = s * w[k];
        }
    }
    return 0;
}

/* First-seen union of the neighbors of the npe element nodes; returns the list length. */
static int cell_union(const int32_t *en, int npe, int max_neighbors, const int32_t *num_neighbors,
                      const int32_t *neighbors, int32_t *out) {
    int len = 0;
    for (int k = 0; k < npe; ++k) {
        const int32_t a = en[k];
        for (int l = 0; l < num_neighbors[a]; ++l) {
            const int32_t j = neighbors[(int64_t)a * max_neighbors + l];
            int found = 0;
            for (int t = 0; t < len; ++t)
                if (out[t] == j) { found = 1; break; }
            if (!found) out[len++] = j;
        }
    }
    return len;
}

int64_t or_rk_cells_count(int64_t n_elem, const int32_t *conn, int npe, int max_neighbors,
                          const int32_t *num_neighbors, const int32_t *neighbors, int64_t *cell_start) {
    if (npe * max_neighbors > 8 * 64) return -1;
    cell_start[0] = 0;
    /* lengths in parallel (cells are independent), then the running sum */
#pragma omp parallel for schedule(dynamic, 4096)
    for (int64_t e = 0; e < n_elem; ++e) {
        int32_t buf[8 * 64];
        cell_start[e + 1] = cell_union(&conn[npe * e], npe, max_neighbors, num_neighbors, neighbors, buf);
    }
    for (int64_t e = 0; e < n_elem; ++e) cell_start[e + 1] += cell_start[e];
    return cell_start[n_elem];
}

int or_rk_cells_build(int64_t n_nodes, const double *X, int64_t n_elem, const int32_t *conn, int npe,
                      int max_neighbors, const int32_t *num_neighbors, const int32_t *neighbors, const double *phi,
                      const int64_t *cell_start, int32_t *cell_nodes, double *cell_dphi, double *cell_volume) {
    (void)n_nodes;
    if (npe != 4 && npe != 8) return -1;
    int bad = 0;
    /* cells are independent: each writes only its own CSR range */
#pragma omp parallel for schedule(dynamic, 4096) reduction(| : bad)
    for (int64_t e = 0; e < n_elem; ++e) {
        int32_t *list = &cell_nodes[cell_start[e]];
        double *dphi = &cell_dphi[3 * cell_start[e]];
        const int len = cell_union(&conn[npe * e], npe, max_neighbors, num_neighbors, neighbors, list);
        if (len != cell_start[e + 1] - cell_start[e]) {
            bad = 1;
            continue;
        }
        double Xe[24], B[24], V;
        for (int k = 0; k < npe; ++k)
            for (int d = 0; d < 3; ++d) Xe[3 * k + d] = X[3 * (int64_t)conn[npe * e + k] + d];
        /* the element's smoothed B and volume (SmoothedCellDataProcessor.h:608 with
         * QuadratureSmoothed.h:32-62 for Tet4, :96-150 for Hex8) */
        if (npe == 4) or_smoothed_tet4_bmatrix(Xe, B, &V);
        else or_smoothed_hex8_bmatrix(Xe, B, &V);
        for (int t = 0; t < 3 * len; ++t) dphi[t] = 0.0;
        for (int k = 0; k < npe; ++k) {
            const int32_t a = conn[npe * e + k];
            for (int l = 0; l < num_neighbors[a]; ++l) {
                const int32_t j = neighbors[(int64_t)a * max_neighbors + l];
                const double fv = phi[(int64_t)a * max_neighbors + l];
                int t = 0;
                while (list[t] != j) ++t;
                for (int m = 0; m < 3; ++m) dphi[3 * t + m] += B[3 * k + m] * V * fv;
            }
        }
        for (int t = 0; t < 3 * len; ++t) dphi[t] /= V;
        cell_volume[e] = V;
    }
    return bad ? -1 : 0;
}

int or_cell_internal_force(int64_t n_nodes, int64_t n_cells, const int64_t *cell_start, const int32_t *cell_nodes,
                           const double *cell_dphi, const double *cell_volume, const or_material *mat,
                           const double *u, const double *state_n, double *state_np1, double *H, double *P,
                           double *f, int nthreads) {
    const int sw = or_state_width(mat->kind);
    const int has_state = sw > 0;
    memset(f, 0, sizeof(double) * 3 * n_nodes);
#pragma omp parallel for schedule(static) num_threads(nthreads > 0 ? nthreads : 1)
    for (int64_t c = 0; c < n_cells; ++c) {
        const double lam = mat->elem_lambda ? mat->elem_lambda[c] : mat->lambda;
        const double mu = mat->elem_mu ? mat->elem_mu[c] : mat->mu;
        double He[9] = {0}, Pe[9], st[9];
        for (int64_t t = cell_start[c]; t < cell_start[c + 1]; ++t) {
            const int64_t j = cell_nodes[t];
            for (int a = 0; a < 3; ++a)
                for (int b = 0; b < 3; ++b) He[3 * a + b] += u[3 * j + a] * cell_dphi[3 * t + b];
        }
        if (or_rate_form(mat->kind))
            or_stress_rate(mat, lam, mu, He, &mat->H_n[9 * c], &mat->P_n[9 * c], mat->dt, &state_n[c * sw],
                           &state_np1[c * sw], Pe);
        else
            or_stress_mat(mat, lam, mu, He, has_state ? &state_n[c * sw] : NULL,
                          has_state ? &state_np1[c * sw] : NULL, Pe);
        /* a new failure's marker becomes the maximum damage in the smoothed-cell path
         * (ComputeInternalForceSmoothedCell.h:219-223, NeoHookeWithDamage.h:247-252) */
        if (mat->kind == OR_NH_DAMAGE && state_np1[c * sw] == 2.0)
            state_np1[c * sw] = mat->maximum_damage < 0.0 ? 0.0 : (mat->maximum_damage > 1.0 ? 1.0 : mat->maximum_damage);
        if (H) memcpy(&H[9 * c], He, sizeof(He));
        if (P) memcpy(&P[9 * c], Pe, sizeof(Pe));
        const double V = cell_volume[c];
        /* stress_term = P^T * (-V) */
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b) st[3 * a + b] = Pe[3 * b + a] * -V;
        for (int64_t t = cell_start[c]; t < cell_start[c + 1]; ++t) {
            const int64_t j = cell_nodes[t];
            for (int b = 0; b < 3; ++b) {
                double s = 0.0;
                for (int a = 0; a < 3; ++a) s += cell_dphi[3 * t + a] * st[3 * a + b];
#pragma omp atomic
                f[3 * j + b] += s;
            }
        }
    }
    return 0;
}

int or_rk_mass(int64_t n_nodes, int64_t n_elem, const int32_t *conn, int npe, const double *elem_volume,
               double density, int max_neighbors, const int32_t *num_neighbors, const int32_t *neighbors, const double *phi,
               double *mass) {
    double *mfe = (double *)calloc((size_t)n_nodes, sizeof(double));
    if (!mfe) return -1;
    for (int64_t e = 0; e < n_elem; ++e) {
        const double m = elem_volume[e] * density;
        for (int k = 0; k < npe; ++k) mfe[conn[npe * e + k]] += m / npe;
    }
    memset(mass, 0, sizeof(double) * n_nodes);
    for (int64_t i = 0; i < n_nodes; ++i) {
        const int n = num_neighbors[i];
        if (n == 0) { mass[i] = mfe[i]; continue; }
        for (int k = 0; k < n; ++k) mass[neighbors[i * max_neighbors + k]] += mfe[i] * phi[i * max_neighbors + k];
    }
    free(mfe);
    return 0;
}

int or_rk_cells_two_pass(int64_t n_elem, const double *X, const int32_t *conn, int npe, int64_t *cell_start,
                         int32_t *cell_nodes, double *cell_dphi, double *cell_volume) {
    if (npe != 4 && npe != 8) return -1;
    cell_start[0] = 0;
    for (int64_t e = 0; e < n_elem; ++e) {
        double Xe[24], B[24], V;
        for (int k = 0; k < npe; ++k)
            for (int d = 0; d < 3; ++d) Xe[3 * k + d] = X[3 * (int64_t)conn[npe * e + k] + d];
        if (npe == 4) or_smoothed_tet4_bmatrix(Xe, B, &V);
        else or_smoothed_hex8_bmatrix(Xe, B, &V);
        for (int k = 0; k < npe; ++k) {
            cell_nodes[npe * e + k] = conn[npe * e + k];
            for (int m = 0; m < 3; ++m) cell_dphi[3 * (npe * e + k) + m] = (B[3 * k + m] * V) / V;
        }
        cell_volume[e] = V;
        cell_start[e + 1] = npe * (e + 1);
    }
    return 0;
}

int or_rk_gather(int64_t n_nodes, int max_neighbors, const int32_t *num_neighbors, const int32_t *neighbors,
                 const double *phi, const double *src, double *dst) {
    for (int64_t i = 0; i < n_nodes; ++i) {
        const int n = num_neighbors[i];
        if (n == 0) {
            for (int d = 0; d < 3; ++d) dst[3 * i + d] = src[3 * i + d];
            continue;
        }
        double s[3] = {0.0, 0.0, 0.0};
        for (int k = n - 1; k >= 0; --k) {
            const int64_t j = neighbors[i * max_neighbors + k];
            const double w = phi[i * max_neighbors + k];
            for (int d = 0; d < 3; ++d) s[d] += src[3 * j + d] * w;
        }
        for (int d = 0; d < 3; ++d) dst[3 * i + d] = s[d];
    }
    return 0;
}

int or_rk_scatter(int64_t n_nodes, int max_neighbors, const int32_t *num_neighbors, const int32_t *neighbors,
                  const double *phi, const double *src, double *dst) {
    memset(dst, 0, sizeof(double) * 3 * n_nodes);
    for (int64_t i = 0; i < n_nodes; ++i) {
        const int n = num_neighbors[i];
        for (int k = 0; k < n; ++k) {
            const int64_t j = neighbors[i * max_neighbors + k];
            const double w = phi[i * max_neighbors + k];
            for (int d = 0; d < 3; ++d) dst[3 * j + d] += src[3 * i + d] * w;
        }
    }
    for (int64_t i = 0; i < n_nodes; ++i)
        if (num_neighbors[i] == 0)
            for (int d = 0; d < 3; ++d) dst[3 * i + d] = src[3 * i + d];
    return 0;
}
