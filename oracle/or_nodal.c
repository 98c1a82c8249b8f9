/*
 * or_nodal.c — nodal strain smoothing on a "thex" mesh (each tetrahedron split into four
 * hexahedra) with subcells and F-bar.  TEST INFRASTRUCTURE (oracle).
 *
 * Restates:
 *   active nodes       include/MeshLabelerProcessor.h:227-282 (each hex marks its node of
 *                      lowest id: the vertices of the original tets) and the check of
 *                      :285-330 (exactly one active node per hex)
 *   cells, subcells    include/MeshLabelerProcessor.h:690-832: one cell per active node made of
 *                      the hexes around it; subdomains 0: every hex its own subcell,
 *                      subdomains 1: the whole cell one subcell (LabelSubcellIds, :807-827)
 *   search range       include/NeighborSearchProcessor.h:196-260: every node is a query point,
 *                      only active nodes are candidates (spheres of the active nodes)
 *   subcell lists      include/SmoothedCellDataProcessor.h:407-539 over all hexes of a subcell,
 *   and derivatives    :542-675 (sum over the subcell's hexes of B_k V_e phi, then / V_sub;
 *                      cell volume = sum of its hexes' volumes)
 *   F-bar force        include/ComputeInternalForceSmoothedCell.h:120-413: subcell H and
 *                      J = DetApIm1(H) + 1; cell J-bar = sum J_s V_s / V_cell; H-bar =
 *                      (H + I) cbrt(J-bar / J_s) - I (unchanged when the ratio is within 3 eps
 *                      of 1); P-bar = stress(H-bar); dual pressure p = sum_s cbrt(.)/(3 J-bar)
 *                      V_s (P-bar : H + tr P-bar) / V_cell; P = P-bar s + (p - s/(3 J_s)
 *                      P-bar : F) J_s F^-T; f_k -= V_s P grad phi_k.  Used only when a cell
 *                      has more subcells than one (:458-459).
 *
 * Pinning (tests/test_oracle_nodal.py): the labelling on generated thex meshes (one active
 * node per hex = the original vertices, cells = the vertices' hex stars, cell volumes tile
 * the domain); the F-bar projection of the reference's FBarTest (test/unit_tests/
 * FbarTest.cpp: every F-bar has J = J-bar, and P is the derivative of the F-bar energy
 * sum_s V_s W(F-bar_s) with respect to F_s, checked by central differences with the
 * neo-Hookean energy whose derivative the stress is); the whole force as minus the gradient
 * of the F-bar strain energy with respect to the nodal displacements (central
 * differences); a linear displacement field gives H = G in every subcell and leaves F-bar
 * inactive (KinematicsTest.cpp:65-133 nodal cases); uniform shear of a thex brick gives the
 * patch-test face resultants (ForceTest.cpp:250-304 nodal and F-bar cases).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

#define REAL_EPS 2.2204460492503131e-16

int64_t or_nodal_active(int64_t n_nodes, int64_t n_hex, const int32_t *hexes, const int64_t *node_id, uint8_t *active) {
    memset(active, 0, (size_t)n_nodes);
    for (int64_t e = 0; e < n_hex; ++e) {
        int best = 0;
        for (int k = 1; k < 8; ++k) {
            const int64_t a = hexes[8 * e + k], b = hexes[8 * e + best];
            if ((node_id ? node_id[a] : a) < (node_id ? node_id[b] : b)) best = k;
        }
        active[hexes[8 * e + best]] = 1;
    }
    int64_t bad = 0;
    for (int64_t e = 0; e < n_hex; ++e) {
        int n = 0;
        for (int k = 0; k < 8; ++k) n += active[hexes[8 * e + k]] != 0;
        bad += n != 1;
    }
    return bad;
}

int64_t or_nodal_cells(int64_t n_nodes, int64_t n_hex, const int32_t *hexes, const uint8_t *active,
                       int32_t *hex_cell) {
    int32_t *rank = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n_nodes > 0 ? n_nodes : 1));
    if (!rank) return -1;
    int64_t nc = 0;
    for (int64_t i = 0; i < n_nodes; ++i) rank[i] = active[i] ? (int32_t)nc++ : -1;
    for (int64_t e = 0; e < n_hex; ++e) {
        hex_cell[e] = -1;
        for (int k = 0; k < 8; ++k)
            if (active[hexes[8 * e + k]]) hex_cell[e] = rank[hexes[8 * e + k]];
    }
    free(rank);
    return nc;
}

/* first-seen union over the subcell's elements, their nodes and the nodes' neighbors */
static int64_t subcell_union(const int32_t *elems, int64_t n_el, const int32_t *conn, int npe, int max_neighbors,
                             const int32_t *num, const int32_t *nbr, int32_t *out) {
    int64_t len = 0;
    for (int64_t q = 0; q < n_el; ++q)
        for (int k = 0; k < npe; ++k) {
            const int32_t a = conn[(int64_t)npe * elems[q] + k];
            for (int l = 0; l < num[a]; ++l) {
                const int32_t j = nbr[(int64_t)a * max_neighbors + l];
                int64_t t = 0;
                while (t < len && out[t] != j) ++t;
                if (t == len) {
                    if (out) out[len] = j;
                    ++len;
                }
            }
        }
    return len;
}

int64_t or_rk_subcells_count(int64_t n_sub, const int64_t *sub_start, const int32_t *sub_elems, const int32_t *conn,
                             int npe, int max_neighbors, const int32_t *num, const int32_t *nbr, int64_t *start) {
    int bad = 0;
#pragma omp parallel for schedule(dynamic, 256) reduction(| : bad)
    for (int64_t s = 0; s < n_sub; ++s) {
        const int64_t ne = sub_start[s + 1] - sub_start[s];
        int32_t *buf = (int32_t *)malloc(sizeof(int32_t) * (size_t)(ne * npe * max_neighbors + 1));
        if (!buf) {
            bad = 1;
            continue;
        }
        start[s + 1] = subcell_union(&sub_elems[sub_start[s]], ne, conn, npe, max_neighbors, num, nbr, buf);
        free(buf);
    }
    if (bad) return -1;
    start[0] = 0;
    for (int64_t s = 0; s < n_sub; ++s) start[s + 1] += start[s];
    return start[n_sub];
}

int or_rk_subcells_build(int64_t n_nodes, const double *X, int64_t n_sub, const int64_t *sub_start,
                         const int32_t *sub_elems, const int32_t *conn, int npe, int max_neighbors,
                         const int32_t *num, const int32_t *nbr, const double *phi, const int64_t *start,
                         int32_t *nodes, double *dphi, double *sub_volume, double *elem_volume) {
    (void)n_nodes;
    if (npe != 4 && npe != 8) return -1;
    int bad = 0;
#pragma omp parallel for schedule(dynamic, 256) reduction(| : bad)
    for (int64_t s = 0; s < n_sub; ++s) {
        const int32_t *els = &sub_elems[sub_start[s]];
        const int64_t n_el = sub_start[s + 1] - sub_start[s];
        int32_t *list = &nodes[start[s]];
        double *dp = &dphi[3 * start[s]];
        const int64_t len = subcell_union(els, n_el, conn, npe, max_neighbors, num, nbr, list);
        if (len != start[s + 1] - start[s]) {
            bad = 1;
            continue;
        }
        for (int64_t t = 0; t < 3 * len; ++t) dp[t] = 0.0;
        double vs = 0.0;
        for (int64_t q = 0; q < n_el; ++q) {
            const int32_t e = els[q];
            double Xe[24], B[24], V;
            for (int k = 0; k < npe; ++k)
                for (int d = 0; d < 3; ++d) Xe[3 * k + d] = X[3 * (int64_t)conn[(int64_t)npe * e + k] + d];
            if (npe == 4) or_smoothed_tet4_bmatrix(Xe, B, &V);
            else or_smoothed_hex8_bmatrix(Xe, B, &V);
            if (elem_volume) elem_volume[e] = V;
            vs += V;
            for (int k = 0; k < npe; ++k) {
                const int32_t a = conn[(int64_t)npe * e + k];
                for (int l = 0; l < num[a]; ++l) {
                    const int32_t j = nbr[(int64_t)a * max_neighbors + l];
                    const double fv = phi[(int64_t)a * max_neighbors + l];
                    int64_t t = 0;
                    while (list[t] != j) ++t;
                    for (int m = 0; m < 3; ++m) dp[3 * t + m] += B[3 * k + m] * V * fv;
                }
            }
        }
        for (int64_t t = 0; t < 3 * len; ++t) dp[t] /= vs;
        sub_volume[s] = vs;
    }
    return bad ? -1 : 0;
}

static void inverse3(const double A[9], double Ai[9]) {
    double det;
    or_inverse3_cofactor(A, Ai, &det);
}

int or_fbar_internal_force(int64_t n_nodes, int64_t n_cells, const int64_t *cell_sub, const double *cell_volume,
                           const int64_t *sub_start, const int32_t *sub_nodes, const double *sub_dphi,
                           const double *sub_volume, const or_material *mat, const double *u, const double *state_n,
                           double *state_np1, double *H, double *Hbar, double *P, double *Pbar, double *f,
                           int nthreads) {
    const int64_t n_sub = cell_sub[n_cells];
    const int sw = or_state_width(mat->kind);
    if (or_rate_form(mat->kind)) return -1;
    double *J = (double *)malloc(sizeof(double) * (size_t)(n_sub > 0 ? n_sub : 1));
    double *Hl = (double *)malloc(sizeof(double) * 9 * (size_t)(n_sub > 0 ? n_sub : 1));
    double *Hb = (double *)malloc(sizeof(double) * 9 * (size_t)(n_sub > 0 ? n_sub : 1));
    double *Pl = (double *)malloc(sizeof(double) * 9 * (size_t)(n_sub > 0 ? n_sub : 1));
    double *Pb = (double *)malloc(sizeof(double) * 9 * (size_t)(n_sub > 0 ? n_sub : 1));
    if (!J || !Hl || !Hb || !Pl || !Pb) {
        free(J); free(Hl); free(Hb); free(Pl); free(Pb);
        return -1;
    }
    const int nt = nthreads > 0 ? nthreads : 1;
    /* subcell displacement gradients and J (ComputeDisplacementGradientAndJ, :186-224) */
#pragma omp parallel for schedule(static) num_threads(nt)
    for (int64_t s = 0; s < n_sub; ++s) {
        double h[9] = {0};
        for (int64_t t = sub_start[s]; t < sub_start[s + 1]; ++t) {
            const int64_t a = sub_nodes[t];
            const double *b = &sub_dphi[3 * t];
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j) h[3 * i + j] += u[3 * a + i] * b[j];
        }
        memcpy(&Hl[9 * s], h, sizeof(h));
        J[s] = or_det_ap_im1(h) + 1.0;
    }
    /* per cell: J-bar, H-bar, stress, dual pressure, un-barred stress */
#pragma omp parallel for schedule(dynamic, 64) num_threads(nt)
    for (int64_t c = 0; c < n_cells; ++c) {
        double jbar = 0.0;
        for (int64_t s = cell_sub[c]; s < cell_sub[c + 1]; ++s) jbar = jbar + J[s] * sub_volume[s];
        jbar = jbar / cell_volume[c];
        for (int64_t s = cell_sub[c]; s < cell_sub[c + 1]; ++s) {
            const double ratio = jbar / J[s];
            double *hb = &Hb[9 * s];
            if (fabs(ratio - 1.0) < 3.0 * REAL_EPS) {
                memcpy(hb, &Hl[9 * s], 9 * sizeof(double));
            } else {
                const double js = cbrt(ratio);
                for (int k = 0; k < 9; ++k) hb[k] = Hl[9 * s + k] * js;
                for (int d = 0; d < 3; ++d) hb[4 * d] += js - 1.0;
            }
            const double lam = mat->elem_lambda ? mat->elem_lambda[s] : mat->lambda;
            const double mu = mat->elem_mu ? mat->elem_mu[s] : mat->mu;
            or_stress_mat(mat, lam, mu, hb, sw ? &state_n[sw * s] : NULL, sw ? &state_np1[sw * s] : NULL, &Pb[9 * s]);
        }
        double dual = 0.0;
        for (int64_t s = cell_sub[c]; s < cell_sub[c + 1]; ++s) {
            const double js = cbrt(jbar / J[s]);
            const double djb = js / (3.0 * jbar);
            const double *pb = &Pb[9 * s], *h = &Hl[9 * s];
            double ph = 0.0;
            for (int k = 0; k < 9; ++k) ph += pb[k] * h[k];
            dual += djb * sub_volume[s] * (ph + (pb[0] + pb[4] + pb[8]));
        }
        dual = dual / cell_volume[c];
        for (int64_t s = cell_sub[c]; s < cell_sub[c + 1]; ++s) {
            const double js = cbrt(jbar / J[s]);
            const double *pb = &Pb[9 * s];
            double F[9], Fi[9], pf = 0.0;
            for (int k = 0; k < 9; ++k) F[k] = Hl[9 * s + k];
            for (int d = 0; d < 3; ++d) F[4 * d] += 1.0;
            for (int k = 0; k < 9; ++k) pf += pb[k] * F[k];
            const double djj = -js / (3.0 * J[s]);
            const double scalar = (dual + djj * pf) * J[s];
            inverse3(F, Fi);
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j) Pl[9 * s + 3 * i + j] = pb[3 * i + j] * js + scalar * Fi[3 * j + i];
        }
    }
    /* force (ComputeInternalForce, :312-351): f_k -= V_s P grad phi_k */
    memset(f, 0, sizeof(double) * 3 * (size_t)n_nodes);
    for (int64_t s = 0; s < n_sub; ++s) {
        const double *p = &Pl[9 * s];
        const double V = sub_volume[s];
        double st[9];
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) st[3 * j + i] = p[3 * i + j] * -V; /* P^T (-V) */
        for (int64_t t = sub_start[s]; t < sub_start[s + 1]; ++t) {
            const int64_t a = sub_nodes[t];
            for (int i = 0; i < 3; ++i) {
                double acc = 0.0;
                for (int m = 0; m < 3; ++m) acc += sub_dphi[3 * t + m] * st[3 * m + i];
                f[3 * a + i] += acc;
            }
        }
    }
    if (H) memcpy(H, Hl, sizeof(double) * 9 * (size_t)n_sub);
    if (Hbar) memcpy(Hbar, Hb, sizeof(double) * 9 * (size_t)n_sub);
    if (P) memcpy(P, Pl, sizeof(double) * 9 * (size_t)n_sub);
    if (Pbar) memcpy(Pbar, Pb, sizeof(double) * 9 * (size_t)n_sub);
    free(J); free(Hl); free(Hb); free(Pl); free(Pb);
    return 0;
}
