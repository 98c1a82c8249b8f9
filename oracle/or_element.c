/*
 * or_element.c — element kinematics and the Gaussian internal-force loop.
 * TEST INFRASTRUCTURE (oracle).
 *
 * Restates:
 *   Tet4 shape-function derivatives       include/ShapeFunctionsFunctorTet4.h:27-35
 *   Hex8 shape-function derivatives       include/ShapeFunctionsFunctorHex8.h:51-62
 *   Gauss points / weights                include/ElementTetrahedron4.h:125-127 (xi=1/4, w=1/6),
 *                                         include/ElementHexahedron8.h:105-117 (+-1/sqrt3, w=1, listed order)
 *   B matrix and weight                   include/QuadratureGaussian.h:17-38
 *   smoothed Tet4 / Hex8 B and volume     include/QuadratureSmoothed.h:32-62, :96-150
 *   element loop (gather, H, P, -B P^T w) include/ComputeInternalForceGaussian.h:57-131
 *   element volume                        include/ComputeElementVolumeFunctor.h:47-69
 *
 * Pinning (tests/test_oracle_element.py): the Tet4 and smoothed Tet4/Hex8 golds of
 * test/unit_tests/QuadratureTests.cpp:17-285 (reference, scaled, translated, rotated,
 * random affine, sheared hex); Hex8 Gauss B has no reference gold and is pinned by
 * closed form: for an affine hex B_k = dN_k(xi) J^-1 exactly, sum_k B_k = 0,
 * sum_q w_q = volume; the whole force loop by the patch test closed forms
 * (PatchTestFixture.h:138-169): uniform H gives zero interior force and face
 * force sums P n A.
 */
#include <math.h>
#include <string.h>

#include "oracle.h"

static const double kTetDN[12] = {-1.0, -1.0, -1.0, 1.0, 0.0, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0, 1.0};

static void hex8_gauss_point(int g, double xi[3]) {
    static const int s[8][3] = {{-1, -1, -1}, {1, -1, -1}, {1, 1, -1}, {-1, 1, -1},
                                {-1, -1, 1}, {1, -1, 1}, {1, 1, 1}, {-1, 1, 1}};
    const double a = 1.0 / sqrt(3.0);
    for (int d = 0; d < 3; ++d) xi[d] = s[g][d] * a;
}

/* dN_k/dxi_j at parametric point p, ShapeFunctionsFunctorHex8.h:51-62 */
static void hex8_dN(const double p[3], double dN[24]) {
    static const int s[8][3] = {{-1, -1, -1}, {1, -1, -1}, {1, 1, -1}, {-1, 1, -1},
                                {-1, -1, 1}, {1, -1, 1}, {1, 1, 1}, {-1, 1, 1}};
    for (int k = 0; k < 8; ++k) {
        const double f0 = s[k][0] < 0 ? (1.0 - p[0]) : (1.0 + p[0]);
        const double f1 = s[k][1] < 0 ? (1.0 - p[1]) : (1.0 + p[1]);
        const double f2 = s[k][2] < 0 ? (1.0 - p[2]) : (1.0 + p[2]);
        dN[3 * k + 0] = (s[k][0] * 0.125) * f1 * f2;
        dN[3 * k + 1] = (s[k][1] * 0.125) * f0 * f2;
        dN[3 * k + 2] = (s[k][2] * 0.125) * f0 * f1;
    }
}

/* J = X^T dN, detJ, J^-1 (cofactor), B = dN J^-1, w = w_q detJ. */
static void bmatrix_from_dN(int nn, const double *X, const double *dN, double wq, double *B, double *w) {
    double J[9], Ji[9], det;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            double s = 0.0;
            for (int k = 0; k < nn; ++k) s += X[3 * k + i] * dN[3 * k + j];
            J[3 * i + j] = s;
        }
    or_inverse3_cofactor(J, Ji, &det);
    for (int k = 0; k < nn; ++k)
        for (int j = 0; j < 3; ++j) {
            double s = 0.0;
            for (int m = 0; m < 3; ++m) s += dN[3 * k + m] * Ji[3 * m + j];
            B[3 * k + j] = s;
        }
    *w = wq * det;
}

void or_tet4_bmatrix(const double X[12], double B[12], double *w) {
    bmatrix_from_dN(4, X, kTetDN, 1.0 / 6.0, B, w);
}

void or_hex8_bmatrix(const double X[24], int g, double B[24], double *w) {
    double p[3], dN[24];
    hex8_gauss_point(g, p);
    hex8_dN(p, dN);
    bmatrix_from_dN(8, X, dN, 1.0, B, w);
}

static void cross3(const double a[3], const double b[3], double c[3]) {
    c[0] = a[1] * b[2] - a[2] * b[1];
    c[1] = a[2] * b[0] - a[0] * b[2];
    c[2] = a[0] * b[1] - a[1] * b[0];
}

/* QuadratureSmoothed.h:32-62; faces per Seacas ordering (:15-18). */
void or_smoothed_tet4_bmatrix(const double X[12], double B[12], double *V) {
    static const int faces[4][3] = {{0, 1, 3}, {1, 2, 3}, {0, 3, 2}, {0, 2, 1}};
    double vol = 0.0;
    memset(B, 0, sizeof(double) * 12);
    for (int f = 0; f < 4; ++f) {
        double e1[3], e2[3], n[3];
        for (int d = 0; d < 3; ++d) {
            e1[d] = X[3 * faces[f][1] + d] - X[3 * faces[f][0] + d];
            e2[d] = X[3 * faces[f][2] + d] - X[3 * faces[f][0] + d];
        }
        cross3(e1, e2, n);
        for (int d = 0; d < 3; ++d) n[d] /= 2.0;
        const double *x0 = &X[3 * faces[f][0]];
        vol += x0[0] * n[0] + x0[1] * n[1] + x0[2] * n[2];
        for (int k = 0; k < 4; ++k)
            for (int l = 0; l < 3; ++l)
                if (k == faces[f][l]) {
                    for (int d = 0; d < 3; ++d) B[3 * k + d] += n[d] / 3.0;
                    break;
                }
    }
    vol /= 3.0;
    for (int i = 0; i < 12; ++i) B[i] /= vol;
    *V = vol;
}

/* QuadratureSmoothed.h:96-150. */
void or_smoothed_hex8_bmatrix(const double X[24], double B[24], double *V) {
    static const int faces[6][4] = {{0, 1, 5, 4}, {1, 2, 6, 5}, {2, 3, 7, 6},
                                    {0, 4, 7, 3}, {0, 3, 2, 1}, {4, 5, 6, 7}};
    double vol = 0.0;
    memset(B, 0, sizeof(double) * 24);
    for (int f = 0; f < 6; ++f) {
        double mid[3] = {0.0, 0.0, 0.0};
        for (int k = 0; k < 4; ++k)
            for (int d = 0; d < 3; ++d) mid[d] += X[3 * faces[f][k] + d];
        for (int d = 0; d < 3; ++d) mid[d] /= 4.0;
        int km1 = 3;
        for (int k = 0; k < 4; ++k) {
            double e1[3], e2[3], n[3];
            for (int d = 0; d < 3; ++d) {
                e1[d] = X[3 * faces[f][km1] + d] - mid[d];
                e2[d] = X[3 * faces[f][k] + d] - mid[d];
            }
            cross3(e1, e2, n);
            for (int d = 0; d < 3; ++d) n[d] /= 2.0;
            vol += mid[0] * n[0] + mid[1] * n[1] + mid[2] * n[2];
            for (int l = 0; l < 4; ++l) {
                const double c = (l == k || l == km1) ? 5.0 / 12.0 : 1.0 / 12.0;
                for (int d = 0; d < 3; ++d) B[3 * faces[f][l] + d] += n[d] * c;
            }
            km1 = k;
        }
    }
    vol /= 3.0;
    for (int i = 0; i < 24; ++i) B[i] /= vol;
    *V = vol;
}

void or_element_volumes(int etype, int64_t n_elem, const int32_t *conn, const double *X, double *vol) {
    const int npe = etype == OR_TET4 ? 4 : 8;
    const int nq = etype == OR_TET4 ? 1 : 8;
    for (int64_t e = 0; e < n_elem; ++e) {
        double Xe[24], B[24], w, v = 0.0;
        for (int k = 0; k < npe; ++k)
            for (int d = 0; d < 3; ++d) Xe[3 * k + d] = X[3 * (int64_t)conn[e * npe + k] + d];
        for (int q = 0; q < nq; ++q) {
            if (etype == OR_TET4) or_tet4_bmatrix(Xe, B, &w);
            else or_hex8_bmatrix(Xe, q, B, &w);
            v += w;
        }
        vol[e] = v;
    }
}

int or_internal_force(int etype, int64_t n_nodes, const double *X, int64_t n_elem, const int32_t *conn,
                      const or_material *mat, const double *u, const double *state_n, double *state_np1,
                      double *H, double *P, double *f, int nthreads) {
    if (etype != OR_TET4 && etype != OR_HEX8) return -1;
    const int npe = etype == OR_TET4 ? 4 : 8;
    const int nq = etype == OR_TET4 ? 1 : 8;
    const int sw = or_state_width(mat->kind);
    const int has_state = sw > 0;
    const int per_qp = has_state && mat->state_per_qp && etype == OR_HEX8;
    memset(f, 0, sizeof(double) * 3 * n_nodes);
#pragma omp parallel for schedule(static) num_threads(nthreads > 0 ? nthreads : 1)
    for (int64_t e = 0; e < n_elem; ++e) {
        double Xe[24], ue[24], fe[24], B[24], w, He[9], Pe[9];
        const double lam = mat->elem_lambda ? mat->elem_lambda[e] : mat->lambda;
        const double mu = mat->elem_mu ? mat->elem_mu[e] : mat->mu;
        for (int k = 0; k < npe; ++k) {
            const int64_t n = conn[e * npe + k];
            for (int d = 0; d < 3; ++d) {
                Xe[3 * k + d] = X[3 * n + d];
                ue[3 * k + d] = u[3 * n + d];
            }
        }
        memset(fe, 0, sizeof(fe));
        for (int q = 0; q < nq; ++q) {
            if (etype == OR_TET4) or_tet4_bmatrix(Xe, B, &w);
            else or_hex8_bmatrix(Xe, q, B, &w);
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j) {
                    double s = 0.0;
                    for (int k = 0; k < npe; ++k) s += ue[3 * k + i] * B[3 * k + j];
                    He[3 * i + j] = s;
                }
            if (mat->updated) { /* H = dH + H_n + dH H_n (ComputeInternalForceBase.h:158-166) */
                const double *Hn = &mat->H_n[9 * e];
                double Ht[9];
                for (int i = 0; i < 3; ++i)
                    for (int j = 0; j < 3; ++j) {
                        double s = 0.0;
                        for (int k = 0; k < 3; ++k) s += He[3 * i + k] * Hn[3 * k + j];
                        Ht[3 * i + j] = He[3 * i + j] + Hn[3 * i + j] + s;
                    }
                memcpy(He, Ht, sizeof(Ht));
            }
            const int64_t sidx = (per_qp ? e * 8 + q : e) * sw;
            if (or_rate_form(mat->kind))  /* the element's H_n, P_n for every Gauss point */
                or_stress_rate(mat, lam, mu, He, &mat->H_n[9 * e], &mat->P_n[9 * e], mat->dt, &state_n[sidx],
                               &state_np1[sidx], Pe);
            else
                or_stress_mat(mat, lam, mu, He, has_state ? &state_n[sidx] : NULL,
                              has_state ? &state_np1[sidx] : NULL, Pe);
            if (mat->updated) { /* B F_r and w / det F_r (ComputeInternalForceGaussian.h:109-123) */
                double Fr[9], Fri[9], dFr, Bp[24];
                for (int i = 0; i < 9; ++i) Fr[i] = mat->H_n[9 * e + i];
                Fr[0] += 1.0; Fr[4] += 1.0; Fr[8] += 1.0;
                or_inverse3_cofactor(Fr, Fri, &dFr);
                for (int k = 0; k < npe; ++k)
                    for (int j = 0; j < 3; ++j) {
                        double s = 0.0;
                        for (int m = 0; m < 3; ++m) s += B[3 * k + m] * Fr[3 * m + j];
                        Bp[3 * k + j] = s;
                    }
                memcpy(B, Bp, sizeof(double) * 3 * npe);
                w = w / dFr;
            }
            for (int k = 0; k < npe; ++k)
                for (int j = 0; j < 3; ++j) {
                    double s = 0.0;
                    for (int m = 0; m < 3; ++m) s += B[3 * k + m] * Pe[3 * j + m];
                    fe[3 * k + j] -= s * w;
                }
        }
        /* the stored H/P are those of the last Gauss point (ComputeInternalForceGaussian.h:89,99-107) */
        if (H) memcpy(&H[9 * e], He, sizeof(He));
        if (P) memcpy(&P[9 * e], Pe, sizeof(Pe));
        for (int k = 0; k < npe; ++k) {
            const int64_t n = conn[e * npe + k];
            for (int d = 0; d < 3; ++d) {
#pragma omp atomic
                f[3 * n + d] += fe[3 * k + d];
            }
        }
    }
    (void)n_nodes;
    return 0;
}
