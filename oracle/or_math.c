/*
 * or_math.c — small dense math used by the path.  TEST INFRASTRUCTURE (oracle).
 *
 * Pinning:
 *   or_elastic_properties  — LinearElasticPropertiesTest.cpp:16-179 values (tests/test_oracle_math.py)
 *   or_det_ap_im1          — closed form det(I+A)-1 on random A, MathUtilsTest.cpp:576-607 (same test)
 *   or_kernel_value        — W(0)=1, W(R)=0, W(R/2)=0.25 (MathUtilsTest.cpp), continuity at q=1/2
 *   or_inverse3_cofactor   — A*Ainv = I and numpy.linalg.inv on random matrices
 *   or_fullpivlu_inverse   — numpy.linalg.inv on random matrices; rank-deficient inputs
 *                            give Eigen's zero-filled solution (checked against a
 *                            closed form on a rank-1 matrix).  Bit pattern: parity unpinned.
 */
#include <math.h>
#include <string.h>

#include "oracle.h"

/* include/Materials/Base.h:47-126 (CompleteProperties), E/nu branch = case 7. */
int or_lame_from_E_nu(double E, double nu, double *lambda, double *mu, double *K) {
    if (!(nu >= -1.0 && nu < 0.5) || !(E > 0.0)) return -1;
    *mu = E / (2 * (1 + nu));
    *lambda = (E * nu) / ((1 + nu) * (1 - 2 * nu));
    if (K) *K = E / (3 * (1 - 2 * nu));
    return 0;
}

/* Property completion from any two of {E, nu, G, lambda, K}; codes:
 * 0=E 1=nu 2=G 3=lambda 4=K.  out = [E, nu, G, lambda, K].
 * Follows include/Materials/Base.h:63-126 case by case, then the physical
 * bound checks of :130-158.  Returns 0, or -1 on an invalid pair / bounds. */
int or_elastic_properties(int ga, double a, int gb, double b, double out[5]) {
    double E = NAN, nu = NAN, G = NAN, lam = NAN, K = NAN;
    double *slot[5] = {&E, &nu, &G, &lam, &K};
    if (ga < 0 || ga > 4 || gb < 0 || gb > 4 || ga == gb) return -1;
    *slot[ga] = a;
    *slot[gb] = b;
    if (!isnan(K) && !isnan(E)) {
        nu = (3 * K - E) / (6 * K);
        G = E / (2 * (1 + nu));
        lam = (E * nu) / ((1 + nu) * (1 - 2 * nu));
    } else if (!isnan(K) && !isnan(lam)) {
        G = (3 * (K - lam)) / 2;
        nu = lam / (2 * (lam + G));
        E = 2 * G * (1 + nu);
    } else if (!isnan(K) && !isnan(G)) {
        lam = K - (2.0 / 3.0) * G;
        nu = lam / (2 * (lam + G));
        E = 2 * G * (1 + nu);
    } else if (!isnan(K) && !isnan(nu)) {
        E = 3 * K * (1 - 2 * nu);
        G = E / (2 * (1 + nu));
        lam = (E * nu) / ((1 + nu) * (1 - 2 * nu));
    } else if (!isnan(E) && !isnan(lam)) {
        const double qa = 2 * lam, qb = lam + E, qc = -lam;
        const double disc = qb * qb - 4 * qa * qc;
        nu = (-qb + sqrt(disc)) / (2 * qa);
        G = E / (2 * (1 + nu));
        K = lam + (2.0 / 3.0) * G;
    } else if (!isnan(E) && !isnan(G)) {
        nu = (E / (2 * G)) - 1;
        lam = (E * nu) / ((1 + nu) * (1 - 2 * nu));
        K = lam + (2.0 / 3.0) * G;
    } else if (!isnan(E) && !isnan(nu)) {
        G = E / (2 * (1 + nu));
        lam = (E * nu) / ((1 + nu) * (1 - 2 * nu));
        K = E / (3 * (1 - 2 * nu));
    } else if (!isnan(lam) && !isnan(G)) {
        K = lam + (2.0 / 3.0) * G;
        nu = lam / (2 * (lam + G));
        E = 2 * G * (1 + nu);
    } else if (!isnan(lam) && !isnan(nu)) {
        E = (lam * (1 + nu) * (1 - 2 * nu)) / nu;
        G = E / (2 * (1 + nu));
        K = lam + (2.0 / 3.0) * G;
    } else if (!isnan(G) && !isnan(nu)) {
        E = 2 * G * (1 + nu);
        lam = (2 * G * nu) / (1 - 2 * nu);
        K = (2 * G * (1 + nu)) / (3 * (1 - 2 * nu));
    } else {
        return -1;
    }
    if (nu < -1.0 || nu >= 0.5) return -1;
    if (E <= 0.0 || G <= 0.0 || K <= 0.0) return -1;
    if (lam <= -(2.0 / 3.0) * G) return -1;
    out[0] = E; out[1] = nu; out[2] = G; out[3] = lam; out[4] = K;
    return 0;
}

/* det(I + A) - 1 by the principal invariants, include/MathUtils.h:333-339 (same term order). */
double or_det_ap_im1(const double A[9]) {
#define a(i, j) A[3 * (i) + (j)]
    return a(0, 0) + a(1, 1) + a(2, 2) - a(0, 1) * a(1, 0) * (1 + a(2, 2)) + a(0, 0) * a(1, 1) * (1 + a(2, 2)) -
           a(0, 2) * a(2, 0) * (1 + a(1, 1)) - a(1, 2) * a(2, 1) * (1 + a(0, 0)) + a(0, 0) * a(2, 2) +
           a(1, 1) * a(2, 2) + a(0, 1) * a(1, 2) * a(2, 0) + a(0, 2) * a(1, 0) * a(2, 1);
#undef a
}

/* Cubic B-spline kernel, include/MathUtils.h:321-330. */
double or_kernel_value(const double d[3], double R) {
    const double q = sqrt(fma(d[2], d[2], fma(d[1], d[1], d[0] * d[0]))) / R;
    if (q < 0.5) return 1.0 + 6.0 * q * q * (-1.0 + q);
    if (q < 1.0) return 2.0 * (1.0 + q * (-3.0 + 3.0 * q - 1.0 * q * q));
    return 0.0;
}

/* 3x3 inverse by cofactors times 1/det (the closed form Eigen's inverse() uses for
 * fixed 3x3; include/QuadratureGaussian.h:30).  Row-major. */
void or_inverse3_cofactor(const double A[9], double Ainv[9], double *det_out) {
    const double c00 = A[4] * A[8] - A[5] * A[7];
    const double c10 = A[5] * A[6] - A[3] * A[8];
    const double c20 = A[3] * A[7] - A[4] * A[6];
    const double det = A[0] * c00 + A[1] * c10 + A[2] * c20;
    const double id = 1.0 / det;
    Ainv[0] = c00 * id;
    Ainv[1] = (A[2] * A[7] - A[1] * A[8]) * id;
    Ainv[2] = (A[1] * A[5] - A[2] * A[4]) * id;
    Ainv[3] = c10 * id;
    Ainv[4] = (A[0] * A[8] - A[2] * A[6]) * id;
    Ainv[5] = (A[2] * A[3] - A[0] * A[5]) * id;
    Ainv[6] = c20 * id;
    Ainv[7] = (A[1] * A[6] - A[0] * A[7]) * id;
    Ainv[8] = (A[0] * A[4] - A[1] * A[3]) * id;
    if (det_out) *det_out = det;
}

/*
 * Inverse by LU with complete pivoting, restating the published algorithm of
 * Eigen 3.4.0 FullPivLU (third party; used by aperi::InvertMatrix on non-CUDA
 * builds, include/MathUtils.h:307-319):
 *   - pivot = largest |a_ij| in the trailing block, first maximum in
 *     column-major scan order; stop when it is exactly zero;
 *   - rank = #{i : |u_ii| > |max pivot| * eps * n};
 *   - inverse = solve(I): unit-lower forward solve on all rows, upper solve on
 *     the leading rank x rank block, remaining unknowns set to zero.
 * n <= 8.  Row-major in and out.  Returns the rank.
 */
int or_fullpivlu_inverse(int n, const double *A, double *Ainv) {
    double lu[64];
    int prow[8], pcol[8];
    memcpy(lu, A, sizeof(double) * n * n);
    for (int i = 0; i < n; ++i) prow[i] = pcol[i] = i;
    int nonzero_pivots = n;
    double maxpivot = 0.0;
    for (int k = 0; k < n; ++k) {
        double biggest = -1.0;
        int bi = k, bj = k;
        for (int j = k; j < n; ++j)
            for (int i = k; i < n; ++i) {
                const double v = fabs(lu[i * n + j]);
                if (v > biggest) { biggest = v; bi = i; bj = j; }
            }
        if (biggest == 0.0) { nonzero_pivots = k; break; }
        if (biggest > maxpivot) maxpivot = biggest;
        if (bi != k) {
            for (int j = 0; j < n; ++j) { double t = lu[k * n + j]; lu[k * n + j] = lu[bi * n + j]; lu[bi * n + j] = t; }
            int t = prow[k]; prow[k] = prow[bi]; prow[bi] = t;
        }
        if (bj != k) {
            for (int i = 0; i < n; ++i) { double t = lu[i * n + k]; lu[i * n + k] = lu[i * n + bj]; lu[i * n + bj] = t; }
            int t = pcol[k]; pcol[k] = pcol[bj]; pcol[bj] = t;
        }
        const double piv = lu[k * n + k];
        for (int i = k + 1; i < n; ++i) lu[i * n + k] /= piv;
        for (int i = k + 1; i < n; ++i)
            for (int j = k + 1; j < n; ++j) lu[i * n + j] -= lu[i * n + k] * lu[k * n + j];
    }
    const double thr = maxpivot * (2.220446049250313e-16 * (double)n);
    int rank = 0;
    for (int i = 0; i < nonzero_pivots; ++i) rank += fabs(lu[i * n + i]) > thr;
    if (rank == 0) {
        memset(Ainv, 0, sizeof(double) * n * n);
        return 0;
    }
    for (int col = 0; col < n; ++col) {
        double c[8];
        for (int i = 0; i < n; ++i) c[i] = (prow[i] == col) ? 1.0 : 0.0;
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < i; ++j) c[i] -= lu[i * n + j] * c[j];
        for (int i = rank - 1; i >= 0; --i) {
            for (int j = i + 1; j < rank; ++j) c[i] -= lu[i * n + j] * c[j];
            c[i] /= lu[i * n + i];
        }
        for (int i = 0; i < n; ++i) Ainv[pcol[i] * n + col] = (i < rank) ? c[i] : 0.0;
    }
    return rank;
}
