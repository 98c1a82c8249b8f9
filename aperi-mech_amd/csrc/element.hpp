// Per-element force math shared by the element kernels (fem.hip, tile.hip).
//
//   Tet4 1-point: B from the constant derivatives, w = detJ / 6
//     (include/ShapeFunctionsFunctorTet4.h:27-35, include/QuadratureGaussian.h:17-38,
//      include/ElementTetrahedron4.h:125-127)
//   Hex8 2x2x2: Gauss points in the order of include/ElementHexahedron8.h:107-115, w = detJ;
//     per point H = (u^T dN) J^-1 and f_k -= dN_k (w J^-1 P^T), i.e. -B P^T w without
//     forming B (include/ComputeInternalForceGaussian.h:85-124)
//   Stored H / P are those of the last Gauss point (ComputeInternalForceGaussian.h:89, :99-107).
#pragma once

#include "device_math.hpp"

namespace aphip {

// Central-difference node updates shared by every kernel that finishes a node, so
// a node finished in the tile kernel and one finished in the boundary pass see
// bit-identical arithmetic.
//   predictor  v1/2 = v + dt/2 a, velocity BC, u1 = u + dt v1/2  (ExplicitTimeIntegrator.h:116-126, :348-357)
//   corrector  a = f/m, acceleration BC, v = v1/2 + dt/2 a       (ExplicitTimeIntegrator.h:267-276, :308-318)
APHIP_DEV void cd_predict(const double *u, const double *v, const double *a, unsigned fixed, double dt, double *vh,
                          double *u1) {
#pragma clang fp contract(off)
    const double h = 0.5 * dt;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        double w = __fma_rn(h, a[d], v[d]);
        if (fixed & (1u << d)) w = 0.0;
        vh[d] = w;
        u1[d] = __fma_rn(dt, w, u[d]);
    }
}

APHIP_DEV void cd_correct(const double *f, double m, unsigned fixed, double dt, const double *vh, double *a1,
                          double *v1) {
#pragma clang fp contract(off)
    const double h = 0.5 * dt;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        double acc = f[d] / m;
        if (fixed & (1u << d)) acc = 0.0;
        a1[d] = acc;
        v1[d] = __fma_rn(h, acc, vh[d]);
    }
}

// H, P and the state are written once per step and not read again in it: a
// non-temporal store keeps that stream (144 B per element) from evicting the node
// data that neighbouring tiles re-read from L2.
APHIP_DEV void store_once(double *p, double v) { __builtin_nontemporal_store(v, p); }

// Tet4: fe[k][j] = -(B_k . P_j) w; sn, snp1, aux, r as for stress<KIND>.  UL: X are the
// start-of-step coordinates, U the step's displacement increments, r->Hn the element's
// previous displacement gradient (updated Lagrangian).
template <int KIND, bool UL = false>
APHIP_DEV void tet4_element(const double (&X)[4][3], const double (&U)[4][3], const MatParams &m, double sn,
                            double &snp1, double &aux, double *H, double *P, double (&fe)[4][3],
                            const RateIn *r = nullptr, double sn2 = 0.0) {
#pragma clang fp contract(off)
    double J[9], Ji[9];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        J[3 * i + 0] = X[1][i] - X[0][i];
        J[3 * i + 1] = X[2][i] - X[0][i];
        J[3 * i + 2] = X[3][i] - X[0][i];
    }
    const double det = inverse3(J, Ji);
    const double w = (1.0 / 6.0) * det;
    double B[4][3];
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        B[1][j] = Ji[j];
        B[2][j] = Ji[3 + j];
        B[3][j] = Ji[6 + j];
        B[0][j] = ((-Ji[j]) - Ji[3 + j]) - Ji[6 + j];
    }
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
            H[3 * i + j] = dot4(U[0][i], B[0][j], U[1][i], B[1][j], U[2][i], B[2][j], U[3][i], B[3][j]);
    if constexpr (UL) ul_gradient(H, r->Hn);
    stress<KIND>(H, m, sn, snp1, aux, P, r, sn2);
    double Pf[9];
    if constexpr (UL) {
        ul_force_stress(P, r->Hn, Pf);
    } else {
#pragma unroll
        for (int i = 0; i < 9; ++i) Pf[i] = P[i];
    }
#pragma unroll
    for (int k = 0; k < 4; ++k)
#pragma unroll
        for (int j = 0; j < 3; ++j)
            fe[k][j] = -(dot3(B[k][0], Pf[3 * j], B[k][1], Pf[3 * j + 1], B[k][2], Pf[3 * j + 2]) * w);
}

APHIP_DEV void hex8_sign(int k, int &s0, int &s1, int &s2) {
    s0 = (k == 1 || k == 2 || k == 5 || k == 6) ? 1 : -1;
    s1 = (k == 2 || k == 3 || k == 6 || k == 7) ? 1 : -1;
    s2 = (k >= 4) ? 1 : -1;
}

// Hex8 2x2x2 in modal form.  A trilinear field is a combination of the eight
// corner modes 1, s0, s1, s2, s0 s1, s0 s2, s1 s2, s0 s1 s2 (s_d = the corner's sign
// along parametric axis d); its Walsh-Hadamard transform over the corners gives the
// mode coefficients, and at a Gauss point p = (p0, p1, p2)
//   8 dX/dxi0 = M[s0] + p1 M[s0s1] + p2 M[s0s2] + p1 p2 M[s0s1s2]   (and cyclically),
// which replaces the 8-corner sums of J and grad u by 3 multiply-adds per entry.
// The nodal force f_k = -sum_q dN_k(q) Q_q is accumulated in the same seven modes
// and returned to the corners by one inverse transform.  The 1/8 of the shape
// function derivatives is carried as exact powers of two (J and grad u scaled by 8,
// J^-1 by 1/8, w = det / 512), which leaves every rounding unchanged.
// Gauss points in the order of include/ElementHexahedron8.h:107-115, w = detJ;
// stored H / P are those of the last point (ComputeInternalForceGaussian.h:89, :99-107).

// corner k (hex8_sign order) -> binary index b0 + 2 b1 + 4 b2 with b_d = (s_d + 1) / 2
APHIP_DEV constexpr int hex8_bin(int k) { return k == 2 ? 3 : k == 3 ? 2 : k == 6 ? 7 : k == 7 ? 6 : k; }

// w[S] = sum over corners of prod_{d in S} s_d * v; S a bitmask of axes (w[0] unused)
APHIP_DEV void hex8_modes(const double (&v)[8], double (&w)[8]) {
#pragma clang fp contract(off)
#pragma unroll
    for (int k = 0; k < 8; ++k) w[hex8_bin(k)] = v[k];
#pragma unroll
    for (int h = 1; h < 8; h <<= 1)
#pragma unroll
        for (int i = 0; i < 8; ++i)
            if (!(i & h)) {
                const double a = w[i], b = w[i | h];
                w[i] = b + a;
                w[i | h] = b - a;
            }
}

// v_k = sum_S prod_{d in S} s_d(k) r[S], with r[0] = 0
APHIP_DEV void hex8_corners(const double (&r)[8], double (&v)[8]) {
#pragma clang fp contract(off)
    double w[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) w[i] = r[i];
#pragma unroll
    for (int h = 1; h < 8; h <<= 1)
#pragma unroll
        for (int i = 0; i < 8; ++i)
            if (!(i & h)) {
                const double a = w[i], b = w[i | h];
                w[i] = a - b;
                w[i | h] = a + b;
            }
#pragma unroll
    for (int k = 0; k < 8; ++k) v[k] = w[hex8_bin(k)];
}

// 8 d(field_i)/dxi at Gauss point (p0, p1, p2); p01 = p0 p1 etc.
APHIP_DEV void hex8_grad(const double (&w)[8], double p0, double p1, double p2, double p01, double p02, double p12,
                         double (&g)[3]) {
#pragma clang fp contract(off)
    g[0] = __fma_rn(p12, w[7], __fma_rn(p2, w[5], __fma_rn(p1, w[3], w[1])));
    g[1] = __fma_rn(p02, w[7], __fma_rn(p2, w[6], __fma_rn(p0, w[3], w[2])));
    g[2] = __fma_rn(p01, w[7], __fma_rn(p1, w[6], __fma_rn(p0, w[5], w[4])));
}

// state_n / state_np1 point at the element's state slot(s) (per_qp: 8 slots), each
// state_width<KIND>() values wide.  r: the element's H_n, P_n (those of its last Gauss
// point in the previous step, like the reference's element fields) for every point.
template <int KIND, bool PER_QP, bool UL = false>
APHIP_DEV void hex8_element(const double (&X)[8][3], const double (&U)[8][3], const MatParams &m, double ga,
                            const double *state_n, double *state_np1, double *H, double *P, double (&fe)[8][3],
                            const RateIn *r = nullptr) {
#pragma clang fp contract(off)
    double XM[3][8], UM[3][8];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        double vx[8], vu[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            vx[k] = X[k][i];
            vu[k] = U[k][i];
        }
        hex8_modes(vx, XM[i]);
        hex8_modes(vu, UM[i]);
    }
    // force modes r[j][S] (S = 1..7)
    double R[3][8];
#pragma unroll
    for (int j = 0; j < 3; ++j)
#pragma unroll
        for (int S = 0; S < 8; ++S) R[j][S] = 0.0;
    const double ga2 = ga * ga;
    double sn = 0.0, sn2 = 0.0;
    if constexpr (!PER_QP) {
        sn = load_state<KIND>(state_n, 0);
        sn2 = load_state2<KIND>(state_n, 0);
    }
    // Two Gauss points in flight where the registers allow it without spilling (the
    // neo-Hookean and SVK updates; measured -3 % on the neo-Hookean step); one otherwise.
    // The order of the force accumulation is the same either way.
    constexpr int kQUnroll = (KIND == APHIP_NEO_HOOKEAN || KIND == APHIP_SVK) && !UL ? 2 : 1;
#pragma unroll kQUnroll
    for (int q = 0; q < 8; ++q) {
        int sq0, sq1, sq2;
        hex8_sign(q, sq0, sq1, sq2);
        const double p0 = sq0 * ga, p1 = sq1 * ga, p2 = sq2 * ga;
        const double p01 = (sq0 * sq1) * ga2, p02 = (sq0 * sq2) * ga2, p12 = (sq1 * sq2) * ga2;
        double J[9], G[9];  // 8 J, 8 grad u (row i = component, column = parametric axis)
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            double gx[3], gu[3];
            hex8_grad(XM[i], p0, p1, p2, p01, p02, p12, gx);
            hex8_grad(UM[i], p0, p1, p2, p01, p02, p12, gu);
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                J[3 * i + d] = gx[d];
                G[3 * i + d] = gu[d];
            }
        }
        double Ji[9];  // (8 J)^-1 = J^-1 / 8
        const double w = inverse3(J, Ji) * (-1.0 / 512.0);  // -detJ
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j)
                H[3 * i + j] = dot3(G[3 * i], Ji[j], G[3 * i + 1], Ji[3 + j], G[3 * i + 2], Ji[6 + j]);
        double snp1 = 0.0, aux = 0.0;
        if constexpr (PER_QP) {
            sn = load_state<KIND>(state_n, q);
            sn2 = load_state2<KIND>(state_n, q);
        }
        if constexpr (UL) ul_gradient(H, r->Hn);
        stress<KIND>(H, m, sn, snp1, aux, P, r, sn2);
        if constexpr (PER_QP) store_state<KIND>(state_np1, q, snp1, aux);
        else if (q == 7) store_state<KIND>(state_np1, 0, snp1, aux);
        double Pf[9];
        if constexpr (UL) {
            ul_force_stress(P, r->Hn, Pf);
        } else {
#pragma unroll
            for (int i = 0; i < 9; ++i) Pf[i] = P[i];
        }
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            // Q[n] = -(w / 8) (J^-1 P^T)[n][j]: the force of axis n's derivative
            double Q[3];
#pragma unroll
            for (int n = 0; n < 3; ++n) Q[n] = w * dot3(Ji[3 * n], Pf[3 * j], Ji[3 * n + 1], Pf[3 * j + 1], Ji[3 * n + 2], Pf[3 * j + 2]);
            R[j][1] += Q[0];
            R[j][2] += Q[1];
            R[j][4] += Q[2];
            R[j][3] = __fma_rn(p1, Q[0], __fma_rn(p0, Q[1], R[j][3]));
            R[j][5] = __fma_rn(p2, Q[0], __fma_rn(p0, Q[2], R[j][5]));
            R[j][6] = __fma_rn(p2, Q[1], __fma_rn(p1, Q[2], R[j][6]));
            R[j][7] = __fma_rn(p12, Q[0], __fma_rn(p02, Q[1], __fma_rn(p01, Q[2], R[j][7])));
        }
    }
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        double v[8];
        hex8_corners(R[j], v);
#pragma unroll
        for (int k = 0; k < 8; ++k) fe[k][j] = v[k];
    }
}

// Hex8 with one lane per Gauss point: lane q (= lane id mod 8) evaluates point q
// and the eight lanes of an element combine their nodal contributions with a
// three-step exchange, after which lane k holds node k's force (summed as
// ((q0+q4)+(q2+q6)) + ((q1+q5)+(q3+q7)), a fixed order).  Node coordinates and
// displacements are streamed from LDS (sX/sU, 3 doubles per local node, lc =
// local node of each corner), so no lane keeps the 48-double element state.
// H, P and a per-element J2 state are those of Gauss point 7 (the last point,
// ComputeInternalForceGaussian.h:89, :99-107); with per-point state each lane
// owns its slot.
template <int KIND, bool PER_QP>
APHIP_DEV void hex8_qp_lane(int q, const int (&lc)[8], const double *sX, const double *sU, const MatParams &m,
                            double ga, const double *state_n, double *state_np1, double *H, double *P,
                            double (&fk)[3], const RateIn *r = nullptr) {
#pragma clang fp contract(off)
    int sq0, sq1, sq2;
    hex8_sign(q, sq0, sq1, sq2);
    const double p0 = sq0 * ga, p1 = sq1 * ga, p2 = sq2 * ga;
    // parametric derivatives of node k at this lane's Gauss point
    auto dn = [&](int k, double (&d)[3]) {
        int s0, s1, s2;
        hex8_sign(k, s0, s1, s2);
        const double f0 = s0 < 0 ? (1.0 - p0) : (1.0 + p0);
        const double f1 = s1 < 0 ? (1.0 - p1) : (1.0 + p1);
        const double f2 = s2 < 0 ? (1.0 - p2) : (1.0 + p2);
        d[0] = (s0 * 0.125) * f1 * f2;
        d[1] = (s1 * 0.125) * f0 * f2;
        d[2] = (s2 * 0.125) * f0 * f1;
    };
    double J[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0}, G[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        double d[3];
        dn(k, d);
        const double *x = sX + 3 * lc[k];
        const double *u = sU + 3 * lc[k];
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            const double xi = x[i], ui = u[i];
#pragma unroll
            for (int dd = 0; dd < 3; ++dd) {
                J[3 * i + dd] = __fma_rn(xi, d[dd], J[3 * i + dd]);
                G[3 * i + dd] = __fma_rn(ui, d[dd], G[3 * i + dd]);
            }
        }
    }
    double Ji[9];
    const double w = inverse3(J, Ji);
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
            H[3 * i + j] = dot3(G[3 * i], Ji[j], G[3 * i + 1], Ji[3 + j], G[3 * i + 2], Ji[6 + j]);
    double sn = 0.0, snp1 = 0.0, aux = 0.0;
    sn = load_state<KIND>(state_n, PER_QP ? q : 0);
    const double sn2 = load_state2<KIND>(state_n, PER_QP ? q : 0);
    stress<KIND>(H, m, sn, snp1, aux, P, r, sn2);
    if constexpr (PER_QP) store_state<KIND>(state_np1, q, snp1, aux);
    else if (q == 7) store_state<KIND>(state_np1, 0, snp1, aux);
    double Q[9];
#pragma unroll
    for (int n = 0; n < 3; ++n)
#pragma unroll
        for (int j = 0; j < 3; ++j)
            Q[3 * n + j] = w * dot3(Ji[3 * n], P[3 * j], Ji[3 * n + 1], P[3 * j + 1], Ji[3 * n + 2], P[3 * j + 2]);
    // reduce-scatter over the 8 lanes: at each step keep the half of the nodes whose
    // index bit matches this lane's bit and add the partner's copy of that half
    const bool b2 = (q >> 2) & 1, b1 = (q >> 1) & 1, b0 = q & 1;
    double r4[4][3];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        double dl[3], dh[3];
        dn(k, dl);
        dn(k + 4, dh);
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            const double cl = -dot3(dl[0], Q[j], dl[1], Q[3 + j], dl[2], Q[6 + j]);
            const double ch = -dot3(dh[0], Q[j], dh[1], Q[3 + j], dh[2], Q[6 + j]);
            const double keep = b2 ? ch : cl;
            const double send = b2 ? cl : ch;
            r4[k][j] = keep + __shfl_xor(send, 4, 8);
        }
    }
    double r2[2][3];
#pragma unroll
    for (int k = 0; k < 2; ++k)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            const double keep = b1 ? r4[k + 2][j] : r4[k][j];
            const double send = b1 ? r4[k][j] : r4[k + 2][j];
            r2[k][j] = keep + __shfl_xor(send, 2, 8);
        }
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        const double keep = b0 ? r2[1][j] : r2[0][j];
        const double send = b0 ? r2[0][j] : r2[1][j];
        fk[j] = keep + __shfl_xor(send, 1, 8);
    }
}

}  // namespace aphip
