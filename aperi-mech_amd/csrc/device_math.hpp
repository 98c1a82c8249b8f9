// Device-side FP64 math of the force step: 3x3 algebra and the stress updates.
//
// Stress updates restate the reference StressFunctor::GetStress implementations:
//   linear elastic  include/Materials/LinearElastic.h:70-88
//   neo-Hookean     include/Materials/NeoHooke.h:67-92 (J-1 from DetApIm1, MathUtils.h:333-339;
//                   F^-1 by cofactors, as the reference's CUDA build does, MathUtils.h:314-316)
//   J2              include/Materials/Plastic.h:88-140
//   SVK             include/Materials/Elastic.h:61-79
//   NH with damage  include/Materials/NeoHookeWithDamage.h:180-243
//   power-law creep include/Materials/PowerLawCreep.h:101-173 (hypoelastic, rate form)
//   Drucker-Prager  include/Materials/DruckerPrager.h:97-216 (rate form, oblique return)
// The material is a compile-time template parameter; the reference's device
// virtual call (ComputeInternalForceGaussian.h:107) is resolved on the host.
#pragma once

#include <hip/hip_runtime.h>

#include "aperi_hip.h"

namespace aphip {

#define APHIP_DEV __device__ __forceinline__

// Rounding is fixed by the source: every device function of the force math turns
// implicit multiply-add contraction off and writes its fused multiply-adds
// explicitly, so an element gives bit-identical results in every kernel that
// inlines it (the compiler's fusion choices otherwise depend on the surrounding
// code).

// a0 b0 + a1 b1 + a2 b2, accumulated left to right with fused multiply-adds
APHIP_DEV double dot3(double a0, double b0, double a1, double b1, double a2, double b2) {
#pragma clang fp contract(off)
    return __fma_rn(a2, b2, __fma_rn(a1, b1, a0 * b0));
}

APHIP_DEV double dot4(double a0, double b0, double a1, double b1, double a2, double b2, double a3, double b3) {
#pragma clang fp contract(off)
    return __fma_rn(a3, b3, __fma_rn(a2, b2, __fma_rn(a1, b1, a0 * b0)));
}

// det(I + A) - 1 by principal invariants (MathUtils.h:333-339); A row-major.  The
// terms are accumulated in the reference's order.
APHIP_DEV double det_ap_im1(const double *A) {
#pragma clang fp contract(off)
    double s = A[0] + A[4] + A[8];
    s = __fma_rn(-(A[1] * A[3]), 1.0 + A[8], s);
    s = __fma_rn(A[0] * A[4], 1.0 + A[8], s);
    s = __fma_rn(-(A[2] * A[6]), 1.0 + A[4], s);
    s = __fma_rn(-(A[5] * A[7]), 1.0 + A[0], s);
    s = __fma_rn(A[0], A[8], s);
    s = __fma_rn(A[4], A[8], s);
    s = __fma_rn(A[1] * A[5], A[6], s);
    s = __fma_rn(A[2] * A[3], A[7], s);
    return s;
}

// a b - c d with one rounding of the product difference's second term
APHIP_DEV double dif2(double a, double b, double c, double d) {
#pragma clang fp contract(off)
    return __fma_rn(a, b, -(c * d));
}

// Cofactor inverse of a row-major 3x3; returns det.
APHIP_DEV double inverse3(const double *A, double *Ai) {
#pragma clang fp contract(off)
    const double c00 = dif2(A[4], A[8], A[5], A[7]);
    const double c10 = dif2(A[5], A[6], A[3], A[8]);
    const double c20 = dif2(A[3], A[7], A[4], A[6]);
    const double det = dot3(A[0], c00, A[1], c10, A[2], c20);
    const double id = 1.0 / det;
    Ai[0] = c00 * id;
    Ai[1] = dif2(A[2], A[7], A[1], A[8]) * id;
    Ai[2] = dif2(A[1], A[5], A[2], A[4]) * id;
    Ai[3] = c10 * id;
    Ai[4] = dif2(A[0], A[8], A[2], A[6]) * id;
    Ai[5] = dif2(A[2], A[3], A[0], A[5]) * id;
    Ai[6] = c20 * id;
    Ai[7] = dif2(A[1], A[6], A[0], A[7]) * id;
    Ai[8] = dif2(A[0], A[4], A[1], A[3]) * id;
    return det;
}

struct MatParams {
    double lambda, mu, yield_stress, hardening;
    // neo-Hookean with damage; max_damage and max_damage_increment lie in [0, 1]
    double I1_critical, I1_failure, alpha, max_damage, max_damage_increment;
    // power-law creep: bulk modulus and the creep rate factor at the constant
    // temperature, a = A exp(-m / T) (A exp(-m) at T = 0), and the exponent n
    double bulk, creep_rate, creep_n;
    // Drucker-Prager: yield s / sqrt(2) = A1 - A2 z, flow direction A2G
    double dp_A1, dp_A2, dp_A2G;
};

// Rate-form materials read the previous step's displacement gradient and stress.
template <int KIND>
constexpr bool rate_form() {
    return KIND == APHIP_POWER_LAW_CREEP || KIND == APHIP_DRUCKER_PRAGER;
}

// Inputs of a rate-form update: H_n, P_n (row-major) and the time increment.
struct RateIn {
    double Hn[9], Pn[9];
    double dt;
};

// H_n, P_n of element e from the SoA (9 x ne) N buffers: both for the rate-form
// materials, H_n alone for the updated Lagrangian formulation (UL), nothing otherwise.
template <int KIND, bool UL = false>
APHIP_DEV void load_rate(RateIn &r, const double *Hn, const double *Pn, int64_t ne, int64_t e, double dt) {
    if constexpr (rate_form<KIND>() || UL) {
#pragma unroll
        for (int i = 0; i < 9; ++i) r.Hn[i] = Hn[i * ne + e];
    }
    if constexpr (rate_form<KIND>()) {
#pragma unroll
        for (int i = 0; i < 9; ++i) r.Pn[i] = Pn[i * ne + e];
        r.dt = dt;
    }
}

// State values per evaluation point (element or Gauss point): J2 the equivalent
// plastic strain; neo-Hookean with damage the damage and I1 (in that order).
template <int KIND>
constexpr int state_width() {
    return KIND == APHIP_J2 ? 1
           : (KIND == APHIP_NEO_HOOKEAN_DAMAGE || KIND == APHIP_POWER_LAW_CREEP || KIND == APHIP_DRUCKER_PRAGER)
               ? 2
               : 0;
}

// The second state value of evaluation point `slot`, for the material that reads it
// (Drucker-Prager's volumetric plastic strain).
template <int KIND>
APHIP_DEV double load_state2(const double *state, int64_t slot) {
    if constexpr (KIND == APHIP_DRUCKER_PRAGER) return state[2 * slot + 1];
    else return 0.0;
}

// The first state value of evaluation point `slot` (the only one a material reads).
template <int KIND>
APHIP_DEV double load_state(const double *state, int64_t slot) {
    if constexpr (state_width<KIND>() == 0) return 0.0;
    else return state[state_width<KIND>() * slot];
}

// New state of evaluation point `slot`: snp1 and, for the damage model, aux (I1).
template <int KIND>
APHIP_DEV void store_state(double *state, int64_t slot, double snp1, double aux) {
    if constexpr (state_width<KIND>() == 1) {
        state[slot] = snp1;
    } else if constexpr (state_width<KIND>() == 2) {
        state[2 * slot] = snp1;
        state[2 * slot + 1] = aux;
    }
}

APHIP_DEV double tr3(const double *A) { return A[0] + A[4] + A[8]; }

// sqrt(A : A), summed in row-major order
APHIP_DEV double magnitude9(const double *A) {
#pragma clang fp contract(off)
    double s = 0.0;
#pragma unroll
    for (int i = 0; i < 9; ++i) s = __fma_rn(A[i], A[i], s);
    return sqrt(s);
}

// Power-law creep (PowerLawCreep.h:101-173).  Cauchy stress from the old PK1 stress,
// sigma = P_n F^T / J; rate of deformation from L = (H - H_n) / dt (I + H_n)^-1
// (ComputeInternalForceBase.h:182-187; L = 0 when dt = 0); one explicit sub-step of
// the deviator over min(critt, dt) (the reference's loop leaves after its first pass:
// tm1 = min(critt, dt) <= dt), pressure from the bulk modulus; P = J sigma F^-T.
// State: effective creep strain (accumulated), effective stress rate.
APHIP_DEV void creep_stress(const double *H, const MatParams &m, const RateIn &r, double sn, double &snp1,
                            double &rate, double *P) {
#pragma clang fp contract(off)
    double F[9], Fi[9], Fn[9], Fni[9], L[9];
#pragma unroll
    for (int i = 0; i < 9; ++i) {
        F[i] = H[i];
        Fn[i] = r.Hn[i];
    }
    F[0] += 1.0;
    F[4] += 1.0;
    F[8] += 1.0;
    Fn[0] += 1.0;
    Fn[4] += 1.0;
    Fn[8] += 1.0;
    const double J = inverse3(F, Fi);
    inverse3(Fn, Fni);
    const double dt = r.dt;
    auto per_dt = [dt](double x) { return dt != 0.0 ? x / dt : 0.0; };
    double dH[9];
#pragma unroll
    for (int i = 0; i < 9; ++i) dH[i] = per_dt(H[i] - r.Hn[i]);
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) L[3 * i + j] = dot3(dH[3 * i], Fni[j], dH[3 * i + 1], Fni[3 + j], dH[3 * i + 2], Fni[6 + j]);
    // sigma = (1/J) P_n F^T
    const double iJ = 1.0 / J;
    double sig[9];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
            sig[3 * i + j] = dot3(iJ * r.Pn[3 * i], F[3 * j], iJ * r.Pn[3 * i + 1], F[3 * j + 1], iJ * r.Pn[3 * i + 2],
                                  F[3 * j + 2]);
    const double G = m.mu;
    double sd[9], de[9], ded[9], cs[9];
    const double sm = tr3(sig) / 3.0;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) de[3 * i + j] = (0.5 * (L[3 * i + j] + L[3 * j + i])) * dt;
    const double dm = tr3(de) / 3.0;
#pragma unroll
    for (int i = 0; i < 9; ++i) {
        sd[i] = sig[i];
        ded[i] = de[i];
    }
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        sd[4 * i] -= sm;
        ded[4 * i] -= dm;
    }
#pragma unroll
    for (int i = 0; i < 9; ++i) cs[i] = ded[i] + sd[i] / 2.0 / G;
    const double s32 = sqrt(3.0 / 2.0);
    double eff = magnitude9(sd) * s32;
    double edot = m.creep_rate * pow(eff, m.creep_n - 1.0);
    double critt = edot == 0.0 ? 1.1 * dt : 4.0 * 0.98 / (6.0 * G * m.creep_n * edot);
    const double delt = fmin(critt, dt);
#pragma unroll
    for (int i = 0; i < 9; ++i) {
        const double dp = 1.5 * sd[i] * edot;
        const double sdot = 2.0 * G * (per_dt(ded[i]) - dp);
        sd[i] = __fma_rn(sdot, delt, sd[i]);
    }
    edot = m.creep_rate * pow(magnitude9(sd) * s32, m.creep_n - 1.0);
    const double mean = __fma_rn(m.bulk, tr3(de), tr3(sig) / 3.0);
#pragma unroll
    for (int i = 0; i < 9; ++i) {
        sig[i] = sd[i];
        cs[i] -= sd[i] / 2.0 / G;
    }
#pragma unroll
    for (int i = 0; i < 3; ++i) sig[4 * i] += mean;
    snp1 = __fma_rn(sqrt(2.0 / 3.0), magnitude9(cs), sn);
    double ds[9];
#pragma unroll
    for (int i = 0; i < 9; ++i) ds[i] = 2.0 * G * (per_dt(ded[i]) - 3.0 * edot * sd[i] / 2.0);
    rate = s32 * magnitude9(ds);
    // P = (J sigma) F^-T
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
            P[3 * i + j] = dot3(J * sig[3 * i], Fi[3 * j], J * sig[3 * i + 1], Fi[3 * j + 1], J * sig[3 * i + 2],
                                Fi[3 * j + 2]);
}

// Lode components of A: z = tr A / sqrt 3, s = |dev A|, E = dev A / s (0 when s <= 1e-10)
// (include/Tensor.h:34-57)
APHIP_DEV void lode(const double *A, double &z, double &s, double *E) {
#pragma clang fp contract(off)
    const double t = tr3(A);
    z = t / sqrt(3.0);
    double d[9];
#pragma unroll
    for (int i = 0; i < 9; ++i) d[i] = A[i];
    const double iso = t / 3.0;
#pragma unroll
    for (int i = 0; i < 3; ++i) d[4 * i] -= iso;
    s = magnitude9(d);
    const bool nz = fabs(s) > 1.0e-10;
#pragma unroll
    for (int i = 0; i < 9; ++i) E[i] = nz ? d[i] / s : 0.0;
}

// L = (H - H_n) / dt (I + H_n)^-1 (ComputeInternalForceBase.h:182-187); 0 when dt = 0
APHIP_DEV void velocity_gradient(const double *H, const RateIn &r, double *L) {
#pragma clang fp contract(off)
    double Fn[9], Fni[9], dH[9];
#pragma unroll
    for (int i = 0; i < 9; ++i) Fn[i] = r.Hn[i];
    Fn[0] += 1.0;
    Fn[4] += 1.0;
    Fn[8] += 1.0;
    inverse3(Fn, Fni);
#pragma unroll
    for (int i = 0; i < 9; ++i) dH[i] = r.dt != 0.0 ? (H[i] - r.Hn[i]) / r.dt : 0.0;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) L[3 * i + j] = dot3(dH[3 * i], Fni[j], dH[3 * i + 1], Fni[3 + j], dH[3 * i + 2], Fni[6 + j]);
}

// Drucker-Prager (DruckerPrager.h:97-216): Kirchhoff stress tau = P_n F^T plus the
// hypoelastic increment of de = dt/2 (L + L^T); beyond g = s / sqrt 2 - (A1 - A2 z) >
// 1e-4 an oblique return in the Lode plane (Newton, at most 25 iterations; the deviatoric
// direction E_s of the trial stress is kept), plastic strain increment from the
// elastic part of dstrain = dt/2 (H + H^T) (the reference's choice); P = tau F^-T.
// State: plastic strain, volumetric plastic strain (both accumulated).  Where the
// reference aborts on a failed return, P and the state become NaN.
APHIP_DEV void dp_stress(const double *H, const MatParams &m, const RateIn &r, double sn, double sn2, double &snp1,
                         double &snp1b, double *P) {
#pragma clang fp contract(off)
    const double K = m.bulk, G = m.mu, dt = r.dt;
    const double ez = 1.0 / sqrt(3.0);  // Ez = I / sqrt 3
    double F[9], Fi[9], L[9], tau[9], de[9];
#pragma unroll
    for (int i = 0; i < 9; ++i) F[i] = H[i];
    F[0] += 1.0;
    F[4] += 1.0;
    F[8] += 1.0;
    inverse3(F, Fi);
    velocity_gradient(H, r, L);
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            tau[3 * i + j] = dot3(r.Pn[3 * i], F[3 * j], r.Pn[3 * i + 1], F[3 * j + 1], r.Pn[3 * i + 2], F[3 * j + 2]);
            de[3 * i + j] = (0.5 * dt) * (L[3 * i + j] + L[3 * j + i]);
        }
    const double trde = tr3(de);
    {
        const double iso = trde / 3.0, kt = K * trde, g2 = 2 * G;
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                const double dd = i == j ? de[3 * i + j] - iso : de[3 * i + j];
                tau[3 * i + j] += (i == j ? kt : 0.0) + g2 * dd;
            }
    }
    double zt, st, Es[9];
    lode(tau, zt, st, Es);
    double dstr[9], dep[9];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            dstr[3 * i + j] = (0.5 * (H[3 * i + j] + H[3 * j + i])) * dt;
            dep[3 * i + j] = 0.0;
        }
    const double rs2 = sqrt(2.0);
    double f = m.dp_A1 - m.dp_A2 * zt;
    double g = st / rs2 - f;
    bool failed = false;
    if (g > 1.0e-4) {
        bool converged = false;
        const double Nz = m.dp_A2, Ns = 1.0 / rs2, Mz = m.dp_A2G, Ms = Ns;
        const double Az = K * Mz, As = 2.0 * G * Ms;
        for (int it = 0; it < 25; ++it) {
            const double fac = sqrt((zt * zt + st * st) / (Mz * Mz + Ms * Ms)) / K;
            const double Pz = fac * Az, Ps = fac * As;
            const double beta = -g / (Nz * Pz + Ns * Ps);
            if (fabs(beta) < 1.0e-10) {
                converged = true;
                break;
            }
            zt += beta * Pz;
            st += beta * Ps;
            f = m.dp_A1 - m.dp_A2 * zt;
            g = st / rs2 - f;
        }
        failed = !converged;
        double dtau[9];
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                const int k = 3 * i + j;
                dtau[k] = zt * (i == j ? ez : 0.0) + st * Es[k] - tau[k];
                tau[k] += dtau[k];
            }
        double dz, ds, Ed[9];
        lode(dtau, dz, ds, Ed);
        const double cz = 1.0 / 3.0 / K * dz, cs = 1.0 / 2.0 / G * ds;
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                const int k = 3 * i + j;
                const double dee = cz * (i == j ? ez : 0.0) + cs * Ed[k];
                dep[k] = dstr[k] - dee;
            }
    }
    const double tp = tr3(dep);
    double dd[9];
#pragma unroll
    for (int i = 0; i < 9; ++i) dd[i] = dep[i];
#pragma unroll
    for (int i = 0; i < 3; ++i) dd[4 * i] -= tp / 3.0;
    double ddot = 0.0;
#pragma unroll
    for (int i = 0; i < 9; ++i) ddot = __fma_rn(dd[i], dd[i], ddot);
    const double deqps = sqrt(ddot + (tp * tp) / 3.0);
    const double nan = __builtin_nan("");
    snp1 = failed ? nan : sn + deqps;
    snp1b = failed ? nan : sn2 + tp;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            const double v = dot3(tau[3 * i], Fi[3 * j], tau[3 * i + 1], Fi[3 * j + 1], tau[3 * i + 2], Fi[3 * j + 2]);
            P[3 * i + j] = failed ? nan : v;
        }
}

// P = stress(H).  sn / snp1: J2 equivalent plastic strain, damage, effective creep
// strain or the Drucker-Prager plastic strain, in / out; aux: the damage model's I1,
// the creep model's effective stress rate or the Drucker-Prager volumetric plastic
// strain (whose previous value is sn2).  The smoothed-cell path turns a new failure's
// marker 2.0 into the maximum damage afterwards (cell_damage_marker).  r: the
// rate-form inputs (creep, Drucker-Prager).
template <int KIND>
APHIP_DEV void stress(const double *H, const MatParams &m, double sn, double &snp1, double &aux, double *P,
                      const RateIn *r = nullptr, double sn2 = 0.0) {
#pragma clang fp contract(off)
    if constexpr (KIND == APHIP_LINEAR_ELASTIC) {
        const double two_mu = m.mu * 2.0;
        double eps[9];
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) eps[3 * i + j] = 0.5 * (H[3 * i + j] + H[3 * j + i]);
        const double tr = eps[0] + eps[4] + eps[8];
#pragma unroll
        for (int i = 0; i < 9; ++i) P[i] = two_mu * eps[i];
#pragma unroll
        for (int i = 0; i < 3; ++i) P[4 * i] = __fma_rn(m.lambda, tr, two_mu * eps[4 * i]);
        snp1 = sn;
    } else if constexpr (KIND == APHIP_NEO_HOOKEAN) {
        double tau[9], F[9], Fi[9];
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                const double hh = dot3(H[3 * i], H[3 * j], H[3 * i + 1], H[3 * j + 1], H[3 * i + 2], H[3 * j + 2]);
                tau[3 * i + j] = m.mu * (hh + H[3 * j + i] + H[3 * i + j]);
            }
        const double lnJ = log1p(det_ap_im1(H));
#pragma unroll
        for (int i = 0; i < 3; ++i) tau[4 * i] = __fma_rn(m.lambda, lnJ, tau[4 * i]);
#pragma unroll
        for (int i = 0; i < 9; ++i) F[i] = H[i];
        F[0] += 1.0;
        F[4] += 1.0;
        F[8] += 1.0;
        inverse3(F, Fi);
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j)
                P[3 * i + j] = dot3(tau[3 * i], Fi[3 * j], tau[3 * i + 1], Fi[3 * j + 1], tau[3 * i + 2], Fi[3 * j + 2]);
        snp1 = sn;
    } else if constexpr (KIND == APHIP_J2) {
        double ep = sn;
        double dev[9];
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) dev[3 * i + j] = 0.5 * (H[3 * i + j] + H[3 * j + i]);
        const double tr = dev[0] + dev[4] + dev[8];
        const double K = m.lambda + 2.0 * m.mu / 3.0;  // (no product to fuse: 2 mu / 3 rounds first)
        const double p = K * tr;
#pragma unroll
        for (int i = 0; i < 3; ++i) dev[4 * i] -= (1.0 / 3.0) * tr;
        double n2 = 0.0;
#pragma unroll
        for (int i = 0; i < 9; ++i) {
            dev[i] *= 2.0 * m.mu;
            n2 = __fma_rn(dev[i], dev[i], n2);
        }
        const double eta_norm = sqrt(n2);
        const double q = sqrt(3.0 / 2.0) * eta_norm;
        const double phi = q - __fma_rn(m.hardening, ep, m.yield_stress);
        if (phi > 0.0) {
            const double dep = phi / (3.0 * m.mu + m.hardening);
            const double s = sqrt(6.0) * m.mu * dep / eta_norm;
#pragma unroll
            for (int i = 0; i < 9; ++i) dev[i] = __fma_rn(-s, dev[i], dev[i]);
            ep += dep;
        }
        snp1 = ep;
#pragma unroll
        for (int i = 0; i < 9; ++i) P[i] = dev[i];
#pragma unroll
        for (int i = 0; i < 3; ++i) P[4 * i] += p;
    } else if constexpr (KIND == APHIP_NEO_HOOKEAN_DAMAGE) {
        // tau = lambda ln J I + mu (B - I), P = tau F^-T as the neo-Hookean branch, then
        // damage from I1 = tr B (NeoHookeWithDamage.h:187-243)
        double BmI[9], tau[9], F[9], Fi[9];
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                const double hh = dot3(H[3 * i], H[3 * j], H[3 * i + 1], H[3 * j + 1], H[3 * i + 2], H[3 * j + 2]);
                BmI[3 * i + j] = hh + H[3 * j + i] + H[3 * i + j];
                tau[3 * i + j] = m.mu * BmI[3 * i + j];
            }
        const double I1 = BmI[0] + BmI[4] + BmI[8] + 3.0;
        const double lnJ = log1p(det_ap_im1(H));
#pragma unroll
        for (int i = 0; i < 3; ++i) tau[4 * i] = __fma_rn(m.lambda, lnJ, tau[4 * i]);
#pragma unroll
        for (int i = 0; i < 9; ++i) F[i] = H[i];
        F[0] += 1.0;
        F[4] += 1.0;
        F[8] += 1.0;
        inverse3(F, Fi);
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j)
                P[3 * i + j] = dot3(tau[3 * i], Fi[3 * j], tau[3 * i + 1], Fi[3 * j + 1], tau[3 * i + 2], Fi[3 * j + 2]);
        const double Dn = fmin(fmax(sn, 0.0), m.max_damage);
        double D = Dn;
        if (I1 > m.I1_failure) {
            D = 1.0;
        } else if (I1 > m.I1_critical) {
            D = pow((I1 - m.I1_critical) / (m.I1_failure - m.I1_critical), m.alpha);
            D = fmin(fmax(D, 0.0), 1.0);
            D = fmax(D, Dn);  // no healing
        }
        if (D != 0.0) {
            D = fmin(D, Dn + m.max_damage_increment);
            const bool failed = D >= m.max_damage;
            if (failed) D = m.max_damage;
            const double keep = 1.0 - D;
#pragma unroll
            for (int i = 0; i < 9; ++i) P[i] = P[i] * keep;
            if (failed && Dn < m.max_damage) D = 2.0;  // marks a new failure
        }
        snp1 = D;
        aux = I1;
    } else if constexpr (KIND == APHIP_POWER_LAW_CREEP) {
        creep_stress(H, m, *r, sn, snp1, aux, P);
    } else if constexpr (KIND == APHIP_DRUCKER_PRAGER) {
        dp_stress(H, m, *r, sn, sn2, snp1, aux, P);
    } else {  // SVK
        const double two_mu = m.mu * 2.0;
        double E[9], S[9];
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                const double hth = dot3(H[i], H[j], H[3 + i], H[3 + j], H[6 + i], H[6 + j]);
                E[3 * i + j] = 0.5 * (H[3 * i + j] + H[3 * j + i] + hth);
            }
        const double tr = E[0] + E[4] + E[8];
#pragma unroll
        for (int i = 0; i < 9; ++i) S[i] = two_mu * E[i];
#pragma unroll
        for (int i = 0; i < 3; ++i) S[4 * i] = __fma_rn(m.lambda, tr, S[4 * i]);
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j)
                P[3 * i + j] = dot3(S[j], (i == 0 ? 1.0 : 0.0) + H[3 * i], S[3 + j], (i == 1 ? 1.0 : 0.0) + H[3 * i + 1],
                                    S[6 + j], (i == 2 ? 1.0 : 0.0) + H[3 * i + 2]);
        snp1 = sn;
    }
}

// Updated Lagrangian (ComputeInternalForceBase.h:158-166): with dH the gradient of the
// step's displacement increment over the start-of-step configuration, the total
// displacement gradient is H = dH + H_n + dH H_n.
APHIP_DEV void ul_gradient(double *H, const double *Hn) {
#pragma clang fp contract(off)
    double dH[9];
#pragma unroll
    for (int i = 0; i < 9; ++i) dH[i] = H[i];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
            H[3 * i + j] = (dH[3 * i + j] + Hn[3 * i + j]) +
                           dot3(dH[3 * i], Hn[j], dH[3 * i + 1], Hn[3 + j], dH[3 * i + 2], Hn[6 + j]);
}

// The reference integrates the force with B F_r and w / det F_r, F_r = I + H_n
// (ComputeInternalForceGaussian.h:109-123).  Since (B F_r) P^T = B (P F_r^T)^T, the
// same force comes from the start-of-step B and w with Pt = P F_r^T / det F_r.
APHIP_DEV void ul_force_stress(const double *P, const double *Hn, double *Pt) {
#pragma clang fp contract(off)
    double Fr[9], Fri[9];
#pragma unroll
    for (int i = 0; i < 9; ++i) Fr[i] = Hn[i];
    Fr[0] += 1.0;
    Fr[4] += 1.0;
    Fr[8] += 1.0;
    const double idet = 1.0 / inverse3(Fr, Fri);
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
            Pt[3 * i + j] = dot3(P[3 * i], Fr[3 * j], P[3 * i + 1], Fr[3 * j + 1], P[3 * i + 2], Fr[3 * j + 2]) * idet;
}

// The smoothed-cell path's handling of a new failure (ComputeInternalForceSmoothedCell.h:
// 219-223 with NeoHookeWithDamage.h:247-252): the marker becomes the maximum damage.
template <int KIND>
APHIP_DEV double cell_damage_marker(double snp1, const MatParams &m) {
    if constexpr (KIND == APHIP_NEO_HOOKEAN_DAMAGE) return snp1 == 2.0 ? m.max_damage : snp1;
    else return snp1;
}

}  // namespace aphip
