/*
 * or_material.c — stress updates.  TEST INFRASTRUCTURE (oracle).
 *
 * Each branch restates one reference StressFunctor::GetStress:
 *   linear elastic  include/Materials/LinearElastic.h:70-88
 *   neo-Hookean     include/Materials/NeoHooke.h:67-92 (InvertMatrix = FullPivLU, MathUtils.h:307-319)
 *   J2 plasticity   include/Materials/Plastic.h:88-140 (small strain, Simo-Hughes box 7.5)
 *   SVK "elastic"   include/Materials/Elastic.h:61-79
 *   power-law creep include/Materials/PowerLawCreep.h:101-173 with the velocity gradient of
 *                   include/ComputeInternalForceBase.h:182-187 (L = (H - H_n) / dt (I + H_n)^-1);
 *                   the sub-stepping loop of :122-144 leaves after its first pass (tm1 =
 *                   min(critt, dt) <= dt), so one explicit sub-step of min(critt, dt)
 *   Drucker-Prager  include/Materials/DruckerPrager.h:97-216 (Kirchhoff stress P_n F^T plus the
 *                   hypoelastic increment, oblique return in the Lode plane by Newton with the
 *                   trial deviatoric direction, Lode components of include/Tensor.h:34-57; the
 *                   reference aborts on 25 failed iterations, the oracle returns NaN)
 *   NH with damage  include/Materials/NeoHookeWithDamage.h:180-243 (state: damage, I1; damage
 *                   from I1 between I1_critical and I1_failure with exponent alpha, never
 *                   decreasing, at most maximum_damage_increment per evaluation, capped at
 *                   maximum_damage, where a new failure is marked 2.0; the constructor's
 *                   clamping of the increment to [0, 1], :153-168)
 *
 * Pinning (tests/test_oracle_materials.py):
 *   LE   textbook Hooke's law on random symmetric strains; P independent of skew(H)
 *   NH   notebook closed forms (test/material_tests/neohookean.ipynb): uniaxial
 *        stretch and simple shear Cauchy stress tau = lambda ln J I + mu (B - I);
 *        rotation invariance P(RF) = R P(F); small-strain limit -> LE
 *   J2   1-element uniaxial radial return with hardening, analytic plastic strain
 *   SVK  PatchTestFixture.h:138-169 closed form P = F S with E = 1, nu = 0
 *   NH with damage (no reference test exists): below I1_critical identical to NH with
 *        state (0, I1); uniaxial stretch F = diag(l, 1, 1): I1 = l^2 + 2 and
 *        P = (1 - ((I1 - Ic) / (If - Ic))^alpha) P_NH; the increment cap, no decrease on
 *        unloading, the failure marker then the clamped damage on the next evaluation
 */
#include <math.h>

#include "oracle.h"

void or_stress(int kind, double lambda, double mu, double yield_stress, double hardening,
               const double H[9], const double *state_n, double *state_np1, double P[9]) {
    if (kind == OR_LINEAR_ELASTIC) {
        const double two_mu = mu * 2.0;
        double eps[9];
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) eps[3 * i + j] = 0.5 * (H[3 * i + j] + H[3 * j + i]);
        const double tr = eps[0] + eps[4] + eps[8];
        for (int i = 0; i < 9; ++i) P[i] = two_mu * eps[i];
        for (int i = 0; i < 3; ++i) P[4 * i] = lambda * tr + two_mu * eps[4 * i];
    } else if (kind == OR_NEO_HOOKEAN) {
        const double m = (mu * 2.0) / 2.0;
        double BmI[9], F[9], Finv[9], tau[9];
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) {
                double hh = 0.0;
                for (int k = 0; k < 3; ++k) hh += H[3 * i + k] * H[3 * j + k];
                BmI[3 * i + j] = hh + H[3 * j + i] + H[3 * i + j];
            }
        const double Jm1 = or_det_ap_im1(H);
        const double lnJ = log1p(Jm1);
        for (int i = 0; i < 9; ++i) tau[i] = m * BmI[i];
        for (int i = 0; i < 3; ++i) tau[4 * i] += lambda * lnJ;
        for (int i = 0; i < 9; ++i) F[i] = H[i];
        F[0] += 1.0; F[4] += 1.0; F[8] += 1.0;
        or_fullpivlu_inverse(3, F, Finv);
        /* P = tau * Finv^T  ->  P_ij = sum_k tau_ik Finv_jk */
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) {
                double s = 0.0;
                for (int k = 0; k < 3; ++k) s += tau[3 * i + k] * Finv[3 * j + k];
                P[3 * i + j] = s;
            }
    } else if (kind == OR_J2) {
        const double m = (mu * 2.0) / 2.0;
        double ep = state_n ? state_n[0] : 0.0;
        double eps[9], dev[9];
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) eps[3 * i + j] = 0.5 * (H[3 * i + j] + H[3 * j + i]);
        const double tr = eps[0] + eps[4] + eps[8];
        const double K = lambda + 2.0 * m / 3.0;
        const double p = K * tr;
        for (int i = 0; i < 9; ++i) dev[i] = eps[i];
        for (int i = 0; i < 3; ++i) dev[4 * i] -= (1.0 / 3.0) * tr;
        for (int i = 0; i < 9; ++i) dev[i] *= 2.0 * m;
        double nrm2 = 0.0;
        for (int i = 0; i < 9; ++i) nrm2 += dev[i] * dev[i];
        const double eta_norm = sqrt(nrm2);
        const double q = sqrt(3.0 / 2.0) * eta_norm;
        const double phi = q - (yield_stress + hardening * ep);
        if (phi > 0.0) {
            const double dep = phi / (3.0 * m + hardening);
            const double s = sqrt(6.0) * m * dep / eta_norm;
            for (int i = 0; i < 9; ++i) dev[i] -= s * dev[i];
            ep += dep;
        }
        if (state_np1) state_np1[0] = ep;
        for (int i = 0; i < 9; ++i) P[i] = dev[i];
        for (int i = 0; i < 3; ++i) P[4 * i] += p;
    } else { /* OR_SVK */
        const double two_mu = mu * 2.0;
        double E[9], S[9];
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) {
                double hth = 0.0;
                for (int k = 0; k < 3; ++k) hth += H[3 * k + i] * H[3 * k + j];
                E[3 * i + j] = 0.5 * (H[3 * i + j] + H[3 * j + i] + hth);
            }
        const double tr = E[0] + E[4] + E[8];
        for (int i = 0; i < 9; ++i) S[i] = two_mu * E[i];
        for (int i = 0; i < 3; ++i) S[4 * i] += lambda * tr;
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) {
                double s = 0.0;
                for (int k = 0; k < 3; ++k) s += ((i == k ? 1.0 : 0.0) + H[3 * i + k]) * S[3 * k + j];
                P[3 * i + j] = s;
            }
    }
}

int or_state_width(int kind) {
    return kind == OR_J2 ? 1 : (kind == OR_NH_DAMAGE || kind == OR_CREEP || kind == OR_DP) ? 2 : 0;
}

int or_rate_form(int kind) { return kind == OR_CREEP || kind == OR_DP; }

static double mag9(const double *A) {
    double s = 0.0;
    for (int i = 0; i < 9; ++i) s += A[i] * A[i];
    return sqrt(s);
}

static void lode(const double *A, double *z, double *s, double *E) {
    const double t = A[0] + A[4] + A[8];
    double d[9];
    *z = t / sqrt(3.0);
    for (int i = 0; i < 9; ++i) d[i] = A[i];
    for (int i = 0; i < 3; ++i) d[4 * i] -= t / 3.0;
    *s = mag9(d);
    for (int i = 0; i < 9; ++i) E[i] = fabs(*s) > 1.0e-10 ? d[i] / *s : 0.0;
}

static void drucker_prager(const or_material *mat, double mu, const double H[9], const double L[9], double dt,
                           const double Pn[9], const double *state_n, double *state_np1, double P[9]) {
    const double K = mat->bulk, G = mu, ez = 1.0 / sqrt(3.0);
    double F[9], Fi[9], J, tau[9], de[9], dstr[9], dep[9] = {0}, Es[9], zt, st;
    for (int i = 0; i < 9; ++i) F[i] = H[i];
    F[0] += 1.0; F[4] += 1.0; F[8] += 1.0;
    or_inverse3_cofactor(F, Fi, &J);
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            double s = 0.0;
            for (int k = 0; k < 3; ++k) s += Pn[3 * i + k] * F[3 * j + k];
            tau[3 * i + j] = s;
            de[3 * i + j] = (0.5 * dt) * (L[3 * i + j] + L[3 * j + i]);
            dstr[3 * i + j] = 0.5 * (H[3 * i + j] + H[3 * j + i]) * dt;
        }
    const double trde = de[0] + de[4] + de[8];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            const double dd = de[3 * i + j] - (i == j ? trde / 3.0 : 0.0);
            tau[3 * i + j] += (i == j ? K * trde : 0.0) + 2 * G * dd;
        }
    lode(tau, &zt, &st, Es);
    double f = mat->dp_A1 - mat->dp_A2 * zt, g = st / sqrt(2.0) - f;
    int failed = 0;
    if (g > 1.0e-4) {
        int converged = 0;
        const double Nz = mat->dp_A2, Ns = 1.0 / sqrt(2.0), Mz = mat->dp_A2G, Ms = Ns;
        for (int it = 0; it < 25; ++it) {
            const double Az = K * Mz, As = 2.0 * G * Ms;
            const double fac = sqrt((zt * zt + st * st) / (Mz * Mz + Ms * Ms)) / K;
            const double Pz = fac * Az, Ps = fac * As;
            const double beta = -g / (Nz * Pz + Ns * Ps);
            if (fabs(beta) < 1.0e-10) { converged = 1; break; }
            zt += beta * Pz;
            st += beta * Ps;
            f = mat->dp_A1 - mat->dp_A2 * zt;
            g = st / sqrt(2.0) - f;
        }
        failed = !converged;
        double dtau[9], dz, ds, Ed[9];
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) {
                const int k = 3 * i + j;
                dtau[k] = zt * (i == j ? ez : 0.0) + st * Es[k] - tau[k];
                tau[k] += dtau[k];
            }
        lode(dtau, &dz, &ds, Ed);
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) {
                const int k = 3 * i + j;
                const double dee = 1.0 / 3.0 / K * dz * (i == j ? ez : 0.0) + 1.0 / 2.0 / G * ds * Ed[k];
                dep[k] = dstr[k] - dee;
            }
    }
    const double tp = dep[0] + dep[4] + dep[8];
    double dd2 = 0.0;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            const double x = dep[3 * i + j] - (i == j ? tp / 3.0 : 0.0);
            dd2 += x * x;
        }
    const double deqps = sqrt(dd2 + (tp * tp) / 3.0);
    if (state_np1) {
        state_np1[0] = failed ? NAN : (state_n ? state_n[0] : 0.0) + deqps;
        state_np1[1] = failed ? NAN : (state_n ? state_n[1] : 0.0) + tp;
    }
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            double s = 0.0;
            for (int k = 0; k < 3; ++k) s += tau[3 * i + k] * Fi[3 * j + k];
            P[3 * i + j] = failed ? NAN : s;
        }
}

void or_stress_rate(const or_material *mat, double lambda, double mu, const double H[9], const double Hn[9],
                    const double Pn[9], double dt, const double *state_n, double *state_np1, double P[9]) {
    (void)lambda;
    if (mat->kind == OR_DP) {
        double Fn[9], Fni[9], L[9];
        for (int i = 0; i < 9; ++i) Fn[i] = Hn[i];
        Fn[0] += 1.0; Fn[4] += 1.0; Fn[8] += 1.0;
        or_fullpivlu_inverse(3, Fn, Fni);
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) {
                double s = 0.0;
                for (int k = 0; k < 3; ++k)
                    s += (dt != 0.0 ? (H[3 * i + k] - Hn[3 * i + k]) / dt : 0.0) * Fni[3 * k + j];
                L[3 * i + j] = s;
            }
        drucker_prager(mat, mu, H, L, dt, Pn, state_n, state_np1, P);
        return;
    }
    const double G = mu, K = mat->bulk;
    double F[9], Fi[9], Fn[9], Fni[9], L[9], J;
    for (int i = 0; i < 9; ++i) {
        F[i] = H[i];
        Fn[i] = Hn[i];
    }
    F[0] += 1.0; F[4] += 1.0; F[8] += 1.0;
    Fn[0] += 1.0; Fn[4] += 1.0; Fn[8] += 1.0;
    or_inverse3_cofactor(F, Fi, &J);
    or_fullpivlu_inverse(3, Fn, Fni);
    /* L = (H - H_n) / dt (I + H_n)^-1; a zero increment gives L = 0 */
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            double s = 0.0;
            for (int k = 0; k < 3; ++k) s += (dt != 0.0 ? (H[3 * i + k] - Hn[3 * i + k]) / dt : 0.0) * Fni[3 * k + j];
            L[3 * i + j] = s;
        }
    /* sigma = (1/J) P_n F^T */
    double sig[9], sd[9], de[9], ded[9], cs[9], ds[9];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            double s = 0.0;
            for (int k = 0; k < 3; ++k) s += (1.0 / J * Pn[3 * i + k]) * F[3 * j + k];
            sig[3 * i + j] = s;
        }
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) de[3 * i + j] = 0.5 * (L[3 * i + j] + L[3 * j + i]) * dt;
    const double sm = (sig[0] + sig[4] + sig[8]) / 3.0, dm = (de[0] + de[4] + de[8]) / 3.0;
    for (int i = 0; i < 9; ++i) {
        sd[i] = sig[i];
        ded[i] = de[i];
    }
    for (int i = 0; i < 3; ++i) {
        sd[4 * i] -= sm;
        ded[4 * i] -= dm;
    }
    for (int i = 0; i < 9; ++i) cs[i] = ded[i] + sd[i] / 2.0 / G;
    const double T = mat->temperature;
    const double a = T == 0.0 ? mat->creep_A * exp(-mat->creep_m) : mat->creep_A * exp(-mat->creep_m / T);
    const double n = mat->creep_n, s32 = sqrt(3.0 / 2.0);
    double edot = a * pow(mag9(sd) * s32, n - 1.0);
    const double critt = edot == 0.0 ? 1.1 * dt : 4.0 * 0.98 / (6.0 * G * n * edot);
    const double delt = critt < dt ? critt : dt;
    for (int i = 0; i < 9; ++i) {
        const double rate = dt != 0.0 ? ded[i] / dt : 0.0;
        const double sdot = 2.0 * G * (rate - 1.5 * sd[i] * edot);
        sd[i] += sdot * delt;
    }
    edot = a * pow(mag9(sd) * s32, n - 1.0);
    const double mean = (sig[0] + sig[4] + sig[8]) / 3.0 + K * (de[0] + de[4] + de[8]);
    for (int i = 0; i < 9; ++i) {
        sig[i] = sd[i];
        cs[i] -= sd[i] / 2.0 / G;
    }
    for (int i = 0; i < 3; ++i) sig[4 * i] += mean;
    if (state_np1) state_np1[0] = (state_n ? state_n[0] : 0.0) + sqrt(2.0 / 3.0) * mag9(cs);
    for (int i = 0; i < 9; ++i) {
        const double rate = dt != 0.0 ? ded[i] / dt : 0.0;
        ds[i] = 2.0 * G * (rate - 3.0 * edot * sd[i] / 2.0);
    }
    if (state_np1) state_np1[1] = s32 * mag9(ds);
    /* P = (J sigma) F^-T */
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            double s = 0.0;
            for (int k = 0; k < 3; ++k) s += (J * sig[3 * i + k]) * Fi[3 * j + k];
            P[3 * i + j] = s;
        }
}

static double clampd(double x, double lo, double hi) { return x < lo ? lo : (x > hi ? hi : x); }

static void nh_damage(const or_material *mat, double lambda, double mu, const double H[9], const double *state_n,
                      double *state_np1, double P[9]) {
    const double dmax = clampd(mat->maximum_damage, 0.0, 1.0);
    const double dinc = clampd(mat->maximum_damage_increment, 0.0, 1.0);
    const double Dn = clampd(state_n ? state_n[0] : 0.0, 0.0, dmax);
    double BmI[9], F[9], Finv[9], tau[9];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            double hh = 0.0;
            for (int k = 0; k < 3; ++k) hh += H[3 * i + k] * H[3 * j + k];
            BmI[3 * i + j] = hh + H[3 * j + i] + H[3 * i + j];
        }
    const double I1 = BmI[0] + BmI[4] + BmI[8] + 3.0;
    if (state_np1) state_np1[1] = I1;
    const double lnJ = log1p(or_det_ap_im1(H));
    for (int i = 0; i < 9; ++i) tau[i] = mu * BmI[i];
    for (int i = 0; i < 3; ++i) tau[4 * i] += lambda * lnJ;
    for (int i = 0; i < 9; ++i) F[i] = H[i];
    F[0] += 1.0; F[4] += 1.0; F[8] += 1.0;
    or_fullpivlu_inverse(3, F, Finv);
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            double s = 0.0;
            for (int k = 0; k < 3; ++k) s += tau[3 * i + k] * Finv[3 * j + k];
            P[3 * i + j] = s;
        }
    double D = Dn;
    if (I1 > mat->I1_failure) {
        D = 1.0;
    } else if (I1 > mat->I1_critical) {
        D = pow((I1 - mat->I1_critical) / (mat->I1_failure - mat->I1_critical), mat->alpha);
        D = clampd(D, 0.0, 1.0);
        D = D > Dn ? D : Dn;
    }
    if (D == 0.0) {
        if (state_np1) state_np1[0] = 0.0;
        return;
    }
    D = D < Dn + dinc ? D : Dn + dinc;
    int failed = 0;
    if (D >= dmax) {
        failed = 1;
        D = dmax;
    }
    for (int i = 0; i < 9; ++i) P[i] *= 1.0 - D;
    if (failed && Dn < dmax) D = 2.0;
    if (state_np1) state_np1[0] = D;
}

void or_stress_mat(const or_material *mat, double lambda, double mu, const double H[9], const double *state_n,
                   double *state_np1, double P[9]) {
    if (mat->kind == OR_NH_DAMAGE) nh_damage(mat, lambda, mu, H, state_n, state_np1, P);
    else or_stress(mat->kind, lambda, mu, mat->yield_stress, mat->hardening_modulus, H, state_n, state_np1, P);
}
