"""Pins the oracle's stress updates (oracle/or_material.c) by closed forms and
invariants that a wrong sign, index or transpose would break."""
import math

import numpy as np
import pytest

import oracle as O

LAM, MU = O.lame(200e9, 0.3)


def _rot(axis, ang):
    k = np.asarray(axis, float) / np.linalg.norm(axis)
    Kx = np.array([[0, -k[2], k[1]], [k[2], 0, -k[0]], [-k[1], k[0], 0]])
    return np.cos(ang) * np.eye(3) + np.sin(ang) * Kx + (1 - np.cos(ang)) * np.outer(k, k)


def test_linear_elastic_hooke():
    rng = np.random.default_rng(1)
    for _ in range(20):
        H = rng.uniform(-1e-3, 1e-3, (3, 3))
        P, _ = O.stress(O.LINEAR_ELASTIC, LAM, MU, H)
        eps = 0.5 * (H + H.T)
        assert np.allclose(P, LAM * np.trace(eps) * np.eye(3) + 2 * MU * eps, rtol=1e-14, atol=1e-3)
        # skew part of H carries no stress
        W = 0.5 * (H - H.T)
        P2, _ = O.stress(O.LINEAR_ELASTIC, LAM, MU, H - W)
        assert np.allclose(P, P2, rtol=1e-13, atol=1e-3)


def test_neo_hookean_simple_shear():
    # test/material_tests/neohookean.ipynb: tau = lambda ln J I + mu (B - I); simple shear
    # F = I + g e1 e2 has J = 1, so sigma_12 = mu g, sigma_11 = mu g^2, sigma_22 = sigma_33 = 0
    g = 0.2
    H = np.zeros((3, 3))
    H[0, 1] = g
    P, _ = O.stress(O.NEO_HOOKEAN, LAM, MU, H)
    F = np.eye(3) + H
    sigma = P @ F.T  # J = 1: sigma = P F^T / J
    assert sigma[0, 1] == pytest.approx(MU * g, rel=1e-13)
    assert sigma[0, 0] == pytest.approx(MU * g * g, rel=1e-13)
    assert abs(sigma[1, 1]) < 1e-13 * MU and abs(sigma[2, 2]) < 1e-13 * MU
    assert np.allclose(sigma, sigma.T, atol=1e-4)


def test_neo_hookean_uniaxial_stretch():
    # F = diag(s, 1, 1): J = s, tau_11 = lambda ln s + mu (s^2 - 1), tau_22 = lambda ln s.
    # ln s is evaluated as log1p(s - 1): near s = 1 that is the accuracy DetApIm1 + log1p
    # exist for (NeoHooke.h:80-86); log(s) itself would be off by ~1e-7 relative here.
    for s in (0.7, 0.95, 1.0 + 1e-9, 1.3):
        e = s - 1.0
        H = np.diag([e, 0.0, 0.0])
        P, _ = O.stress(O.NEO_HOOKEAN, LAM, MU, H)
        assert P[0, 0] == pytest.approx((LAM * math.log1p(e) + MU * e * (2.0 + e)) / s, rel=1e-12)
        assert P[1, 1] == pytest.approx(LAM * math.log1p(e), rel=1e-12, abs=1e-6)
        assert abs(P[0, 1]) + abs(P[1, 0]) + abs(P[0, 2]) == 0.0


def test_neo_hookean_objectivity_and_small_strain():
    rng = np.random.default_rng(2)
    H = rng.uniform(-0.05, 0.05, (3, 3))
    P, _ = O.stress(O.NEO_HOOKEAN, LAM, MU, H)
    R = _rot([0.3, -1.0, 0.4], 0.7)
    P_rot, _ = O.stress(O.NEO_HOOKEAN, LAM, MU, R @ (np.eye(3) + H) - np.eye(3))
    assert np.allclose(P_rot, R @ P, rtol=1e-10, atol=1e-12 * np.abs(P).max())
    h = 1e-7 * rng.uniform(-1, 1, (3, 3))
    Pnh, _ = O.stress(O.NEO_HOOKEAN, LAM, MU, h)
    Ple, _ = O.stress(O.LINEAR_ELASTIC, LAM, MU, h)
    assert np.allclose(Pnh, Ple, rtol=1e-5, atol=1e-6 * np.abs(Ple).max())


def test_j2_uniaxial_radial_return():
    # isochoric uniaxial strain: s_trial = 2 mu dev(eps), q_trial = 3 mu e.
    E, nu, sy, Hh = 70e9, 0.33, 300e6, 1e9
    lam, mu = O.lame(E, nu)
    for e, ep_n in ((0.5 * sy / (3 * mu), 0.0), (4 * sy / (3 * mu), 0.0), (6 * sy / (3 * mu), 1e-3)):
        H = np.diag([e, -e / 2, -e / 2])
        P, ep = O.stress(O.J2, lam, mu, H, state_n=ep_n, yield_stress=sy, hardening=Hh)
        q_trial = 3 * mu * e
        if q_trial <= sy + Hh * ep_n:
            assert ep == ep_n
            assert P[0, 0] == pytest.approx(2 * mu * e, rel=1e-13)
        else:
            dep = (q_trial - sy - Hh * ep_n) / (3 * mu + Hh)
            assert ep == pytest.approx(ep_n + dep, rel=1e-12)
            s = P - np.trace(P) / 3 * np.eye(3)
            q = math.sqrt(1.5) * np.linalg.norm(s)
            assert q == pytest.approx(sy + Hh * ep, rel=1e-12)  # consistency: back on the yield surface
        assert abs(np.trace(P)) < 1e-6 * sy  # isochoric: no pressure


def test_j2_pressure_is_elastic():
    E, nu = 70e9, 0.33
    lam, mu = O.lame(E, nu)
    H = 1e-2 * np.eye(3)  # pure dilatation never yields in J2
    P, ep = O.stress(O.J2, lam, mu, H, yield_stress=300e6, hardening=1e9)
    assert ep == 0.0
    assert np.allclose(P, (lam + 2 * mu / 3) * 3e-2 * np.eye(3), rtol=1e-13)


def test_svk_patch_closed_form():
    # PatchTestFixture.h:138-169: P = F S, S = lambda tr(E) I + 2 mu E, E = 1/2 (F^T F - I)
    lam, mu = O.lame(1.0, 0.0)
    H = np.array([[0.1, 0.02, 0.0], [0.01, -0.05, 0.03], [0.0, 0.04, 0.2]])
    P, _ = O.stress(O.SVK, lam, mu, H)
    F = np.eye(3) + H
    Egl = 0.5 * (F.T @ F - np.eye(3))
    assert np.allclose(P, F @ (lam * np.trace(Egl) * np.eye(3) + 2 * mu * Egl), rtol=1e-14, atol=1e-15)


# ---- neo-Hookean with damage (include/Materials/NeoHookeWithDamage.h:180-243) ----
# The reference has no test of this material; the closed forms below pin it.

DMG = dict(I1_critical=3.1, I1_failure=4.0, alpha=1.5)


def _uniaxial_nh(lam, mu, l):
    # F = diag(l, 1, 1): tau = lam ln(l) I + mu (B - I), P = tau F^-T
    P = np.zeros((3, 3))
    P[0, 0] = (lam * np.log(l) + mu * (l * l - 1.0)) / l
    P[1, 1] = P[2, 2] = lam * np.log(l)
    return P


def test_damage_free_is_neo_hookean():
    lam, mu = O.lame(1e9, 0.3)
    rng = np.random.default_rng(7)
    for _ in range(20):
        H = 0.01 * rng.standard_normal((3, 3))
        P, st = O.stress_damage(lam, mu, H, DMG)
        Pn, _ = O.stress(O.NEO_HOOKEAN, lam, mu, H)
        assert np.array_equal(P, Pn)
        F = np.eye(3) + H
        assert st[0] == 0.0 and st[1] == pytest.approx(np.trace(F @ F.T), rel=1e-14)


@pytest.mark.parametrize("l", [1.05, 1.1, 1.2])
def test_damage_uniaxial_closed_form(l):
    lam, mu = O.lame(2e9, 0.25)
    H = np.diag([l - 1.0, 0.0, 0.0])
    P, st = O.stress_damage(lam, mu, H, DMG)
    I1 = l * l + 2.0
    D = ((I1 - DMG["I1_critical"]) / (DMG["I1_failure"] - DMG["I1_critical"])) ** DMG["alpha"] if I1 > 3.1 else 0.0
    assert st[1] == pytest.approx(I1, rel=1e-15)
    assert st[0] == pytest.approx(D, rel=1e-14)
    assert np.allclose(P, (1.0 - D) * _uniaxial_nh(lam, mu, l), rtol=1e-13, atol=1e-6)


def test_damage_increment_cap_and_no_healing():
    lam, mu = O.lame(2e9, 0.25)
    H = np.diag([0.2, 0.0, 0.0])  # I1 = 3.44: D = (0.34 / 0.9)^1.5 = 0.2322
    full = ((1.2 ** 2 + 2 - 3.1) / 0.9) ** 1.5
    P, st = O.stress_damage(lam, mu, H, dict(DMG, maximum_damage_increment=0.05), state_n=(0.1, 0.0))
    assert st[0] == pytest.approx(0.15, rel=1e-15)  # min(D, Dn + 0.05)
    assert np.allclose(P, 0.85 * _uniaxial_nh(lam, mu, 1.2), rtol=1e-13)
    P, st = O.stress_damage(lam, mu, H, DMG, state_n=(0.1, 0.0))
    assert st[0] == pytest.approx(full, rel=1e-14)
    # unloading below the critical I1 keeps the damage and its softening
    Hs = np.diag([0.01, 0.0, 0.0])
    P, st = O.stress_damage(lam, mu, Hs, DMG, state_n=(0.4, 0.0))
    assert st[0] == 0.4 and np.allclose(P, 0.6 * _uniaxial_nh(lam, mu, 1.01), rtol=1e-13)
    # a larger stored damage is never reduced by a smaller evolution value
    P, st = O.stress_damage(lam, mu, H, DMG, state_n=(0.5, 0.0))
    assert st[0] == 0.5


def test_damage_failure_marker_and_cap():
    lam, mu = O.lame(2e9, 0.25)
    H = np.diag([0.5, 0.0, 0.0])  # I1 = 4.25 > I1_failure
    P, st = O.stress_damage(lam, mu, H, DMG)
    assert st[0] == 2.0  # new failure (NeoHookeWithDamage.h:238-240)
    assert np.allclose(P, 0.01 * _uniaxial_nh(lam, mu, 1.5), rtol=1e-13)
    P, st = O.stress_damage(lam, mu, H, DMG, state_n=(2.0, 0.0))
    assert st[0] == 0.99  # clamped on input, no longer new
    assert np.allclose(P, 0.01 * _uniaxial_nh(lam, mu, 1.5), rtol=1e-13)
    P, st = O.stress_damage(lam, mu, H, dict(DMG, maximum_damage=0.9))
    assert st[0] == 2.0 and np.allclose(P, 0.1 * _uniaxial_nh(lam, mu, 1.5), rtol=1e-13)


def test_damage_in_elements_and_cells():
    # the element loop stores (damage, I1) per element; the smoothed-cell path turns a
    # new failure's marker into the maximum damage (ComputeInternalForceSmoothedCell.h:219-223)
    X = np.array([[0.0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]])
    t = np.array([[0, 1, 2, 3]])
    lam, mu = O.lame(2e9, 0.25)
    u = X @ np.diag([0.5, 0.0, 0.0]).T
    f, H, P, st = O.internal_force(O.TET4, X, t, O.NH_DAMAGE, lam, mu, u, damage=DMG)
    assert st.shape == (2,) and st[0] == 2.0 and st[1] == pytest.approx(4.25, rel=1e-14)
    B, V = O.smoothed_tet4_bmatrix(X)
    cells = (np.array([0, 4]), np.arange(4, dtype=np.int32), B, np.array([V]))
    fc, Hc, Pc, sc = O.cell_internal_force(4, cells, O.NH_DAMAGE, lam, mu, u, damage=DMG)
    assert sc[0] == 0.99 and sc[1] == pytest.approx(4.25, rel=1e-14)
    assert np.allclose(fc, f, rtol=1e-12, atol=1e-9 * np.abs(f).max())


# ---- power-law creep (include/Materials/PowerLawCreep.h:101-173) ----
# The reference has no test of this material; the closed forms below pin it.

def _creep(E, nu, A, n, m=0.0, T=0.0):
    return dict(A=A, n=n, m=m, constant_temperature=T, bulk=E / (3 * (1 - 2 * nu)))


def test_creep_without_creep_is_a_hooke_increment():
    # A = 0: from a stress-free state, sigma = lambda tr(eps) I + 2 mu eps with eps = sym(H),
    # P = J sigma F^-T; no creep strain; stress rate sqrt(3/2) |2 mu dev(eps)| / dt
    E, nu = 1e9, 0.3
    lam, mu = O.lame(E, nu)
    rng = np.random.default_rng(3)
    for _ in range(10):
        H = 1e-3 * rng.standard_normal((3, 3))
        dt = 1e-5
        P, st = O.stress_creep(lam, mu, H, np.zeros(9), np.zeros(9), dt, _creep(E, nu, 0.0, 3.0))
        eps = 0.5 * (H + H.T)
        sig = lam * np.trace(eps) * np.eye(3) + 2 * mu * eps
        F = np.eye(3) + H
        Pc = np.linalg.det(F) * sig @ np.linalg.inv(F).T
        assert np.allclose(P, Pc, rtol=0, atol=1e-12 * np.abs(Pc).max())
        dev = eps - np.trace(eps) / 3 * np.eye(3)
        assert abs(st[0]) < 1e-18
        assert st[1] == pytest.approx(np.sqrt(1.5) * np.linalg.norm(2 * mu * dev / dt), rel=1e-12)


@pytest.mark.parametrize("n", [1.0, 3.0])
def test_creep_relaxation_is_one_norton_euler_step(n):
    # no deformation (H = H_n = 0, F = I): the deviator relaxes by one explicit step of
    # s' = -3 G a seq^(n-1) s / ... over min(critt, dt); pressure is unchanged; the creep
    # strain is the relaxed elastic strain
    E, nu = 1e9, 0.3
    lam, mu = O.lame(E, nu)
    s0 = np.array([[2e6, 5e5, 0.0], [5e5, -1e6, 3e5], [0.0, 3e5, -1e6]]) + 4e6 * np.eye(3)
    sd0 = s0 - np.trace(s0) / 3 * np.eye(3)
    seq = np.sqrt(1.5) * np.linalg.norm(sd0)
    a = 1e-3 / seq ** (n - 1) / (3 * mu)  # edot such that 3 G edot = 1e-3 per unit time
    for dt in (1.0, 1e3):  # below and above the sub-step limit critt
        P, st = O.stress_creep(lam, mu, np.zeros((3, 3)), np.zeros(9), s0.reshape(9), dt,
                               _creep(E, nu, a, n), state_n=(0.25, 0.0))
        edot = a * seq ** (n - 1)
        critt = 4 * 0.98 / (6 * mu * n * edot)
        delt = min(critt, dt)
        sd1 = sd0 * (1 - 3 * mu * edot * delt)
        assert np.allclose(P - np.trace(P) / 3 * np.eye(3), sd1, rtol=0, atol=1e-9 * np.abs(sd0).max())
        assert np.trace(P) / 3 == pytest.approx(np.trace(s0) / 3, rel=1e-14)
        dcreep = np.sqrt(2 / 3) * np.linalg.norm((sd0 - sd1) / (2 * mu))
        assert st[0] == pytest.approx(0.25 + dcreep, rel=1e-12)


def test_creep_pressure_does_not_creep():
    E, nu = 1e9, 0.3
    lam, mu = O.lame(E, nu)
    rng = np.random.default_rng(5)
    H, Hn = 2e-3 * rng.standard_normal((3, 3)), 1e-3 * rng.standard_normal((3, 3))
    Pn = 1e6 * rng.standard_normal(9)
    p0 = O.stress_creep(lam, mu, H, Hn, Pn, 1e-4, _creep(E, nu, 0.0, 2.0))[0]
    p1 = O.stress_creep(lam, mu, H, Hn, Pn, 1e-4, _creep(E, nu, 1e-12, 2.0))[0]
    F = np.eye(3) + H
    J = np.linalg.det(F)
    s0, s1 = p0 @ F.T / J, p1 @ F.T / J
    assert np.trace(s0) == pytest.approx(np.trace(s1), rel=1e-12)
    assert np.abs(s0 - s1).max() > 1e-6 * np.abs(s0).max()  # the deviator did creep


# ---- Drucker-Prager (include/Materials/DruckerPrager.h:97-216) ----
# The reference has no test of this material; the closed forms below pin it.

def _dp(E, nu, A1, A2, A2G):
    return dict(A1=A1, A2=A2, A2G=A2G, bulk=E / (3 * (1 - 2 * nu)))


def _dp_stress(lam, mu, H, Pn, dt, prm, state_n=(0.0, 0.0)):
    return O.stress_creep(lam, mu, H, np.zeros(9), Pn, dt, prm, state_n, kind=O.DP)


def test_dp_elastic_is_a_kirchhoff_hooke_increment():
    E, nu = 1e9, 0.3
    lam, mu = O.lame(E, nu)
    rng = np.random.default_rng(8)
    for _ in range(10):
        H = 1e-4 * rng.standard_normal((3, 3))
        P, st = _dp_stress(lam, mu, H, np.zeros(9), 1e-5, _dp(E, nu, 1e9, 0.1, 0.1), state_n=(0.2, -0.1))
        eps = 0.5 * (H + H.T)
        tau = lam * np.trace(eps) * np.eye(3) + 2 * mu * eps
        Pc = tau @ np.linalg.inv(np.eye(3) + H).T
        assert np.allclose(P, Pc, rtol=0, atol=1e-12 * np.abs(Pc).max())
        assert st[0] == 0.2 and st[1] == -0.1  # no plastic increment


def _lode(A):
    z = np.trace(A) / np.sqrt(3)
    d = A - np.trace(A) / 3 * np.eye(3)
    return z, np.linalg.norm(d), d / np.linalg.norm(d)


def test_dp_cylinder_returns_radially():
    # A2 = A2G = 0: the trial pressure is kept and s goes to sqrt(2) A1 along the trial
    # deviator; the plastic strain has the trace of dstrain = dt sym(H)
    E, nu = 1e9, 0.3
    lam, mu = O.lame(E, nu)
    H = np.array([[2e-3, 5e-4, 0.0], [1e-4, -1e-3, 3e-4], [0.0, 2e-4, 5e-4]])
    dt = 1e-5
    A1 = 1e6
    P, st = _dp_stress(lam, mu, H, np.zeros(9), dt, _dp(E, nu, A1, 0.0, 0.0))
    eps = 0.5 * (H + H.T)
    trial = lam * np.trace(eps) * np.eye(3) + 2 * mu * eps
    z0, s0, E0 = _lode(trial)
    assert s0 / np.sqrt(2) > A1  # the trial state yields
    tau = P @ (np.eye(3) + H).T
    z1, s1, E1 = _lode(tau)
    assert z1 == pytest.approx(z0, rel=1e-12)
    assert s1 == pytest.approx(np.sqrt(2) * A1, rel=1e-12)
    assert np.allclose(E1, E0, atol=1e-12)
    assert st[1] == pytest.approx(np.trace(eps) * dt, rel=1e-10, abs=1e-22)


@pytest.mark.parametrize("A2,A2G", [(0.2, 0.2), (0.3, 0.1)])
def test_dp_return_lands_on_the_yield_surface(A2, A2G):
    E, nu = 1e9, 0.3
    lam, mu = O.lame(E, nu)
    H = np.array([[3e-3, 1e-3, 0.0], [0.0, -2e-3, 5e-4], [1e-4, 0.0, 1e-3]])
    A1 = 2e6
    P, st = _dp_stress(lam, mu, H, np.zeros(9), 1e-5, _dp(E, nu, A1, A2, A2G))
    eps = 0.5 * (H + H.T)
    z0, s0, E0 = _lode(lam * np.trace(eps) * np.eye(3) + 2 * mu * eps)
    assert s0 / np.sqrt(2) - (A1 - A2 * z0) > 1e-4
    z1, s1, E1 = _lode(P @ (np.eye(3) + H).T)
    assert abs(s1 / np.sqrt(2) - (A1 - A2 * z1)) < 1e-9 * A1  # consistency
    assert np.allclose(E1, E0, atol=1e-10)  # the trial deviatoric direction
    assert st[0] > 0.0
