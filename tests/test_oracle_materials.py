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
