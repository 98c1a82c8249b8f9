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
