"""Pins the oracle's element kinematics (oracle/or_element.c) against the golds of
test/unit_tests/QuadratureTests.cpp and closed forms, and the whole force loop
against the patch-test closed forms of test/unit_tests/PatchTestFixture.h."""
import math

import numpy as np
import pytest

import oracle as O
from aperi_hip import meshgen as M

TET_REF = np.array([[0.0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]])
TET_B = np.array([[-1.0, -1, -1], [1, 0, 0], [0, 1, 0], [0, 0, 1]])
HEX_REF = np.array([[0.0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0], [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]])
HEX_SB = 0.25 * np.array([[-1.0, -1, -1], [1, -1, -1], [1, 1, -1], [-1, 1, -1], [-1, -1, 1], [1, -1, 1],
                          [1, 1, 1], [-1, 1, 1]])


def rotation(axis, angle):
    # Eigen::AngleAxisd(angle, axis).toRotationMatrix() (Rodrigues)
    k = np.asarray(axis, float) / np.linalg.norm(axis)
    Kx = np.array([[0, -k[2], k[1]], [k[2], 0, -k[0]], [-k[1], k[0], 0]])
    return np.cos(angle) * np.eye(3) + np.sin(angle) * Kx + (1 - np.cos(angle)) * np.outer(k, k)


def tet_cases():
    # QuadratureTests.cpp:17-123
    out = [(TET_REF, TET_B, 1 / 6)]
    f = 0.78
    out.append((TET_REF * f, TET_B / f, f ** 3 / 6))
    out.append((TET_REF * f + [0.512, -4.79, 103.1], TET_B / f, f ** 3 / 6))
    R = rotation([1.2, -2.3, 1.4], 0.36 * math.pi)
    out.append((TET_REF @ R, TET_B @ R, 1 / 6))
    rng = np.random.default_rng(42)
    F = rng.uniform(-1, 1, (3, 3)) + 1.5 * np.eye(3)
    X = TET_REF @ F
    lo, hi = X.min(0), X.max(0)
    X = (X - lo) / (hi - lo)
    F = F @ np.diag(1 / (hi - lo))
    out.append((X, TET_B @ np.linalg.inv(F).T, np.linalg.det(F) / 6))
    return out


def hex_cases():
    # QuadratureTests.cpp:127-285 (smoothed hex golds)
    out = [(HEX_REF, HEX_SB, 1.0)]
    f = 0.78
    out.append((HEX_REF * f, HEX_SB / f, f ** 3))
    out.append((HEX_REF * f + [0.512, -4.79, 103.1], HEX_SB / f, f ** 3))
    R = rotation([1.2, -2.3, 1.4], 0.36 * math.pi)
    out.append((HEX_REF @ R, HEX_SB @ R, 1.0))
    rng = np.random.default_rng(43)
    F = rng.uniform(-1, 1, (3, 3)) + 1.5 * np.eye(3)
    out.append((HEX_REF @ F, HEX_SB @ np.linalg.inv(F).T, np.linalg.det(F)))
    s = 0.2
    Xs = np.array([[s, 0, 0], [1 + s, 0, 0], [1 - s, 1, 0], [-s, 1, 0], [s, 0, 1], [1 + s, 0, 1], [1 - s, 1, 1],
                   [-s, 1, 1]])
    b, bs = 0.25, s / 2
    Bs = np.array([[-b, -(b + bs), -b], [b, -(b - bs), -b], [b, b + bs, -b], [-b, b - bs, -b],
                   [-b, -(b + bs), b], [b, -(b - bs), b], [b, b + bs, b], [-b, b - bs, b]])
    out.append((Xs, Bs, 1.0))
    return out


@pytest.mark.parametrize("case", range(5))
def test_tet4_gauss_bmatrix_golds(case):
    X, Bexp, wexp = tet_cases()[case]
    B, w = O.tet4_bmatrix(X)
    assert np.allclose(B, Bexp, atol=1e-12, rtol=0)
    assert abs(w - wexp) <= 1e-12
    assert np.allclose(B.sum(0), 0.0, atol=1e-12)


@pytest.mark.parametrize("case", range(5))
def test_tet4_smoothed_bmatrix_golds(case):
    X, Bexp, wexp = tet_cases()[case]
    B, V = O.smoothed_tet4_bmatrix(X)
    assert np.allclose(B, Bexp, atol=1e-12, rtol=0)
    assert abs(V - wexp) <= 1e-12


@pytest.mark.parametrize("case", range(6))
def test_hex8_smoothed_bmatrix_golds(case):
    X, Bexp, wexp = hex_cases()[case]
    B, V = O.smoothed_hex8_bmatrix(X)
    assert np.allclose(B, Bexp, atol=1e-12, rtol=0)
    assert abs(V - wexp) <= 1e-12


def _hex_dN(xi):
    s = np.array([[-1, -1, -1], [1, -1, -1], [1, 1, -1], [-1, 1, -1], [-1, -1, 1], [1, -1, 1], [1, 1, 1], [-1, 1, 1]])
    dN = np.zeros((8, 3))
    for k in range(8):
        f = 1 + s[k] * xi
        dN[k] = 0.125 * s[k] * np.array([f[1] * f[2], f[0] * f[2], f[0] * f[1]])
    return dN


def test_hex8_gauss_bmatrix_affine_closed_form():
    # no reference gold: an affine hex has constant J, so B_q = dN(xi_q) J^-1 exactly
    a = 1 / math.sqrt(3)
    pts = a * np.array([[-1, -1, -1], [1, -1, -1], [1, 1, -1], [-1, 1, -1], [-1, -1, 1], [1, -1, 1], [1, 1, 1],
                        [-1, 1, 1]])
    rng = np.random.default_rng(7)
    F = rng.uniform(-0.3, 0.3, (3, 3)) + np.eye(3)
    Xref = 2 * HEX_REF - 1  # [-1,1]^3: J = I for the parent map
    X = Xref @ F.T + [0.3, -1.2, 4.0]
    V = 0.0
    for q in range(8):
        B, w = O.hex8_bmatrix(X, q)
        # parent map x = F xi + c  ->  J = F, B = dN F^-1
        assert np.allclose(B, _hex_dN(pts[q]) @ np.linalg.inv(F), atol=1e-13)
        assert np.allclose(B.sum(0), 0.0, atol=1e-13)
        assert w == pytest.approx(np.linalg.det(F), rel=1e-13)
        V += w
    assert V == pytest.approx(8 * np.linalg.det(F), rel=1e-13)


def test_element_volumes_sum_to_box():
    X, hc = M.hex_lattice(3, 2, 4, 1.5, 0.5, 2.0)
    assert O.element_volumes(1, X, hc).sum() == pytest.approx(1.5 * 0.5 * 2.0, rel=1e-13)
    tets = M.split_tets(X, hc)
    v = O.element_volumes(0, X, tets)
    assert (v > 0).all()
    assert v.sum() == pytest.approx(1.5, rel=1e-13)


@pytest.mark.parametrize("etype", [0, 1])
def test_patch_test_uniform_deformation(etype):
    # PatchTestFixture.h:138-169: a uniform H on an affine patch gives
    # zero force on interior nodes and face force sums -P n A; SVK E=1, nu=0.
    X, hc = M.hex_lattice(3, 3, 3)
    conn = hc if etype == 1 else M.split_tets(X, hc)
    Hu = np.array([[0.02, 0.01, 0.0], [0.0, -0.01, 0.005], [0.003, 0.0, 0.015]])
    u = X @ Hu.T
    lam, mu = O.lame(1.0, 0.0)
    f, H, P, _ = O.internal_force(etype, X, conn, O.SVK, lam, mu, u)
    assert np.allclose(H, Hu.ravel(), atol=1e-13)
    F = np.eye(3) + Hu
    Egl = 0.5 * (F.T @ F - np.eye(3))
    Pexp = F @ (2 * mu * Egl + lam * np.trace(Egl) * np.eye(3))
    assert np.allclose(P, Pexp.ravel(), atol=1e-13)
    interior = np.all((X > 1e-9) & (X < 1 - 1e-9), axis=1)
    assert np.abs(f[interior]).max() < 1e-14
    assert np.allclose(f.sum(0), 0.0, atol=1e-13)  # a uniform stress is self-equilibrated
    # Sum of f over the nodes of face x_d = 1 is -int grad(phi) P dV with phi the
    # interpolant of the face indicator (exactly linear in the last layer):
    # -P e_d * area; the opposite face gives +P e_d.
    for d in range(3):
        n = np.zeros(3)
        n[d] = 1.0
        f1 = f[np.abs(X[:, d] - 1) < 1e-12].sum(0)
        f0 = f[np.abs(X[:, d]) < 1e-12].sum(0)
        assert np.allclose(f1, -(Pexp @ n), atol=1e-13)
        assert np.allclose(f0, Pexp @ n, atol=1e-13)


def test_force_sign_and_volume_weighting():
    # one reference tet stretched in x: f = -B P^T w, so the node at x=1 is pulled back
    X = TET_REF.copy()
    u = np.zeros((4, 3))
    u[1, 0] = 0.01
    lam, mu = O.lame(1.0, 0.0)
    f, H, P, _ = O.internal_force(0, X, np.array([[0, 1, 2, 3]]), O.LINEAR_ELASTIC, lam, mu, u)
    assert H[0, 0] == pytest.approx(0.01)
    assert P[0, 0] == pytest.approx(0.01)  # E = 1, nu = 0 -> P = eps
    assert f[1, 0] == pytest.approx(-0.01 / 6)
    assert f[0, 0] == pytest.approx(0.01 / 6)
