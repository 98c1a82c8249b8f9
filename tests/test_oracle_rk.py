"""Pins the oracle's RKPM set-up and smoothed-cell force (oracle/or_rk.c) against
the neighbor counts of test/unit_tests/NeighborSearchProcessorTest.cpp and the
reproducing conditions of the RK functions."""
from collections import Counter

import numpy as np
import pytest

import oracle as O
from aperi_hip import meshgen as M


def generated(nx, ny, nz):
    # "generated:NXxNYxNZ|tets": unit spacing, 6 tets per hex
    X, hc = M.hex_lattice(nx, ny, nz, nx, ny, nz)
    return X, M.split_tets(X, hc)


def test_ball_search_small_and_large():
    X, _ = generated(1, 1, 4)
    num, nbr, tr = O.neighbor_search(X, np.full(len(X), 0.5))
    assert (num == 1).all() and (nbr[:, 0] == np.arange(len(X))).all()
    diag = np.sqrt(1 + 1 + 16)
    num, nbr, tr = O.neighbor_search(X, np.full(len(X), 1.01 * diag))
    assert (num == len(X)).all() and tr == 0


def test_ball_search_mid_counts():
    # NeighborSearchProcessorTest.cpp:123-158
    X, _ = generated(2, 2, 4)
    num, _, _ = O.neighbor_search(X, np.full(len(X), 1.01))
    assert sorted(Counter(num.tolist()).items()) == [(4, 8), (5, 20), (6, 14), (7, 3)]


def test_kernel_radius_counts():
    # NeighborSearchProcessorTest.cpp:183-207
    X, t = generated(1, 1, 4)
    r = O.kernel_radius(O.max_edge_length(X, t), 1.01)
    radii = Counter(np.round(r, 12).tolist())
    assert radii == {round(np.sqrt(2) * 1.01, 12): 12, round(np.sqrt(3) * 1.01, 12): 8}
    num, _, _ = O.neighbor_search(X, r)
    assert sorted(Counter(num.tolist()).items()) == [(7, 4), (8, 4), (10, 6), (11, 2), (12, 4)]


def test_variable_ball_search_stats():
    # NeighborSearchProcessorTest.cpp:160-181: min 8, max 27, avg 18.713
    X, t = generated(5, 5, 5)
    num, _, _ = O.neighbor_search(X, O.kernel_radius(O.max_edge_length(X, t), 1.1))
    assert num.min() == 8 and num.max() == 27
    assert num.mean() == pytest.approx(18.713, abs=1e-3)


def test_kernel_radius_scale_one_is_nudged():
    assert O.kernel_radius(np.array([2.0]), 1.0)[0] == 2.0 * (1.0 + 1.0e-6)
    with pytest.raises(ValueError):
        O.kernel_radius(np.array([2.0]), -1.0)


def test_inactive_nodes_double_edge_length():
    X, hc = M.hex_lattice(1, 1, 1)
    t = M.split_tets(X, hc)
    n0, n6 = hc[0, 0], hc[0, 6]  # ends of the body diagonal every tet shares
    active = np.ones(len(X), dtype=np.uint8)
    active[n0] = 0
    l_all = O.max_edge_length(X, t)
    l_act = O.max_edge_length(X, t, active)
    assert l_all[n6] == pytest.approx(np.sqrt(3))
    assert l_act[n6] == pytest.approx(2 * np.sqrt(3))  # the inactive end counts double


def jittered(n, seed=3):
    X, hc = M.hex_lattice(n, n, n)
    X = M.jitter(X, 1.0 / n, 0.1, seed, 1)
    return X, M.split_tets(X, hc)


@pytest.mark.parametrize("scale", [1.1, 1.5, 2.2])
def test_grid_search_equals_brute_force(scale):
    X, t = jittered(5)
    r = O.kernel_radius(O.max_edge_length(X, t), scale)
    a = O.neighbor_search(X, r)
    b = O.neighbor_search(X, r, brute_force=True)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]) and a[2] == b[2]


def test_neighbor_cap_sorting_and_truncation():
    X, t = jittered(6)
    r = O.kernel_radius(O.max_edge_length(X, t), 2.2)
    num, nbr, tr = O.neighbor_search(X, r)
    full, fnbr, _ = O.neighbor_search(X, r, max_neighbors=400)
    assert (num <= 40).all() and tr == int(np.maximum(full - 40, 0).sum()) and tr > 0
    tw = np.array([1 + 1e-6, 1 + 1e-7, 1 + 1e-8])
    for i in range(0, len(X), 7):
        ids = fnbr[i, :full[i]]
        d2 = ((X[ids] * tw - X[i]) ** 2).sum(1)
        order = np.lexsort((ids, d2))
        assert np.array_equal(nbr[i, :num[i]], ids[order][:40])  # nearest (tweaked) 40, ties by id
        assert np.all(np.linalg.norm(X[ids] - X[i], axis=1) <= r[ids] * (1 + 1e-15))


def _rk(n=6, scale=1.5, seed=3):
    X, t = jittered(n, seed)
    r = O.kernel_radius(O.max_edge_length(X, t), scale)
    num, nbr, _ = O.neighbor_search(X, r)
    phi = O.rk_function_values(X, r, num, nbr)
    return X, t, r, num, nbr, phi


def test_rk_partition_of_unity_and_linear_reproduction():
    # ShapeFunctionsFunctorsTest.cpp:182-272 (tolerance 4e-12 * 10^(order-1))
    X, t, r, num, nbr, phi = _rk()
    for i in range(len(X)):
        ids = nbr[i, :num[i]]
        assert phi[i, :num[i]].sum() == pytest.approx(1.0, abs=4e-12)
        assert np.allclose(phi[i, :num[i]] @ X[ids], X[i], atol=4e-12)
        assert not phi[i, num[i]:].any()


def test_rk_single_neighbor_and_shepard_fallback():
    X = np.array([[0.0, 0, 0], [10.0, 0, 0]])
    r = np.array([0.5, 0.5])
    num, nbr, _ = O.neighbor_search(X, r)
    phi = O.rk_function_values(X, r, num, nbr)
    assert (num == 1).all() and (phi[:, 0] == 1.0).all()
    # collinear neighbours: the moment matrix is rank deficient; Eigen's FullPivLU
    # zero-fills the null space and the functions still reproduce constants
    X = np.array([[0.0, 0, 0], [0.5, 0, 0], [1.0, 0, 0]])
    r = np.full(3, 1.2)
    num, nbr, _ = O.neighbor_search(X, r)
    phi = O.rk_function_values(X, r, num, nbr)
    assert phi[1, :num[1]].sum() == pytest.approx(1.0, abs=1e-12)


def test_cells_partition_of_nullity_and_linear_patch():
    # KinematicsTestFixture.h:338-357: a linear displacement gives the exact gradient
    X, t, r, num, nbr, phi = _rk()
    cells = O.rk_cells(X, t, num, nbr, phi)
    start, nodes, dphi, vol = cells
    assert vol.sum() == pytest.approx(O.element_volumes(0, X, t).sum(), rel=1e-12)
    for c in range(0, len(vol), 13):
        assert np.allclose(dphi[start[c]:start[c + 1]].sum(0), 0.0, atol=1e-10)
    Hu = np.array([[0.01, 0.002, 0.0], [0.0, -0.004, 0.003], [0.001, 0.0, 0.02]])
    u = X @ Hu.T
    lam, mu = O.lame(70e9, 0.33)
    f, H, P, _ = O.cell_internal_force(len(X), cells, O.LINEAR_ELASTIC, lam, mu, u)
    assert np.allclose(H, Hu.ravel(), atol=1e-12)


def test_cell_lists_are_first_seen_unions():
    X, t, r, num, nbr, phi = _rk(n=4)
    start, nodes, dphi, vol = O.rk_cells(X, t, num, nbr, phi)
    for c in range(len(t)):
        seen = []
        for a in t[c]:
            for j in nbr[a, :num[a]]:
                if j not in seen:
                    seen.append(j)
        assert nodes[start[c]:start[c + 1]].tolist() == seen


def test_rk_mass_sums_to_rho_v():
    X, t, r, num, nbr, phi = _rk(n=4)
    start, nodes, dphi, vol = O.rk_cells(X, t, num, nbr, phi)
    m = O.rk_mass(t, vol, 2700.0, num, nbr, phi)
    assert m.sum() == pytest.approx(2700.0 * vol.sum(), rel=1e-11)


# ---- two-pass method (ExplicitSolver::ComputeForce with force_two_pass_method) ----

def test_two_pass_cells_are_element_b_matrices():
    # a cell of the two-pass method lists its element's nodes; the face-smoothed
    # derivative of a linear tet is its exact (constant) gradient
    X, t, r, num, nbr, phi = _rk(n=4)
    start, nodes, dphi, vol = O.rk_cells_two_pass(X, t)
    assert np.array_equal(start, 4 * np.arange(len(t) + 1)) and np.array_equal(nodes, t.ravel())
    assert np.allclose(vol, O.element_volumes(0, X, t), rtol=1e-13, atol=0)
    for e in range(0, len(t), 7):
        B, V = O.tet4_bmatrix(X[t[e]])
        assert np.allclose(dphi[4 * e:4 * e + 4], B, rtol=0, atol=1e-11 * np.abs(B).max())


def test_gather_reproduces_linear_fields_and_scatter_is_its_transpose():
    X, t, r, num, nbr, phi = _rk()
    G = np.array([[1e-3, 2e-3, -1e-3], [4e-4, -2e-3, 3e-3], [1e-3, 0.0, 5e-4]])
    d = X @ G.T + np.array([0.1, -0.2, 0.3])
    assert np.abs(O.rk_gather(num, nbr, phi, d) - d).max() < 1e-14
    # the transpose applied as a gather does not reproduce (a planted Phi^T is caught)
    assert np.abs(O.rk_scatter(num, nbr, phi, d) - d).max() > 1e-3
    rng = np.random.default_rng(4)
    a, b = rng.standard_normal(d.shape), rng.standard_normal(d.shape)
    lhs = (O.rk_gather(num, nbr, phi, a) * b).sum()
    rhs = (a * O.rk_scatter(num, nbr, phi, b)).sum()
    assert lhs == pytest.approx(rhs, rel=1e-13)
    # dense check: gather is Phi, scatter is Phi^T
    n = len(X)
    Phi = np.zeros((n, n))
    for i in range(n):
        Phi[i, nbr[i, :num[i]]] = phi[i, :num[i]]
    assert np.abs(O.rk_gather(num, nbr, phi, a) - Phi @ a).max() < 1e-13 * np.abs(a).max()
    assert np.abs(O.rk_scatter(num, nbr, phi, b) - Phi.T @ b).max() < 1e-13 * np.abs(b).max()


def test_two_pass_linear_patch():
    X, t, r, num, nbr, phi = _rk()
    cells2 = O.rk_cells_two_pass(X, t)
    G = np.array([[0.01, 0.002, 0.0], [0.0, -0.004, 0.003], [0.001, 0.0, 0.02]])
    d = X @ G.T + 0.05
    lam, mu = O.lame(70e9, 0.33)
    f, H, P, _, up, fl = O.rk_two_pass_force(len(X), cells2, num, nbr, phi, O.LINEAR_ELASTIC, lam, mu, d)
    assert np.allclose(up, d, atol=1e-14)
    assert np.allclose(H, G.ravel(), atol=1e-12)


@pytest.mark.parametrize("kind", [O.NEO_HOOKEAN, O.J2])
def test_two_pass_force_equals_one_pass(kind):
    # one-pass: grad phi_J = sum_k grad N_k phi_kJ per cell; two-pass: Phi^T f(Phi d).
    # The same bilinear form in another order, so H, P and f agree to rounding;
    # a transposed Phi on either side breaks the agreement.
    X, t, r, num, nbr, phi = _rk()
    cells1 = O.rk_cells(X, t, num, nbr, phi)
    cells2 = O.rk_cells_two_pass(X, t)
    lam, mu = O.lame(70e9, 0.33)
    d = 4e-3 * (2 * M.uniform01(8, 7, X.size).reshape(X.shape) - 1)
    kw = dict(yield_stress=100e6, hardening=1e9)
    f1, H1, P1, s1 = O.cell_internal_force(len(X), cells1, kind, lam, mu, d, **kw)
    f2, H2, P2, s2, up, fl = O.rk_two_pass_force(len(X), cells2, num, nbr, phi, kind, lam, mu, d, **kw)
    if kind == O.J2:
        assert s1.max() > 0
    for a, b in ((H2, H1), (P2, P1), (f2, f1)):
        assert np.abs(a - b).max() < 1e-12 * np.abs(b).max()
    assert np.abs(s2 - s1).max() < 1e-12 * s1.max() if kind == O.J2 else True
    wrong = O.rk_scatter(num, nbr, phi, d)  # Phi^T in the gather
    fw = O.cell_internal_force(len(X), cells2, kind, lam, mu, wrong, **kw)[0]
    assert np.abs(O.rk_gather(num, nbr, phi, fw) - f1).max() > 1e-3 * np.abs(f1).max()


def test_two_pass_solver_matches_one_pass_over_steps():
    X, t, r, num, nbr, phi = _rk(n=4)
    cells1 = O.rk_cells(X, t, num, nbr, phi)
    cells2 = O.rk_cells_two_pass(X, t)
    lam, mu = O.lame(70e9, 0.33)
    mass = O.rk_mass(t, cells1[3], 2700.0, num, nbr, phi)
    assert np.array_equal(O.rk_mass(t, cells2[3], 2700.0, num, nbr, phi), mass)
    c = X - 0.5
    v0 = 60.0 * np.stack([-0.5 * c[:, 0], -0.5 * c[:, 1], c[:, 2]], axis=1)
    dt = 0.5 * M.stable_dt_estimate(X, t, 70e9, 0.33, 2700.0)
    a = O.Solver(-1, X, None, O.J2, lam, mu, mass, None, 300e6, 1e9, cells=cells1)
    b = O.Solver(-1, X, None, O.J2, lam, mu, mass, None, 300e6, 1e9, cells=cells2, two_pass=(num, nbr, phi))
    a.v[:] = v0
    b.v[:] = v0
    a.run(dt, 50)
    b.run(dt, 50)
    assert a.state.max() > 0
    for x, y in ((b.u, a.u), (b.v, a.v), (b.f, a.f), (b.P, a.P)):
        assert np.abs(x - y).max() < 1e-10 * np.abs(y).max()


# ---- Hex8 background (ElementReproducingKernelHex8: one cell per hex) ----------

def _rk_hex(n=4, scale=1.5, seed=11):
    X, hc = M.hex_lattice(n, n, n)
    X = M.jitter(X, 1.0 / n, 0.1, seed, 1)
    r = O.kernel_radius(O.max_edge_length(X, hc), scale)
    num, nbr, tr = O.neighbor_search(X, r)
    phi = O.rk_function_values(X, r, num, nbr)
    return X, hc, r, num, nbr, phi


def test_hex8_background_cells_linear_patch_volume_and_mass():
    # the face-smoothed Hex8 cell (QuadratureSmoothed.h:96-150) with RK functions:
    # derivatives sum to zero, a linear field gives its exact gradient, cell volumes
    # are the elements' smoothed volumes (which tile the unit cube when no node is
    # jittered), cell lists are first-seen unions over the eight nodes, and the RK mass
    # (rho V / 8 per element node, scattered by phi) sums to rho V
    X0, hc0 = M.hex_lattice(4, 4, 4)
    assert sum(O.smoothed_hex8_bmatrix(X0[h])[1] for h in hc0) == pytest.approx(1.0, rel=1e-13)
    X, hc, r, num, nbr, phi = _rk_hex()
    start, nodes, dphi, vol = O.rk_cells(X, hc, num, nbr, phi)
    assert np.allclose(vol, [O.smoothed_hex8_bmatrix(X[h])[1] for h in hc], rtol=1e-14, atol=0)
    for c in range(0, len(vol), 5):
        assert np.allclose(dphi[start[c]:start[c + 1]].sum(0), 0.0, atol=1e-10)
        seen = []
        for a in hc[c]:
            for j in nbr[a, :num[a]]:
                if j not in seen:
                    seen.append(j)
        assert nodes[start[c]:start[c + 1]].tolist() == seen
    Hu = np.array([[0.01, 0.002, 0.0], [0.0, -0.004, 0.003], [0.001, 0.0, 0.02]])
    u = X @ Hu.T + np.array([0.1, -0.2, 0.05])
    lam, mu = O.lame(70e9, 0.33)
    f, H, P, _ = O.cell_internal_force(len(X), (start, nodes, dphi, vol), O.LINEAR_ELASTIC, lam, mu, u)
    assert np.allclose(H, Hu.ravel(), atol=1e-12)
    m = O.rk_mass(hc, vol, 2700.0, num, nbr, phi)
    assert m.sum() == pytest.approx(2700.0 * vol.sum(), rel=1e-11)


def test_hex8_background_small_support_is_the_smoothed_element():
    # with supports smaller than any node spacing each node is its own only neighbor
    # (phi = 1), so a one-pass cell lists its hex's nodes and their derivatives are the
    # smoothed Hex8 B, the same as the two-pass cell
    X, hc, r, num, nbr, phi = _rk_hex(scale=0.3)
    assert (num == 1).all() and np.array_equal(nbr[:, 0], np.arange(len(X)))
    assert np.allclose(phi[:, 0], 1.0, rtol=0, atol=1e-15)
    start, nodes, dphi, vol = O.rk_cells(X, hc, num, nbr, phi)
    s2, n2, d2, v2 = O.rk_cells_two_pass(X, hc)
    assert np.array_equal(start, 8 * np.arange(len(hc) + 1)) and np.array_equal(nodes, hc.ravel())
    assert np.array_equal(s2, start) and np.array_equal(n2, nodes)
    for e in range(len(hc)):
        B, V = O.smoothed_hex8_bmatrix(X[hc[e]])
        assert vol[e] == pytest.approx(V, rel=1e-15) and v2[e] == pytest.approx(V, rel=1e-15)
        assert np.allclose(dphi[8 * e:8 * e + 8], B, rtol=0, atol=1e-13 * np.abs(B).max())
        assert np.allclose(d2[8 * e:8 * e + 8], B, rtol=0, atol=1e-13 * np.abs(B).max())
