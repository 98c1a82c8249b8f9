"""Pins of the oracle's nodal strain smoothing on thex meshes, its subcells and F-bar
(oracle/or_nodal.c), independent of the restatement itself:

  labelling     the active nodes of a thex mesh are exactly the tet vertices (each hex has
                one, its lowest id), each cell is the star of hexes around its vertex, the
                cells tile the domain; a plain hex lattice is refused (MeshLabelerProcessor.h
                :285-330)
  search        candidates are active nodes only; the grid search equals brute force
  kinematics    a linear displacement field gives H = G in every subcell and leaves F-bar
                inactive (KinematicsTest.cpp:65-133, nodal cases)
  F-bar         every F-bar has det = J-bar of its cell (FbarTest.cpp); the force is minus
                the gradient of the F-bar strain energy sum_s V_s W(I + H-bar_s) in the
                nodal displacements (central differences, neo-Hookean and linear elastic);
                without F-bar the same holds with H_s
  patch test    uniform shear: zero net force and sum_I f_I (x) x_I = -V P exactly, for the
                plain and the F-bar force (ForceTest.cpp:250-304 nodal / F-bar cases)
  mass          sum m = rho V (MassUtils.cpp:79-112 through the RK functions)
  integrator    nodes without degrees of freedom keep their values
"""
import numpy as np
import pytest

import oracle as O
from aperi_hip import meshgen as M

E, NU, RHO = 1e9, 0.3, 1000.0


def _block(n=3, seed=11):
    X, hc = M.hex_lattice(n, n, n)
    X = M.jitter(X, 1.0 / n, 0.1, seed, 1)
    tets = M.split_tets(X, hc)
    Xh, hx = M.thex(X, tets)
    return X, tets, Xh, hx


def _energy_density(kind, lam, mu, F):
    """Strain energy per reference volume whose derivative the oracle's stress is."""
    if kind == O.NEO_HOOKEAN:
        J = np.linalg.det(F)
        return 0.5 * mu * (np.einsum("...ij,...ij", F, F) - 3) - mu * np.log(J) + 0.5 * lam * np.log(J) ** 2
    eps = 0.5 * (F + np.swapaxes(F, -1, -2)) - np.eye(3)
    tr = np.trace(eps, axis1=-2, axis2=-1)
    return 0.5 * lam * tr ** 2 + mu * np.einsum("...ij,...ij", eps, eps)


def test_thex_labelling_and_cells():
    X, tets, Xh, hx = _block()
    act, bad = O.nodal_active(hx, len(Xh))
    assert bad == 0
    assert np.array_equal(np.flatnonzero(act), np.arange(len(X)))  # exactly the tet vertices
    hc, cs, ch, ss, se, csub = O.nodal_cells(hx, act, 0)
    assert len(cs) - 1 == len(X)
    for c in range(len(X)):
        members = ch[cs[c]:cs[c + 1]]
        assert np.all(np.any(hx[members] == c, axis=1))          # every hex of the cell holds vertex c
        assert len(members) == np.count_nonzero(np.any(tets == c, axis=1))  # one hex per tet at c
    assert np.array_equal(np.sort(ch), np.arange(len(hx)))      # the cells partition the hexes
    # subdomains 0: every hex a subcell; 1: every cell one subcell
    assert np.array_equal(ss, np.arange(len(hx) + 1)) and np.array_equal(csub, cs)
    _, cs1, ch1, ss1, se1, csub1 = O.nodal_cells(hx, act, 1)
    assert np.array_equal(ss1, cs1) and np.array_equal(csub1, np.arange(len(X) + 1))
    # a plain hex lattice is not a refined mesh: hexes with two or more lowest-id nodes
    Xp, hp = M.hex_lattice(3, 3, 3)
    assert O.nodal_active(hp, len(Xp))[1] > 0


@pytest.mark.parametrize("subdomains", [0, 1])
def test_search_volumes_and_mass(subdomains):
    X, tets, Xh, hx = _block(3, 5)
    d = O.nodal_setup(Xh, hx, 0.9, subdomains, density=RHO)
    act = d["active"].astype(bool)
    valid = np.arange(40)[None, :] < d["num"][:, None]
    assert np.all(act[d["nbr"][valid]])  # only active nodes are neighbors
    num_b, nbr_b, tr_b = O.neighbor_search(Xh, d["radius"], brute_force=True, range_mask=d["active"])
    assert np.array_equal(num_b, d["num"]) and np.array_equal(nbr_b, d["nbr"]) and tr_b == d["trunc"]
    assert d["num"].min() >= 4
    vol_t = O.element_volumes(O.TET4, X, tets).sum()
    assert d["cells"][3].sum() == pytest.approx(vol_t, rel=1e-13)
    assert d["cell_volume"].sum() == pytest.approx(vol_t, rel=1e-13)
    assert d["mass"].sum() == pytest.approx(RHO * vol_t, rel=1e-12)
    assert np.all(d["mass"][~act] == 0.0) and np.all(d["mass"][act] > 0.0)


@pytest.mark.parametrize("subdomains", [0, 1])
def test_linear_field_gives_exact_gradient_and_inactive_fbar(subdomains):
    X, tets, Xh, hx = _block(3, 7)
    d = O.nodal_setup(Xh, hx, 0.9, subdomains)
    G = np.array([[0.01, 0.02, -0.01], [0.005, -0.02, 0.0], [0.0, 0.01, 0.03]])
    u = Xh @ G.T + np.array([1e-3, -2e-3, 0.5e-3])
    lam, mu = O.lame(E, NU)
    f, H, P, _ = O.cell_internal_force(len(Xh), d["cells"], O.NEO_HOOKEAN, lam, mu, u)
    assert np.abs(H - G.reshape(1, 9)).max() < 1e-14
    fb, Hf, Hb, Pf, Pb, _ = O.fbar_internal_force(len(Xh), d["cells"], (d["cell_sub"], d["cell_volume"]),
                                                  O.NEO_HOOKEAN, lam, mu, u)
    assert np.abs(Hb - Hf).max() < 1e-15 and np.abs(Pf - P).max() <= 1e-12 * np.abs(P).max()


def test_fbar_projection_has_the_cell_jbar():
    X, tets, Xh, hx = _block(3, 9)
    d = O.nodal_setup(Xh, hx, 0.9, 0)
    u = 2e-2 * (2 * M.uniform01(3, 4, Xh.size).reshape(Xh.shape) - 1)
    lam, mu = O.lame(E, NU)
    _, H, Hb, _, _, _ = O.fbar_internal_force(len(Xh), d["cells"], (d["cell_sub"], d["cell_volume"]),
                                              O.NEO_HOOKEAN, lam, mu, u)
    I = np.eye(3)
    J = np.linalg.det(I + H.reshape(-1, 3, 3))
    Jb = np.linalg.det(I + Hb.reshape(-1, 3, 3))
    V = d["cells"][3]
    cs = d["cell_sub"]
    jbar = np.array([np.sum(J[cs[c]:cs[c + 1]] * V[cs[c]:cs[c + 1]]) / d["cell_volume"][c] for c in range(len(cs) - 1)])
    cell_of_sub = np.repeat(np.arange(len(cs) - 1), np.diff(cs))
    assert np.abs(Jb - jbar[cell_of_sub]).max() < 1e-12
    assert np.abs(J - jbar[cell_of_sub]).max() > 1e-4  # the field is far from isochoric-uniform


@pytest.mark.parametrize("kind", [O.NEO_HOOKEAN, O.LINEAR_ELASTIC], ids=["nh", "le"])
@pytest.mark.parametrize("fbar", [True, False], ids=["fbar", "plain"])
def test_force_is_minus_energy_gradient(kind, fbar):
    X, tets, Xh, hx = _block(2, 13)
    d = O.nodal_setup(Xh, hx, 0.9, 0)
    lam, mu = O.lame(E, NU)
    V = d["cells"][3]
    act = np.flatnonzero(d["active"])

    def force_and_H(u):
        if fbar:
            f, H, Hb, *_ = O.fbar_internal_force(len(Xh), d["cells"], (d["cell_sub"], d["cell_volume"]), kind, lam,
                                                 mu, u)
            return f, Hb
        f, H, _, _ = O.cell_internal_force(len(Xh), d["cells"], kind, lam, mu, u)
        return f, H

    def energy(u):
        _, Hs = force_and_H(u)
        return np.sum(V * _energy_density(kind, lam, mu, np.eye(3) + Hs.reshape(-1, 3, 3)))

    u = 1e-2 * (2 * M.uniform01(5, 6, Xh.size).reshape(Xh.shape) - 1)
    f, _ = force_and_H(u)
    rng = np.random.default_rng(3)
    for trial in range(4):
        du = np.zeros_like(u)
        du[act] = rng.normal(size=(len(act), 3))
        h = 1e-6
        fd = (energy(u + h * du) - energy(u - h * du)) / (2 * h)
        assert -np.sum(f * du) == pytest.approx(fd, rel=1e-6)
    # single degrees of freedom
    for i in rng.choice(act, 3, replace=False):
        for k in range(3):
            du = np.zeros_like(u)
            du[i, k] = 1.0
            h = 1e-7
            fd = (energy(u + h * du) - energy(u - h * du)) / (2 * h)
            assert -f[i, k] == pytest.approx(fd, rel=1e-5, abs=1e-7 * np.abs(f).max())


@pytest.mark.parametrize("fbar", [True, False], ids=["fbar", "plain"])
def test_uniform_shear_patch_identities(fbar):
    X, tets, Xh, hx = _block(3, 17)
    d = O.nodal_setup(Xh, hx, 0.9, 0)
    gamma = 0.01
    u = np.zeros_like(Xh)
    u[:, 0] = gamma * Xh[:, 1]
    lam, mu = O.lame(E, NU)
    if fbar:
        f, H, _, P, _, _ = O.fbar_internal_force(len(Xh), d["cells"], (d["cell_sub"], d["cell_volume"]),
                                                 O.NEO_HOOKEAN, lam, mu, u)
    else:
        f, H, P, _ = O.cell_internal_force(len(Xh), d["cells"], O.NEO_HOOKEAN, lam, mu, u)
    G = np.zeros((3, 3))
    G[0, 1] = gamma
    assert np.abs(H - G.reshape(1, 9)).max() < 1e-14
    P0 = P[0].reshape(3, 3)
    assert np.abs(P - P[0]).max() < 1e-12 * np.abs(P0).max()
    Vt = d["cells"][3].sum()
    assert np.abs(f.sum(axis=0)).max() < 1e-12 * np.abs(f).max()
    moment = f.T @ Xh  # sum_I f_I (x) x_I
    assert np.abs(moment + Vt * P0).max() < 1e-11 * np.abs(Vt * P0).max()


def test_inactive_nodes_are_not_integrated():
    X, tets, Xh, hx = _block(2, 19)
    d = O.nodal_setup(Xh, hx, 0.9, 0, density=RHO)
    lam, mu = O.lame(E, NU)
    act = d["active"].astype(bool)
    mass = np.where(act, d["mass"], 0.0)
    ref = O.Solver(-1, Xh, None, O.NEO_HOOKEAN, lam, mu, mass, cells=d["cells"],
                   fbar_cells=(d["cell_sub"], d["cell_volume"]), active=d["active"])
    v0 = 3.0 * M.normal(2, 3, Xh.size).reshape(Xh.shape)
    ref.v[:] = v0
    dt = 0.2 * M.stable_dt_estimate(Xh, hx, E, NU, RHO)
    ref.run(dt, 20)
    assert np.array_equal(ref.v[~act], v0[~act]) and np.all(ref.u[~act] == 0.0)
    assert np.all(np.isfinite(ref.u)) and np.abs(ref.u[act]).max() > 0.0
