"""GPU parity of the RKPM set-up and the smoothed-cell force step.

Neighbor lists must be bit-exact (same sets, same order); RK functions, cell
derivatives, mass and forces within 1e-12 relative per evaluation and 1e-10
after 100 J2 steps.  Through the C ABI.
"""
from collections import Counter

import numpy as np
import pytest

import oracle as O
import aperi_hip as A
from aperi_hip import meshgen as M
from conftest import rel_err

pytestmark = pytest.mark.gpu

J2 = dict(E=70e9, nu=0.33, rho=2700.0, yield_stress=300e6, hardening=1e9)


def _mesh(n, seed=3, jit=0.1, spacing=None):
    L = n if spacing else 1.0
    X, hc = M.hex_lattice(n, n, n, L, L, L)
    if jit:
        X = M.jitter(X, L / n, jit, seed, 1)
    return X, M.split_tets(X, hc)


def _oracle_rk(X, t, scale):
    r = O.kernel_radius(O.max_edge_length(X, t), scale)
    num, nbr, tr = O.neighbor_search(X, r)
    phi = O.rk_function_values(X, r, num, nbr)
    return r, num, nbr, tr, phi


def _gpu(X, t, scale, kind=A.J2, p=J2):
    s = A.Solver()
    s.set_mesh(A.TET4, X, t)
    s.set_material(kind, p["rho"], p["E"], p["nu"], p.get("yield_stress", 0.0), p.get("hardening", 0.0))
    s.rk_setup(scale)
    return s


@pytest.mark.parametrize("n,scale", [(6, 1.5), (7, 1.1), (5, 2.2)])
def test_neighbors_bit_exact_and_phi(gpu_available, n, scale):
    X, t = _mesh(n)
    r, num, nbr, tr, phi = _oracle_rk(X, t, scale)
    print(f"rk n={n} scale={scale} nodes={len(X)} trunc={tr}", flush=True)
    with _gpu(X, t, scale) as s:
        gnum, gnbr, gtr = s.rk_neighbors()
        assert np.array_equal(gnum, num)
        assert np.array_equal(gnbr, np.where(np.arange(40)[None, :] < num[:, None], nbr, -1))
        assert gtr == tr
        gphi = s.rk_function_values()
        assert np.abs(gphi - phi).max() < 1e-12
        start, nodes, dphi, vol = s.rk_cells()
        cs, cn, cd, cv = O.rk_cells(X, t, num, nbr, phi)
        assert np.array_equal(start, cs) and np.array_equal(nodes, cn)
        assert rel_err(dphi, cd) < 1e-12
        assert rel_err(vol, cv) < 1e-14


def test_reference_neighbor_counts(gpu_available):
    # NeighborSearchProcessorTest.cpp:160-207 on the unit-spacing generated meshes
    X, t = _mesh(5, jit=0, spacing=True)
    with _gpu(X, t, 1.1) as s:
        num, _, _ = s.rk_neighbors()
        assert num.min() == 8 and num.max() == 27 and abs(num.mean() - 18.713) < 1e-3
    X, hc = M.hex_lattice(1, 1, 4, 1, 1, 4)
    t = M.split_tets(X, hc)
    with _gpu(X, t, 1.01) as s:
        num, _, _ = s.rk_neighbors()
        assert sorted(Counter(num.tolist()).items()) == [(7, 4), (8, 4), (10, 6), (11, 2), (12, 4)]


def test_rk_mass_and_force(gpu_available):
    X, t = _mesh(6, seed=5)
    scale = 1.5
    r, num, nbr, tr, phi = _oracle_rk(X, t, scale)
    cells = O.rk_cells(X, t, num, nbr, phi)
    mass = O.rk_mass(t, cells[3], J2["rho"], num, nbr, phi)
    lam, mu = O.lame(J2["E"], J2["nu"])
    u = 4e-3 * (2 * M.uniform01(8, 7, X.size).reshape(X.shape) - 1)
    f_ref, H_ref, P_ref, s_ref = O.cell_internal_force(len(X), cells, O.J2, lam, mu, u,
                                                        yield_stress=J2["yield_stress"], hardening=J2["hardening"])
    with _gpu(X, t, scale) as s:
        tot = s.compute_mass()
        assert rel_err(s.get(A.FIELD_MASS), mass) < 1e-13
        assert abs(tot - mass.sum()) < 1e-12 * mass.sum()
        s.set_field(A.FIELD_U, u)
        s.internal_force()
        assert s_ref.max() > 0
        assert rel_err(s.get(A.FIELD_F), f_ref) < 1e-12
        assert rel_err(s.get(A.FIELD_H), H_ref) < 1e-12
        assert rel_err(s.get(A.FIELD_P), P_ref) < 1e-12
        assert rel_err(s.get(A.FIELD_STATE), s_ref) < 1e-12


def test_rk_j2_100_steps(gpu_available):
    X, t = _mesh(6, seed=9)
    scale = 1.5
    r, num, nbr, tr, phi = _oracle_rk(X, t, scale)
    cells = O.rk_cells(X, t, num, nbr, phi)
    mass = O.rk_mass(t, cells[3], J2["rho"], num, nbr, phi)
    lam, mu = O.lame(J2["E"], J2["nu"])
    c = X - 0.5
    v0 = 60.0 * np.stack([-0.5 * c[:, 0], -0.5 * c[:, 1], c[:, 2]], axis=1)
    dt = 0.5 * M.stable_dt_estimate(X, t, J2["E"], J2["nu"], J2["rho"])
    ref = O.Solver(-1, X, None, O.J2, lam, mu, mass, None, J2["yield_stress"], J2["hardening"], cells=cells)
    ref.v[:] = v0
    ref.run(dt, 1)
    with _gpu(X, t, scale) as s:
        s.compute_mass()
        s.set_field(A.FIELD_V, v0)
        s.initialize()
        s.step(dt, 1)
        for fld, rr in ((A.FIELD_U, ref.u), (A.FIELD_V, ref.v), (A.FIELD_F, ref.f), (A.FIELD_P, ref.P)):
            assert rel_err(s.get(fld), rr) < 1e-12, fld
        s.step(dt, 99)
        ref.run(dt, 99)
        assert ref.state.max() > 0.0  # the ramp drives yielding
        for fld, rr in ((A.FIELD_U, ref.u), (A.FIELD_V, ref.v), (A.FIELD_A, ref.a), (A.FIELD_F, ref.f),
                        (A.FIELD_H, ref.H), (A.FIELD_P, ref.P)):
            assert rel_err(s.get(fld), rr) < 1e-10, fld
        assert rel_err(s.get(A.FIELD_STATE), ref.state) < 1e-10


# ---- two-pass method ---------------------------------------------------------

NH = dict(E=1e9, nu=0.3, rho=1000.0)


def _gpu2(X, t, scale, kind=A.J2, p=J2):
    s = A.Solver()
    s.set_mesh(A.TET4, X, t)
    s.set_material(kind, p["rho"], p["E"], p["nu"], p.get("yield_stress", 0.0), p.get("hardening", 0.0))
    s.rk_setup(scale, one_pass=False)
    return s


@pytest.mark.parametrize("kind,p", [(A.J2, J2), (A.NEO_HOOKEAN, NH)])
def test_two_pass_cells_mass_and_force(gpu_available, kind, p):
    X, t = _mesh(6, seed=5)
    scale = 1.5
    r, num, nbr, tr, phi = _oracle_rk(X, t, scale)
    cells2 = O.rk_cells_two_pass(X, t)
    mass = O.rk_mass(t, cells2[3], p["rho"], num, nbr, phi)
    lam, mu = O.lame(p["E"], p["nu"])
    u = 4e-3 * (2 * M.uniform01(8, 7, X.size).reshape(X.shape) - 1)
    kw = dict(yield_stress=p.get("yield_stress", 0.0), hardening=p.get("hardening", 0.0))
    f_ref, H_ref, P_ref, s_ref, up_ref, fl_ref = O.rk_two_pass_force(len(X), cells2, num, nbr, phi, kind, lam, mu,
                                                                      u, **kw)
    print(f"rk two-pass kind={kind} nodes={len(X)} cells={len(t)}", flush=True)
    with _gpu2(X, t, scale, kind, p) as s:
        start, nodes, dphi, vol = s.rk_cells()
        assert np.array_equal(start, cells2[0]) and np.array_equal(nodes, cells2[1])
        assert rel_err(dphi, cells2[2]) < 1e-12 and rel_err(vol, cells2[3]) < 1e-14
        s.compute_mass()
        assert rel_err(s.get(A.FIELD_MASS), mass) < 1e-13
        s.set_field(A.FIELD_U, u)
        s.internal_force()
        assert rel_err(s.get(A.FIELD_RK_DISPLACEMENT), up_ref) < 1e-12
        assert rel_err(s.get(A.FIELD_RK_FORCE_LOCAL), fl_ref) < 1e-12
        assert rel_err(s.get(A.FIELD_F), f_ref) < 1e-12
        assert rel_err(s.get(A.FIELD_H), H_ref) < 1e-12
        assert rel_err(s.get(A.FIELD_P), P_ref) < 1e-12
        if kind == A.J2:
            assert s_ref.max() > 0
            assert rel_err(s.get(A.FIELD_STATE), s_ref) < 1e-12
        f2 = s.get(A.FIELD_F)
    # the one-pass method on the GPU gives the same force up to rounding
    with _gpu(X, t, scale, kind, p) as s:
        s.set_field(A.FIELD_U, u)
        s.internal_force()
        assert rel_err(f2, s.get(A.FIELD_F)) < 1e-12


def test_two_pass_small_radius_is_the_tet4_element(gpu_available):
    # ElementTest.cpp:154-162 (ReproducingKernelOnTet4TwoPass): a kernel radius scale of
    # 0.03 leaves every node its own only neighbor, so the force is the smoothed Tet4's
    X, t = _mesh(5, seed=2)
    lam, mu = O.lame(NH["E"], NH["nu"])
    u = 2e-3 * (2 * M.uniform01(3, 1, X.size).reshape(X.shape) - 1)
    f_fe = O.internal_force(O.TET4, X, t, O.NEO_HOOKEAN, lam, mu, u)[0]
    with _gpu2(X, t, 0.03, A.NEO_HOOKEAN, NH) as s:
        num, _, _ = s.rk_neighbors()
        assert (num == 1).all() and (s.rk_function_values()[:, 0] == 1.0).all()
        s.set_field(A.FIELD_U, u)
        s.internal_force()
        assert np.array_equal(s.get(A.FIELD_RK_DISPLACEMENT), u)
        assert rel_err(s.get(A.FIELD_F), f_fe) < 1e-12


def test_two_pass_j2_100_steps(gpu_available):
    X, t = _mesh(6, seed=9)
    scale = 1.5
    r, num, nbr, tr, phi = _oracle_rk(X, t, scale)
    cells2 = O.rk_cells_two_pass(X, t)
    mass = O.rk_mass(t, cells2[3], J2["rho"], num, nbr, phi)
    lam, mu = O.lame(J2["E"], J2["nu"])
    c = X - 0.5
    v0 = 60.0 * np.stack([-0.5 * c[:, 0], -0.5 * c[:, 1], c[:, 2]], axis=1)
    dt = 0.5 * M.stable_dt_estimate(X, t, J2["E"], J2["nu"], J2["rho"])
    ref = O.Solver(-1, X, None, O.J2, lam, mu, mass, None, J2["yield_stress"], J2["hardening"], cells=cells2,
                   two_pass=(num, nbr, phi))
    ref.v[:] = v0
    ref.run(dt, 1)
    with _gpu2(X, t, scale) as s:
        s.compute_mass()
        s.set_field(A.FIELD_V, v0)
        s.initialize()
        s.step(dt, 1)
        for fld, rr in ((A.FIELD_U, ref.u), (A.FIELD_V, ref.v), (A.FIELD_F, ref.f), (A.FIELD_P, ref.P)):
            assert rel_err(s.get(fld), rr) < 1e-12, fld
        s.step(dt, 99)
        ref.run(dt, 99)
        assert ref.state.max() > 0.0
        for fld, rr in ((A.FIELD_U, ref.u), (A.FIELD_V, ref.v), (A.FIELD_A, ref.a), (A.FIELD_F, ref.f),
                        (A.FIELD_H, ref.H), (A.FIELD_P, ref.P)):
            assert rel_err(s.get(fld), rr) < 1e-10, fld
        assert rel_err(s.get(A.FIELD_STATE), ref.state) < 1e-10


def test_two_pass_graph_replay_and_power_method(gpu_available):
    X, t = _mesh(5, seed=4)
    scale = 1.5
    lam, mu = O.lame(NH["E"], NH["nu"])
    dt = 0.3 * M.stable_dt_estimate(X, t, NH["E"], NH["nu"], NH["rho"])
    v0 = 0.5 * (2 * M.uniform01(6, 2, X.size).reshape(X.shape) - 1)
    out = []
    for graph in (False, True):
        with _gpu2(X, t, scale, A.NEO_HOOKEAN, NH) as s:
            s.compute_mass()
            s.set_graph(graph)
            s.set_field(A.FIELD_V, v0)
            s.initialize()
            for _ in range(3):
                s.step(dt, 10)
                out.append(s.get(A.FIELD_U))
    for k in range(3):
        assert np.array_equal(out[k], out[3 + k])
    # the power-method increment of the two-pass operator equals the one-pass one
    res = []
    for one in (True, False):
        with A.Solver() as s:
            s.set_mesh(A.TET4, X, t)
            s.set_material(A.NEO_HOOKEAN, NH["rho"], NH["E"], NH["nu"])
            s.rk_setup(scale, one_pass=one)
            s.compute_mass()
            res.append(s.power_method_time_increment(num_iterations=50))
    assert res[0][1] == res[1][1] and abs(res[0][0] - res[1][0]) < 1e-9 * res[0][0]


# ---- Hex8 background mesh (ElementReproducingKernelHex8) ------------------------

def _gpu_hex(X, hc, scale, one_pass):
    s = A.Solver()
    s.set_mesh(A.HEX8, X, hc)
    s.set_material(A.J2, J2["rho"], J2["E"], J2["nu"], J2["yield_stress"], J2["hardening"])
    s.rk_setup(scale, one_pass=one_pass)
    return s


@pytest.mark.parametrize("one_pass", [True, False], ids=["one_pass", "two_pass"])
def test_hex8_background_setup_force_and_100_steps(gpu_available, one_pass):
    """One smoothed cell per hex (element_smoothing_cell on a hex mesh): neighbors
    bit-exact, RK functions, cell CSR, mass and the J2 force of a random field at
    1e-12, then 100 yielding steps at 1e-10, one- and two-pass."""
    n = 6
    X, hc = M.hex_lattice(n, n, n)
    X = M.jitter(X, 1.0 / n, 0.1, 13, 1)
    scale = 1.5
    r, num, nbr, tr, phi = _oracle_rk(X, hc, scale)
    cells = O.rk_cells(X, hc, num, nbr, phi) if one_pass else O.rk_cells_two_pass(X, hc)
    mass = O.rk_mass(hc, cells[3], J2["rho"], num, nbr, phi)
    lam, mu = O.lame(J2["E"], J2["nu"])
    kw = dict(yield_stress=J2["yield_stress"], hardening=J2["hardening"])
    u = 4e-3 * (2 * M.uniform01(8, 7, X.size).reshape(X.shape) - 1)
    if one_pass:
        f_ref, H_ref, P_ref, s_ref = O.cell_internal_force(len(X), cells, O.J2, lam, mu, u, **kw)
    else:
        f_ref, H_ref, P_ref, s_ref, _, _ = O.rk_two_pass_force(len(X), cells, num, nbr, phi, O.J2, lam, mu, u, **kw)
    print(f"rk hex8 background one_pass={one_pass} nodes={len(X)} cells={len(hc)} trunc={tr}", flush=True)
    with _gpu_hex(X, hc, scale, one_pass) as s:
        gnum, gnbr, gtr = s.rk_neighbors()
        assert np.array_equal(gnum, num) and gtr == tr
        assert np.array_equal(gnbr, np.where(np.arange(40)[None, :] < num[:, None], nbr, -1))
        assert np.abs(s.rk_function_values() - phi).max() < 1e-12
        start, nodes, dphi, vol = s.rk_cells()
        assert np.array_equal(start, cells[0]) and np.array_equal(nodes, cells[1])
        assert rel_err(dphi, cells[2]) < 1e-12 and rel_err(vol, cells[3]) < 1e-14
        s.compute_mass()
        assert rel_err(s.get(A.FIELD_MASS), mass) < 1e-13
        s.set_field(A.FIELD_U, u)
        s.internal_force()
        assert s_ref.max() > 0
        for fld, rr in ((A.FIELD_F, f_ref), (A.FIELD_H, H_ref), (A.FIELD_P, P_ref), (A.FIELD_STATE, s_ref)):
            assert rel_err(s.get(fld), rr) < 1e-12, fld
    c = X - 0.5
    v0 = 60.0 * np.stack([-0.5 * c[:, 0], -0.5 * c[:, 1], c[:, 2]], axis=1)
    dt = 0.5 * M.stable_dt_estimate(X, hc, J2["E"], J2["nu"], J2["rho"])
    ref = O.Solver(-1, X, None, O.J2, lam, mu, mass, None, J2["yield_stress"], J2["hardening"], cells=cells,
                   two_pass=None if one_pass else (num, nbr, phi))
    ref.v[:] = v0
    ref.run(dt, 100)
    assert ref.state.max() > 0.0
    with _gpu_hex(X, hc, scale, one_pass) as s:
        s.compute_mass()
        s.set_field(A.FIELD_V, v0)
        s.initialize()
        s.step(dt, 100)
        for fld, rr in ((A.FIELD_U, ref.u), (A.FIELD_V, ref.v), (A.FIELD_A, ref.a), (A.FIELD_F, ref.f),
                        (A.FIELD_H, ref.H), (A.FIELD_P, ref.P)):
            assert rel_err(s.get(fld), rr) < 1e-10, fld
        assert rel_err(s.get(A.FIELD_STATE), ref.state) < 1e-10


# ---- finite-element strain smoothing (ElementSmoothedTetrahedron4) ---------------

@pytest.mark.parametrize("kind,p", [(A.J2, J2), (A.NEO_HOOKEAN, NH)])
def test_fe_strain_smoothing_tet4(gpu_available, kind, p):
    """One smoothed cell per Tet4 with the element's own shape functions: the cells are
    the element-node cells, the mass the lumped rho V / 4, the force at a random field
    the oracle's cell force (1e-12); over 100 steps it follows the oracle (1e-10) and,
    the smoothed B of a linear tet being its exact gradient, the Gauss Tet4 step (1e-9)."""
    X, t = _mesh(6, seed=5)
    cells = O.rk_cells_two_pass(X, t)
    mass = O.lumped_mass(A.TET4, X, t, p["rho"])
    lam, mu = O.lame(p["E"], p["nu"])
    kw = dict(yield_stress=p.get("yield_stress", 0.0), hardening=p.get("hardening", 0.0))
    u = 4e-3 * (2 * M.uniform01(8, 7, X.size).reshape(X.shape) - 1)
    f_ref, H_ref, P_ref, s_ref = O.cell_internal_force(len(X), cells, kind, lam, mu, u, **kw)
    print(f"fe strain smoothing kind={kind} nodes={len(X)} cells={len(t)}", flush=True)

    def make():
        s = A.Solver()
        s.set_mesh(A.TET4, X, t)
        s.set_material(kind, p["rho"], p["E"], p["nu"], kw["yield_stress"], kw["hardening"])
        s.strain_smoothing_setup()
        return s

    if kind == A.J2:
        assert s_ref.max() > 0  # the random field yields some cells
    with make() as s:
        num, nbr, tr = s.rk_neighbors()
        assert (num == 1).all() and np.array_equal(nbr[:, 0], np.arange(len(X))) and tr == 0
        assert (nbr[:, 1:] == -1).all()
        ident = np.zeros((len(X), 40))
        ident[:, 0] = 1.0
        assert np.array_equal(s.rk_function_values(), ident)
        start, nodes, dphi, vol = s.rk_cells()
        assert np.array_equal(start, cells[0]) and np.array_equal(nodes, cells[1])
        assert rel_err(dphi, cells[2]) < 1e-12 and rel_err(vol, cells[3]) < 1e-14
        s.compute_mass()
        assert rel_err(s.get(A.FIELD_MASS), mass) < 1e-13
        s.set_field(A.FIELD_U, u)
        s.internal_force()
        fields = [(A.FIELD_F, f_ref), (A.FIELD_H, H_ref), (A.FIELD_P, P_ref)]
        if kind == A.J2:
            fields.append((A.FIELD_STATE, s_ref))
        for fld, rr in fields:
            assert rel_err(s.get(fld), rr) < 1e-12, fld
    c = X - 0.5
    v0 = 60.0 * np.stack([-0.5 * c[:, 0], -0.5 * c[:, 1], c[:, 2]], axis=1)
    dt = 0.5 * M.stable_dt_estimate(X, t, p["E"], p["nu"], p["rho"])
    ref = O.Solver(-1, X, None, kind, lam, mu, mass, None, kw["yield_stress"], kw["hardening"], cells=cells)
    ref.v[:] = v0
    ref.run(dt, 100)
    with make() as s, A.Solver() as g:
        s.compute_mass()
        s.set_field(A.FIELD_V, v0)
        s.initialize()
        s.step(dt, 100)
        for fld, rr in ((A.FIELD_U, ref.u), (A.FIELD_V, ref.v), (A.FIELD_F, ref.f), (A.FIELD_H, ref.H),
                        (A.FIELD_P, ref.P)):
            assert rel_err(s.get(fld), rr) < 1e-10, fld
        if kind == A.J2:
            assert ref.state.max() > 0.0
            assert rel_err(s.get(A.FIELD_STATE), ref.state) < 1e-10
        g.set_mesh(A.TET4, X, t)
        g.set_material(kind, p["rho"], p["E"], p["nu"], kw["yield_stress"], kw["hardening"])
        g.compute_mass()
        g.set_field(A.FIELD_V, v0)
        g.initialize()
        g.step(dt, 100)
        assert rel_err(s.get(A.FIELD_U), g.get(A.FIELD_U)) < 1e-9


@pytest.mark.parametrize("tile_cells", ["1", "7", "64", "300"])
def test_one_pass_tiles_are_deterministic_and_size_independent(gpu_available, monkeypatch, tile_cells):
    """The one-pass force runs on cell tiles (runs of consecutive cells with a shared node
    set, forces added in LDS by one wave, slots summed per node in ascending tile order).
    Two runs give the same bits; any tile size (APHIP_RK_TILE_CELLS, read at set-up)
    matches the oracle after 20 J2 steps at 1e-10."""
    X, t = _mesh(6, seed=9)
    scale = 1.5
    r, num, nbr, tr, phi = _oracle_rk(X, t, scale)
    cells = O.rk_cells(X, t, num, nbr, phi)
    mass = O.rk_mass(t, cells[3], J2["rho"], num, nbr, phi)
    lam, mu = O.lame(J2["E"], J2["nu"])
    dt = 0.4 * M.stable_dt_estimate(X, t, J2["E"], J2["nu"], J2["rho"])
    v0 = 30.0 * (2 * M.uniform01(12, 2, X.size).reshape(X.shape) - 1)
    ref = O.Solver(-1, X, None, O.J2, lam, mu, mass, None, J2["yield_stress"], J2["hardening"], cells=cells)
    ref.v[:] = v0
    ref.run(dt, 20)
    monkeypatch.setenv("APHIP_RK_TILE_CELLS", tile_cells)
    print(f"one-pass tiles of {tile_cells} cells", flush=True)
    out = []
    for _ in range(2):
        with _gpu(X, t, scale) as s:
            s.compute_mass()
            s.set_field(A.FIELD_V, v0)
            s.initialize()
            s.step(dt, 20)
            out.append([s.get(f) for f in (A.FIELD_U, A.FIELD_V, A.FIELD_F, A.FIELD_P, A.FIELD_STATE)])
    for a, b in zip(*out):
        assert np.array_equal(a, b)
    u, v, f, P, st = out[0]
    assert ref.state.max() > 0.0
    for g, rf in ((u, ref.u), (v, ref.v), (f, ref.f), (P, ref.P), (st, ref.state)):
        assert rel_err(g, rf) < 1e-10
