"""GPU parity at the full size of each benchmark workload, through the C ABI.

Each test runs one named workload at its real size and checks the HIP path
against the oracle on the same seeded inputs:

  plain Hex8 cube      20^3 Hex8 (8,000 el, 9,261 nodes), neo-Hookean, z = 0 face fixed,
                       dt = 0.5 x the power-method increment, 100 steps
  large Tet4 block     100^3 hexes x 6 Kuhn tets (6,000,000 Tet4, 1,030,301 nodes),
                       linear elastic and neo-Hookean, 100 steps
  RKPM lattice         64^3 nodes (262,144), 1,500,282 Tet4 cells, J2, kernel scale 1.5,
                       one-pass and two-pass force, 100 steps of the velocity ramp
  Hex8 ensemble        256 cubes with 8..48 hexes per edge, neo-Hookean, per-specimen
                       E, nu and dt, 50 steps: bitwise equal to the per-specimen loop,
                       and the smallest, median and largest specimens against the oracle
  adjoint block        64^3 Hex8 with per-element E / nu fields, 200 steps with checkpoints
                       every 20: forward against the oracle, gradient against central
                       differences

Bars (force-step contract): every field within 1e-12 relative (max-abs error over the
field's max-abs) after the first step, within 1e-10 after 100 steps (200 for the adjoint
block); neighbor lists identical (same sets, same order); gradients within 1e-6 of
central differences.
"""
import os

import numpy as np
import pytest

import oracle as O
import aperi_hip as A
from aperi_hip import meshgen as M
from conftest import rel_err

pytestmark = pytest.mark.gpu

NT = max(1, min(16, int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)))


def _say(msg):
    print(msg, flush=True)


def _compare(s, ref, tol, with_state=False):
    pairs = [(A.FIELD_U, ref.u), (A.FIELD_V, ref.v), (A.FIELD_A, ref.a), (A.FIELD_F, ref.f),
             (A.FIELD_H, ref.H), (A.FIELD_P, ref.P)]
    if with_state:
        pairs.append((A.FIELD_STATE, ref.state))
    errs = {}
    for fld, r in pairs:
        errs[fld] = rel_err(s.get(fld), r)
    _say(f"  max rel err per field: { {k: f'{v:.2e}' for k, v in errs.items()} }")
    for fld, e in errs.items():
        assert e < tol, (fld, e)


# ---- plain Hex8 cube ---------------------------------------------------------------

def test_plain_hex8_cube_20(gpu_available):
    w = M.plain_hex_cube(20)
    X, conn, fixed, v0 = w["X"], w["conn"], w["fixed"], w["v0"]
    assert conn.shape == (8000, 8) and X.shape == (9261, 3)
    lam, mu = O.lame(w["E"], w["nu"])
    mass = O.lumped_mass(A.HEX8, X, conn, w["rho"])
    ref = O.Solver(A.HEX8, X, conn, O.NEO_HOOKEAN, lam, mu, mass, fixed)
    _say("plain Hex8 cube 20^3: power-method increment, 100 neo-Hookean steps")
    with A.Solver() as s:
        s.set_mesh(A.HEX8, X, conn)
        s.set_material(A.NEO_HOOKEAN, w["rho"], w["E"], w["nu"])
        s.compute_mass()
        assert rel_err(s.get(A.FIELD_MASS), mass) < 1e-13
        s.set_fixed(fixed)
        # the stable increment of the reference's power method, same start vector both sides
        dt_ref, it_ref, cv_ref = ref.power_method(O.max_edge_length(X, conn), u_in=np.zeros_like(X), nthreads=NT)
        dt_cfl, it, cv = s.power_method_time_increment()
        assert (it, cv) == (it_ref, cv_ref)
        assert dt_cfl == pytest.approx(dt_ref, rel=1e-9)
        dt = 0.5 * dt_ref
        ref.v[:] = v0
        s.set_field(A.FIELD_V, v0)
        s.initialize()
        s.step(dt, 1)
        ref.run(dt, 1, nthreads=NT)
        _compare(s, ref, 1e-12)
        s.step(dt, 99)
        ref.run(dt, 99, nthreads=NT)
        _compare(s, ref, 1e-10)
        assert np.all(s.get(A.FIELD_V)[fixed.astype(bool)] == 0.0)
        assert np.abs(ref.u).max() > 0.0


# ---- large Tet4 block --------------------------------------------------------------

@pytest.mark.parametrize("mat", ["le", "nh"])
def test_large_tet4_block_6M(gpu_available, mat):
    kind = {"le": A.LINEAR_ELASTIC, "nh": A.NEO_HOOKEAN}[mat]
    w = M.large_tet_block(100)
    X, conn, v0 = w["X"], w["conn"], w["v0"]
    assert conn.shape == (6000000, 4) and X.shape == (1030301, 3)
    lam, mu = O.lame(w["E"], w["nu"])
    mass = O.lumped_mass(A.TET4, X, conn, w["rho"])
    dt = 0.5 * M.stable_dt_estimate(X, conn, w["E"], w["nu"], w["rho"])
    _say(f"large Tet4 block 6M {mat}: 1 step at 1e-12, 100 steps at 1e-10 (oracle on {NT} threads)")
    ref = O.Solver(A.TET4, X, conn, {"le": O.LINEAR_ELASTIC, "nh": O.NEO_HOOKEAN}[mat], lam, mu, mass)
    ref.v[:] = v0
    with A.Solver() as s:
        s.set_mesh(A.TET4, X, conn)
        s.set_material(kind, w["rho"], w["E"], w["nu"])
        s.compute_mass()
        assert rel_err(s.get(A.FIELD_MASS), mass) < 1e-13
        s.set_field(A.FIELD_V, v0)
        s.initialize()
        s.step(dt, 1)
        ref.run(dt, 1, nthreads=NT)
        _compare(s, ref, 1e-12)
        s.step(dt, 99)
        ref.run(dt, 99, nthreads=NT)
        _compare(s, ref, 1e-10)


# ---- RKPM lattice ------------------------------------------------------------------

@pytest.mark.parametrize("one_pass", [True, False], ids=["one_pass", "two_pass"])
def test_rkpm_lattice_64(gpu_available, one_pass):
    w = M.rkpm_lattice(64)
    X, conn = w["X"], w["conn"]
    assert X.shape == (262144, 3) and conn.shape == (1500282, 4)
    scale = 1.5
    _say(f"RKPM lattice 64^3 one_pass={one_pass}: set-up against the oracle (threads {NT})")
    r = O.kernel_radius(O.max_edge_length(X, conn), scale)
    num, nbr, trunc = O.neighbor_search(X, r)
    phi = O.rk_function_values(X, r, num, nbr)
    cells = O.rk_cells(X, conn, num, nbr, phi) if one_pass else O.rk_cells_two_pass(X, conn)
    mass = O.rk_mass(conn, cells[3], w["rho"], num, nbr, phi)
    lam, mu = O.lame(w["E"], w["nu"])
    steps = 100
    dt = 0.5 * M.stable_dt_estimate(X, conn, w["E"], w["nu"], w["rho"])
    vstar = 2.0 * w["yield_stress"] / w["E"] / (steps * dt)
    v0 = M.rkpm_velocity(X, vstar, w["ramp_scale"])
    with A.Solver() as s:
        s.set_mesh(A.TET4, X, conn)
        s.set_material(A.J2, w["rho"], w["E"], w["nu"], w["yield_stress"], w["hardening"])
        s.rk_setup(scale, one_pass=one_pass)
        gnum, gnbr, gtr = s.rk_neighbors()
        # identical lists: same counts, same entries in the same order, same truncation count
        assert np.array_equal(gnum, num) and gtr == trunc
        valid = np.arange(40)[None, :] < num[:, None]
        assert np.array_equal(gnbr, np.where(valid, nbr, -1))
        assert np.array_equal(np.sort(np.where(valid, gnbr, -1), axis=1), np.sort(np.where(valid, nbr, -1), axis=1))
        assert num.min() >= 4 and num.max() == 40 and trunc > 0
        assert np.abs(s.rk_function_values() - phi).max() < 1e-12
        start, nodes, dphi, cvol = s.rk_cells()
        assert np.array_equal(start, cells[0]) and np.array_equal(nodes, cells[1])
        assert rel_err(dphi, cells[2]) < 1e-12 and rel_err(cvol, cells[3]) < 1e-12
        del start, nodes, dphi, cvol
        s.compute_mass()
        assert rel_err(s.get(A.FIELD_MASS), mass) < 1e-13
        _say(f"  set-up identical (mean neighbors {num.mean():.2f}, {trunc} extra candidates dropped); "
             f"{steps} J2 steps")
        ref = O.Solver(-1, X, None, O.J2, lam, mu, mass, None, w["yield_stress"], w["hardening"], cells=cells,
                       two_pass=None if one_pass else (num, nbr, phi))
        ref.v[:] = v0
        s.set_field(A.FIELD_V, v0)
        s.initialize()
        s.step(dt, 1)
        ref.run(dt, 1, nthreads=NT)
        _compare(s, ref, 1e-12, with_state=True)
        s.step(dt, steps - 1)
        ref.run(dt, steps - 1, nthreads=NT)
        assert ref.state.max() > 0.0  # the ramp has driven cells past yield
        _compare(s, ref, 1e-10, with_state=True)


# ---- Hex8 ensemble ----------------------------------------------------------------

def test_hex8_ensemble_256(gpu_available):
    steps = 50
    specs = M.ensemble(256)
    dts = [0.5 * M.stable_dt_estimate(sp["X"], sp["conn"], sp["E"], sp["nu"], sp["rho"]) for sp in specs]
    n_el = np.array([len(sp["conn"]) for sp in specs])
    assert n_el.min() == 8 ** 3 and n_el.max() == 48 ** 3
    _say(f"Hex8 ensemble: 256 specimens, {n_el.sum()} Hex8, {steps} graph-replayed steps")
    fields = (A.FIELD_U, A.FIELD_V, A.FIELD_A, A.FIELD_F, A.FIELD_H, A.FIELD_P)
    with A.Solver() as g:
        g.batch_begin(A.HEX8, A.NEO_HOOKEAN)
        for sp, dt in zip(specs, dts):
            g.batch_add(sp["X"], sp["conn"], sp["rho"], sp["E"], sp["nu"], dt)
        g.batch_end()
        g.compute_mass()
        g.set_field(A.FIELD_V, np.concatenate([sp["v0"] for sp in specs]))
        g.initialize()
        g.set_graph(True)
        g.step(0.0, steps)
        got = {f: g.get(f) for f in fields}
        ranges = [g.batch_ranges(k) for k in range(len(specs))]
    # every specimen bitwise equal to its own solver run alone
    for k, (sp, dt) in enumerate(zip(specs, dts)):
        n0, nn, e0, ne = ranges[k]
        with A.Solver() as s:
            s.set_mesh(A.HEX8, sp["X"], sp["conn"])
            s.set_material(A.NEO_HOOKEAN, sp["rho"], sp["E"], sp["nu"])
            s.compute_mass()
            s.set_field(A.FIELD_V, sp["v0"])
            s.initialize()
            s.step(dt, steps)
            for f in fields:
                lo, hi = (e0, e0 + ne) if f in (A.FIELD_H, A.FIELD_P) else (n0, n0 + nn)
                assert np.array_equal(got[f][lo:hi], s.get(f)), (k, f)
    # the smallest, median and largest specimens against the oracle
    order = np.argsort(n_el, kind="stable")
    for k in (order[0], order[len(order) // 2], order[-1]):
        sp, dt = specs[k], dts[k]
        n0, nn, e0, ne = ranges[k]
        lam, mu = O.lame(sp["E"], sp["nu"])
        mass = O.lumped_mass(A.HEX8, sp["X"], sp["conn"], sp["rho"])
        ref = O.Solver(A.HEX8, sp["X"], sp["conn"], O.NEO_HOOKEAN, lam, mu, mass)
        ref.v[:] = sp["v0"]
        ref.run(dt, steps, nthreads=NT)
        _say(f"  specimen {k}: {sp['n']}^3 Hex8 against the oracle")
        for f, r in ((A.FIELD_U, ref.u), (A.FIELD_V, ref.v), (A.FIELD_A, ref.a), (A.FIELD_F, ref.f),
                     (A.FIELD_H, ref.H), (A.FIELD_P, ref.P)):
            lo, hi = (e0, e0 + ne) if f in (A.FIELD_H, A.FIELD_P) else (n0, n0 + nn)
            assert rel_err(got[f][lo:hi], r) < 1e-10, (k, f)


# ---- adjoint block ----------------------------------------------------------------

def _fd_step(kind, theta, scale):
    """Central-difference step per parameter kind: E_e and nu_e, never near zero, take
    1e-6 of their own value.  An initial-velocity component can sit arbitrarily close to
    zero, so its step is taken relative to the velocity field's amplitude, and it is 1e-5
    of that: the smallest velocity sensitivities of this loss (about 1e-12, four decades
    below the median) are otherwise at the rounding floor of two 200-step FP64 forwards
    (1.1e-6 relative measured with 1e-6); the response to a 1e-5 m/s change of one
    velocity component is linear far beyond that step, so truncation stays negligible."""
    return 1e-6 * abs(theta) if kind in ("E", "nu") else 1e-5 * scale


def test_adjoint_block_64(gpu_available):
    steps, interval = 200, 20
    w = M.adjoint_block(64)
    X, conn, rho = w["X"], w["conn"], w["rho"]
    E, nu, v0 = w["E"], w["nu"], w["v0"]
    assert conn.shape == (262144, 8)
    dt = 0.5 * M.stable_dt_estimate(X, conn, float(E.max()), float(nu.min()), rho)
    lam, mu = M.lame(E, nu)
    mass = O.lumped_mass(A.HEX8, X, conn, rho)
    _say(f"adjoint block 64^3 Hex8: forward against the oracle over {steps} steps (threads {NT})")
    ref = O.Solver(A.HEX8, X, conn, O.NEO_HOOKEAN, 0.0, 0.0, mass, elem_lambda=lam, elem_mu=mu)
    ref.v[:] = v0
    with A.Solver() as s:
        s.set_mesh(A.HEX8, X, conn)
        s.set_material(A.NEO_HOOKEAN, rho, 200e9, 0.3)
        s.set_material_fields(E, nu)
        s.compute_mass()
        assert rel_err(s.get(A.FIELD_MASS), mass) < 1e-13

        def forward(E_, nu_, v_, checkpoints=False):
            s.set_material_fields(E_, nu_)
            s.set_field(A.FIELD_U, np.zeros_like(X))
            s.set_field(A.FIELD_V, v_)
            s.initialize()
            if checkpoints:
                s.adjoint_forward(dt, steps, interval)
            else:
                s.step(dt, steps)
            return s.get(A.FIELD_U)

        u_star = forward(w["E_target"], nu, v0)
        # the checkpointed forward is the plain forward, and both follow the oracle
        s.set_material_fields(E, nu)
        s.set_field(A.FIELD_U, np.zeros_like(X))
        s.set_field(A.FIELD_V, v0)
        s.initialize()
        s.step(dt, 1)
        ref.run(dt, 1, nthreads=NT)
        _compare(s, ref, 1e-12)
        s.step(dt, steps - 1)
        ref.run(dt, steps - 1, nthreads=NT)
        _compare(s, ref, 1e-10)
        u_plain = s.get(A.FIELD_U)
        u_T = forward(E, nu, v0, checkpoints=True)
        assert np.array_equal(u_T, u_plain)
        Eb, nub, v0b = s.adjoint_backward(u_T - u_star)
        # 32 parameters: 12 E, 10 nu, 10 initial-velocity components, central differences
        # with L+ - L- formed as (u+ - u-) . ((u+ + u-)/2 - u*), free of the loss's cancellation
        rng = np.random.default_rng(20261202)
        picks = ([("E", int(i)) for i in rng.choice(len(conn), 12, replace=False)] +
                 [("nu", int(i)) for i in rng.choice(len(conn), 10, replace=False)] +
                 [("v0", int(i)) for i in rng.choice(X.size, 10, replace=False)])
        vscale = float(np.abs(v0).max())
        worst = 0.0
        for kind, i in picks:
            Ec, nuc, vc = E.copy(), nu.copy(), v0.copy()
            arr = {"E": Ec, "nu": nuc, "v0": vc.reshape(-1)}[kind]
            h = _fd_step(kind, arr[i], vscale)
            base = arr[i]
            arr[i] = base + h
            up = forward(Ec, nuc, vc)
            arr[i] = base - h
            um = forward(Ec, nuc, vc)
            fd = np.sum((up - um) * (0.5 * (up + um) - u_star)) / (2.0 * h)
            ad = {"E": Eb, "nu": nub, "v0": v0b.reshape(-1)}[kind][i]
            err = abs(ad - fd) / abs(fd)
            worst = max(worst, err)
            assert err < 1e-6, (kind, i, ad, fd, err)
        _say(f"  32 central differences: max rel err {worst:.2e}")
