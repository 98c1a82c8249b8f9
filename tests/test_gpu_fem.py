"""GPU parity of the Gaussian element force step against the CPU oracle.

Bar (force step contract): every field within 1e-12 relative (max-abs error over
the field's max-abs) after one force evaluation / one step, and within 1e-10
after 100 central-difference steps.  All through the C ABI (libaperi_hip.so).
"""
import numpy as np
import pytest

import oracle as O
import aperi_hip as A
from aperi_hip import meshgen as M
from conftest import rel_err

pytestmark = pytest.mark.gpu

TOL_STEP = 1e-12
TOL_100 = 1e-10

MATS = {
    "le": (A.LINEAR_ELASTIC, dict(E=200e9, nu=0.3, rho=7850.0)),
    "nh": (A.NEO_HOOKEAN, dict(E=200e9, nu=0.3, rho=7850.0)),
    "j2": (A.J2, dict(E=70e9, nu=0.33, rho=2700.0, yield_stress=300e6, hardening=1e9)),
    "svk": (A.SVK, dict(E=200e9, nu=0.3, rho=7850.0)),
}


def _mesh(etype, n, seed):
    X, hc = M.hex_lattice(n, n, n)
    X = M.jitter(X, 1.0 / n, 0.1, seed, 1)
    conn = hc if etype == A.HEX8 else M.split_tets(X, hc)
    return X, conn


def _gpu(etype, X, conn, kind, p, state_per_qp=0, flags=0):
    s = A.Solver(flags=flags)
    s.set_mesh(etype, X, conn)
    s.set_material(kind, p["rho"], p["E"], p["nu"], p.get("yield_stress", 0.0), p.get("hardening", 0.0),
                   state_per_qp)
    return s


@pytest.mark.parametrize("etype", [A.TET4, A.HEX8], ids=["tet4", "hex8"])
@pytest.mark.parametrize("mat", list(MATS))
def test_internal_force_parity(gpu_available, etype, mat):
    kind, p = MATS[mat]
    X, conn = _mesh(etype, 5, 11)
    lam, mu = O.lame(p["E"], p["nu"])
    amp = 5e-3 if mat == "j2" else 2e-2
    u = amp * (2.0 * M.uniform01(5, 7, X.size).reshape(X.shape) - 1.0)
    print(f"internal_force etype={etype} mat={mat} nn={len(X)} ne={len(conn)}", flush=True)
    f_ref, H_ref, P_ref, s_ref = O.internal_force(etype, X, conn, kind, lam, mu, u,
                                                   yield_stress=p.get("yield_stress", 0.0),
                                                   hardening=p.get("hardening", 0.0))
    with _gpu(etype, X, conn, kind, p) as s:
        s.set_field(A.FIELD_U, u)
        s.internal_force(0.0, 1e-7)
        assert rel_err(s.get(A.FIELD_F), f_ref) < TOL_STEP
        assert rel_err(s.get(A.FIELD_H), H_ref) < TOL_STEP
        assert rel_err(s.get(A.FIELD_P), P_ref) < TOL_STEP
        if kind == A.J2:
            assert s_ref.max() > 0.0  # the case yields
            assert rel_err(s.get(A.FIELD_STATE), s_ref) < TOL_STEP


@pytest.mark.parametrize("etype", [A.TET4, A.HEX8], ids=["tet4", "hex8"])
@pytest.mark.parametrize("mat", list(MATS))
def test_100_steps_parity(gpu_available, etype, mat):
    kind, p = MATS[mat]
    X, conn = _mesh(etype, 4, 23)
    lam, mu = O.lame(p["E"], p["nu"])
    vscale = 30.0 if mat == "j2" else 5.0
    v0 = vscale * (2.0 * M.uniform01(9, 2, X.size).reshape(X.shape) - 1.0)
    fixed = np.zeros(X.shape, dtype=np.uint8)
    fixed[X[:, 2] < 1e-3] = 1
    v0[fixed.astype(bool)] = 0.0
    mass = O.lumped_mass(etype, X, conn, p["rho"])
    dt = 0.5 * M.stable_dt_estimate(X, conn, p["E"], p["nu"], p["rho"])
    print(f"100 steps etype={etype} mat={mat} dt={dt:.3e}", flush=True)
    ref = O.Solver(etype, X, conn, kind, lam, mu, mass, fixed, p.get("yield_stress", 0.0), p.get("hardening", 0.0))
    ref.v[:] = v0
    with _gpu(etype, X, conn, kind, p) as s:
        tot = s.compute_mass()
        assert abs(tot - mass.sum()) <= 1e-12 * mass.sum()
        assert rel_err(s.get(A.FIELD_MASS), mass) < 1e-14
        s.set_fixed(fixed)
        s.set_field(A.FIELD_V, v0)
        s.initialize()
        ref.run(dt, 0)
        # one step at the per-step bar
        s.step(dt, 1)
        ref.run(dt, 1)
        for fld, r in ((A.FIELD_U, ref.u), (A.FIELD_V, ref.v), (A.FIELD_A, ref.a), (A.FIELD_F, ref.f),
                       (A.FIELD_H, ref.H), (A.FIELD_P, ref.P)):
            assert rel_err(s.get(fld), r) < TOL_STEP, fld
        s.step(dt, 99)
        ref.run(dt, 99)
        for fld, r in ((A.FIELD_U, ref.u), (A.FIELD_V, ref.v), (A.FIELD_A, ref.a), (A.FIELD_F, ref.f),
                       (A.FIELD_H, ref.H), (A.FIELD_P, ref.P)):
            assert rel_err(s.get(fld), r) < TOL_100, fld
        if kind == A.J2:
            assert ref.state.max() > 0.0
            assert rel_err(s.get(A.FIELD_STATE), ref.state) < TOL_100
        # fixed DOFs stay at rest
        assert np.all(s.get(A.FIELD_V)[fixed.astype(bool)] == 0.0)


def test_hex8_state_per_qp(gpu_available):
    kind, p = MATS["j2"]
    X, conn = _mesh(A.HEX8, 3, 5)
    lam, mu = O.lame(p["E"], p["nu"])
    u = 5e-3 * (2.0 * M.uniform01(3, 7, X.size).reshape(X.shape) - 1.0)
    f_ref, H_ref, P_ref, s_ref = O.internal_force(A.HEX8, X, conn, kind, lam, mu, u, yield_stress=p["yield_stress"],
                                                   hardening=p["hardening"], state_per_qp=1)
    with _gpu(A.HEX8, X, conn, kind, p, state_per_qp=1) as s:
        s.set_field(A.FIELD_U, u)
        s.internal_force()
        assert rel_err(s.get(A.FIELD_F), f_ref) < TOL_STEP
        assert rel_err(s.get(A.FIELD_STATE), s_ref) < TOL_STEP


def test_per_element_material_fields(gpu_available):
    X, conn = _mesh(A.TET4, 4, 3)
    ne = len(conn)
    E = 200e9 * (0.8 + 0.4 * M.uniform01(1, 20, ne))
    nu = 0.25 + 0.1 * M.uniform01(1, 21, ne)
    lam, mu = M.lame(E, nu)
    u = 1e-2 * (2.0 * M.uniform01(2, 7, X.size).reshape(X.shape) - 1.0)
    f_ref, H_ref, P_ref, _ = O.internal_force(A.TET4, X, conn, A.NEO_HOOKEAN, 0.0, 0.0, u, elem_lambda=lam,
                                               elem_mu=mu)
    with _gpu(A.TET4, X, conn, A.NEO_HOOKEAN, MATS["nh"][1]) as s:
        s.set_material_fields(E, nu)
        s.set_field(A.FIELD_U, u)
        s.internal_force()
        assert rel_err(s.get(A.FIELD_F), f_ref) < TOL_STEP
        assert rel_err(s.get(A.FIELD_P), P_ref) < TOL_STEP


@pytest.mark.parametrize("etype", [A.TET4, A.HEX8], ids=["tet4", "hex8"])
def test_rigid_translation_gives_zero_force(gpu_available, etype):
    # Sum_k dN_k = 0, so a rigid translation has H = 0 up to the rounding of the
    # fused multiply-adds; the force must be at machine precision of the force
    # scale E h^2 |u| / h of the same displacement magnitude.
    X, conn = _mesh(etype, 3, 8)
    t = np.array([0.3, -0.2, 0.1])
    E, h = MATS["nh"][1]["E"], 1.0 / 3.0
    with _gpu(etype, X, conn, A.NEO_HOOKEAN, MATS["nh"][1]) as s:
        s.set_field(A.FIELD_U, np.tile(t, (len(X), 1)))
        s.internal_force()
        scale = E * h * np.linalg.norm(t)
        assert np.abs(s.get(A.FIELD_F)).max() < 1e-14 * scale
        assert np.abs(s.get(A.FIELD_H)).max() < 1e-14


def test_errors_are_reported_not_crashes(gpu_available):
    X, conn = _mesh(A.TET4, 2, 1)
    s = A.Solver()
    with pytest.raises(A.AphipError):
        s.step(1e-6, 1)  # no mesh yet
    bad = conn.copy()
    bad[0, 0] = len(X) + 5
    with pytest.raises(A.AphipError):
        s.set_mesh(A.TET4, X, bad)
    s.set_mesh(A.TET4, X, conn)
    with pytest.raises(A.AphipError):
        s.set_material(A.NEO_HOOKEAN, 7850.0, 200e9, 0.5)  # nu must be < 0.5
    s.set_material(A.NEO_HOOKEAN, 7850.0, 200e9, 0.3)
    with pytest.raises(A.AphipError):
        s.step(1e-6, 1)  # no mass yet
    s.close()


TILINGS = [(A.TET4, (64, 64)), (A.TET4, (128, 128)), (A.TET4, (256, 256)), (A.TET4, (512, 512)), (A.TET4, (1024, 1024)),
           (A.TET4, "two_pass"), (A.HEX8, (64, 64)), (A.HEX8, (128, 128)), (A.HEX8, (256, 256)), (A.HEX8, (8, 64)),
           (A.HEX8, (16, 128)), (A.HEX8, (32, 256)), (A.HEX8, (64, 512)), (A.HEX8, "two_pass")]


@pytest.mark.parametrize("etype,tiling", TILINGS, ids=[f"{'tet4' if e == 0 else 'hex8'}-{t}" for e, t in TILINGS])
def test_tilings_and_two_pass_path(gpu_available, etype, tiling):
    """Every tile shape (incl. a partial last tile) and the two-pass path give the oracle's
    20-step trajectory; the tile sums differ only in summation grouping.  The Hex8 mesh
    (7^3 = 343) fills both force rounds of a 256-element tile and leaves a partial one."""
    kind, p = MATS["nh"]
    X, conn = _mesh(etype, 7 if etype == A.HEX8 else 5, 31)
    lam, mu = O.lame(p["E"], p["nu"])
    v0 = 5.0 * (2.0 * M.uniform01(4, 2, X.size).reshape(X.shape) - 1.0)
    mass = O.lumped_mass(etype, X, conn, p["rho"])
    dt = 0.5 * M.stable_dt_estimate(X, conn, p["E"], p["nu"], p["rho"])
    ref = O.Solver(etype, X, conn, kind, lam, mu, mass)
    ref.v[:] = v0
    ref.run(dt, 20)
    flags = A.FLAG_TWO_PASS if tiling == "two_pass" else 0
    print(f"tiling {tiling} etype {etype}", flush=True)
    with _gpu(etype, X, conn, kind, p, flags=flags) as s:
        if tiling != "two_pass":
            s.set_tiling(*tiling)
        s.compute_mass()
        s.set_field(A.FIELD_V, v0)
        s.initialize()
        s.step(dt, 20)
        for fld, r in ((A.FIELD_U, ref.u), (A.FIELD_V, ref.v), (A.FIELD_A, ref.a), (A.FIELD_F, ref.f),
                       (A.FIELD_H, ref.H), (A.FIELD_P, ref.P)):
            assert rel_err(s.get(fld), r) < 1e-11, fld


def test_tile_results_are_deterministic(gpu_available):
    kind, p = MATS["nh"]
    X, conn = _mesh(A.TET4, 6, 2)
    v0 = 5.0 * (2.0 * M.uniform01(4, 2, X.size).reshape(X.shape) - 1.0)
    dt = 0.5 * M.stable_dt_estimate(X, conn, p["E"], p["nu"], p["rho"])
    outs = []
    for _ in range(2):
        with _gpu(A.TET4, X, conn, kind, p) as s:
            s.compute_mass()
            s.set_field(A.FIELD_V, v0)
            s.initialize()
            s.step(dt, 10)
            outs.append((s.get(A.FIELD_U), s.get(A.FIELD_F), s.get(A.FIELD_P)))
    for a, b in zip(*outs):
        assert np.array_equal(a, b)  # bitwise: no atomics anywhere in the step


@pytest.mark.parametrize("etype,tiling", [(A.TET4, (128, 128)), (A.HEX8, (128, 128)), (A.HEX8, (16, 128))],
                         ids=["tet4", "hex8", "hex8-gauss-lanes"])
@pytest.mark.parametrize("mat", ["nh", "j2"])
def test_steps_of_one_call_match_single_steps(gpu_available, etype, tiling, mat):
    """Within one aphip_step call the shared-node pass hands the next step's tile kernel
    the shared nodes' predicted displacement; a call of one step predicts it from u, v
    and a.  Thirteen single-step calls, one eager call of thirteen steps and one call
    through the two-step graph give bitwise equal fields (fixed DOFs, many tiles)."""
    kind, p = MATS[mat]
    X, conn = _mesh(etype, 8 if etype == A.HEX8 else 6, 12)
    fixed = np.zeros(X.shape, dtype=np.uint8)
    fixed[X[:, 0] < 1e-3] = 1
    fixed[X[:, 2] < 1e-3, 2] = 1
    vscale = 30.0 if mat == "j2" else 5.0
    v0 = vscale * (2.0 * M.uniform01(7, 2, X.size).reshape(X.shape) - 1.0)
    v0[fixed.astype(bool)] = 0.0
    dt = 0.4 * M.stable_dt_estimate(X, conn, p["E"], p["nu"], p["rho"])
    flds = [A.FIELD_U, A.FIELD_V, A.FIELD_A, A.FIELD_F, A.FIELD_H, A.FIELD_P]
    if kind == A.J2:
        flds.append(A.FIELD_STATE)
    outs = []
    for mode in ("single", "eager", "graph"):
        print(f"one call vs single steps etype={etype} tiling={tiling} mat={mat} mode={mode}", flush=True)
        with _gpu(etype, X, conn, kind, p) as s:
            s.set_tiling(*tiling)
            s.set_fixed(fixed)
            s.compute_mass()
            s.set_field(A.FIELD_V, v0)
            s.initialize()
            s.set_graph(mode == "graph")
            if mode == "single":
                for _ in range(13):
                    s.step(dt, 1)
            else:
                s.step(dt, 13)
            outs.append([s.get(f) for f in flds])
    for other in outs[1:]:
        for f, a, b in zip(flds, outs[0], other):
            assert np.array_equal(a, b), f
    assert np.abs(outs[0][0]).max() > 0.0


@pytest.mark.parametrize("mat", ["nh", "le", "j2"])
@pytest.mark.parametrize("te", [64, 128, 256])
def test_pipelined_step_matches_one_tile_per_workgroup(gpu_available, mat, te):
    """The persistent pipelined Tet4 step and the one-tile-per-workgroup step run the
    same arithmetic (explicit fused multiply-adds, no implicit contraction) in the
    same summation order: bitwise equal fields after 15 steps, and each matches the
    oracle.  The mesh has a partial last tile, fixed DOFs and per-element moduli."""
    kind, p = MATS[mat]
    X, conn = _mesh(A.TET4, 7, 41)
    ne = len(conn)  # 2058 = 32 x 64 + 10: the last tile is partial for every te
    E = p["E"] * (0.9 + 0.2 * M.uniform01(3, 20, ne))
    nu = np.full(ne, p["nu"])
    vscale = 30.0 if mat == "j2" else 5.0
    v0 = vscale * (2.0 * M.uniform01(6, 2, X.size).reshape(X.shape) - 1.0)
    fixed = np.zeros(X.shape, dtype=np.uint8)
    fixed[X[:, 0] < 1e-3] = 1
    v0[fixed.astype(bool)] = 0.0
    dt = 0.4 * M.stable_dt_estimate(X, conn, p["E"], p["nu"], p["rho"])
    print(f"pipelined vs one-tile mat={mat} te={te} ne={ne}", flush=True)
    outs = []
    for flags in (A.FLAG_PIPELINED, 0):
        with _gpu(A.TET4, X, conn, kind, p, flags=flags) as s:
            s.set_tiling(te, te)
            s.set_material_fields(E, nu)
            s.set_fixed(fixed)
            s.compute_mass()
            s.set_field(A.FIELD_V, v0)
            s.initialize()
            s.step(dt, 15)
            flds = [A.FIELD_U, A.FIELD_V, A.FIELD_A, A.FIELD_F, A.FIELD_H, A.FIELD_P]
            if kind == A.J2:
                flds.append(A.FIELD_STATE)
            outs.append([s.get(f) for f in flds])
    lam, mu = M.lame(E, nu)
    mass = O.lumped_mass(A.TET4, X, conn, p["rho"])
    ref = O.Solver(A.TET4, X, conn, kind, 0.0, 0.0, mass, fixed, p.get("yield_stress", 0.0),
                   p.get("hardening", 0.0), elem_lambda=lam, elem_mu=mu)
    ref.v[:] = v0
    ref.run(dt, 15)
    for k, out in enumerate(outs):
        for fld, r in zip(range(4), (ref.u, ref.v, ref.a, ref.f)):
            assert rel_err(out[fld], r) < 1e-11, (k, fld)
    for fld, (a, b) in enumerate(zip(*outs)):
        if not np.array_equal(a, b):
            d = np.abs(a - b)
            i = np.unravel_index(np.argmax(d), d.shape)
            print(f"field {fld}: {np.count_nonzero(d)} differ, max {d.max():.3e} at {i}")
        assert np.array_equal(a, b), fld
    assert np.abs(outs[0][0]).max() > 0.0
