"""GPU parity of the neo-Hookean-with-damage material (include/Materials/NeoHookeWithDamage.h)
against the CPU oracle, through the C ABI: Tet4 and Hex8 elements (per-element and
per-Gauss-point state), the element-partials path, the RKPM cells (one-pass and
two-pass, where a new failure's marker becomes the maximum damage), 100 steps of a
stretching block that accrues damage (capped increments) or fails, and the reference's
parameter checks.  Bars: 1e-12 per evaluation, 1e-10 after 100 steps.
"""
import re

import numpy as np
import pytest

import oracle as O
import aperi_hip as A
from aperi_hip import meshgen as M
from conftest import rel_err

pytestmark = pytest.mark.gpu

P_MAT = dict(E=1e9, nu=0.3, rho=1000.0)
GRADUAL = dict(I1_critical=3.004, I1_failure=3.1, alpha=1.5, maximum_damage_increment=0.05)
FAILING = dict(I1_critical=3.004, I1_failure=3.03, alpha=1.0, maximum_damage=0.9)


def _mesh(etype, n, seed):
    X, hc = M.hex_lattice(n, n, n)
    X = M.jitter(X, 1.0 / n, 0.1, seed, 1)
    return X, hc, (hc if etype == A.HEX8 else M.split_tets(X, hc))


def _gpu(etype, X, conn, dmg, state_per_qp=0, flags=0):
    s = A.Solver(flags=flags)
    s.set_mesh(etype, X, conn)
    s.set_material(A.NEO_HOOKEAN_DAMAGE, P_MAT["rho"], P_MAT["E"], P_MAT["nu"], state_per_qp=state_per_qp,
                   damage=dmg)
    return s


CASES = [(A.TET4, 0, 0), (A.HEX8, 0, 0), (A.HEX8, 1, 0), (A.TET4, 0, A.FLAG_TWO_PASS), (A.HEX8, 0, A.FLAG_TWO_PASS)]
IDS = ["tet4", "hex8", "hex8-qp", "tet4-partials", "hex8-partials"]


@pytest.mark.parametrize("etype,per_qp,flags", CASES, ids=IDS)
@pytest.mark.parametrize("dmg", [GRADUAL, FAILING], ids=["gradual", "failing"])
def test_damage_force_parity(gpu_available, etype, per_qp, flags, dmg):
    # one evaluation from an undamaged state; the prior damage (read from the N state)
    # is exercised by the step tests below
    X, hc, conn = _mesh(etype, 5, 11)
    lam, mu = O.lame(P_MAT["E"], P_MAT["nu"])
    u = 4e-2 * (2.0 * M.uniform01(5, 7, X.size).reshape(X.shape) - 1.0)
    f_ref, H_ref, P_ref, s_ref = O.internal_force(etype, X, conn, O.NH_DAMAGE, lam, mu, u, state_per_qp=per_qp,
                                                   damage=dmg)
    d = s_ref[0::2]
    assert (d > 0).sum() > 0 and (d == 0).sum() > 0
    if dmg is FAILING:
        assert (d == 2.0).sum() > 0  # new failures keep their marker on the Gaussian path
    print(f"damage force etype={etype} per_qp={per_qp} flags={flags}", flush=True)
    with _gpu(etype, X, conn, dmg, per_qp, flags) as s:
        s.set_field(A.FIELD_U, u)
        s.internal_force(0.0, 1e-7)
        assert rel_err(s.get(A.FIELD_F), f_ref) < 1e-12
        assert rel_err(s.get(A.FIELD_H), H_ref) < 1e-12
        assert rel_err(s.get(A.FIELD_P), P_ref) < 1e-12
        st = s.get(A.FIELD_STATE)
        assert st.shape == s_ref.shape and rel_err(st, s_ref) < 1e-12
        assert np.array_equal(st[0::2] == 2.0, d == 2.0)


@pytest.mark.parametrize("etype,per_qp,flags", CASES, ids=IDS)
@pytest.mark.parametrize("dmg", [GRADUAL, FAILING], ids=["gradual", "failing"])
def test_damage_100_steps(gpu_available, etype, per_qp, flags, dmg):
    X, hc, conn = _mesh(etype, 4, 23)
    lam, mu = O.lame(P_MAT["E"], P_MAT["nu"])
    mass = O.lumped_mass(etype, X, conn, P_MAT["rho"])
    dt = 0.5 * M.stable_dt_estimate(X, M.split_tets(X, hc), P_MAT["E"], P_MAT["nu"], P_MAT["rho"])
    v0 = np.zeros_like(X)
    v0[:, 0] = 0.2 / (100 * dt) * (X[:, 0] - 0.5)
    ref = O.Solver(etype, X, conn, O.NH_DAMAGE, lam, mu, mass, state_per_qp=per_qp, damage=dmg)
    ref.v[:] = v0
    ref.run(dt, 100)
    d = ref.state[0::2]
    assert d.max() > 0.0
    print(f"damage 100 steps etype={etype} per_qp={per_qp} flags={flags} dmax={d.max():.3f}", flush=True)
    with _gpu(etype, X, conn, dmg, per_qp, flags) as s:
        s.compute_mass()
        s.set_field(A.FIELD_V, v0)
        s.initialize()
        s.step(dt, 100)
        for fld, rr in ((A.FIELD_U, ref.u), (A.FIELD_V, ref.v), (A.FIELD_A, ref.a), (A.FIELD_F, ref.f),
                        (A.FIELD_H, ref.H), (A.FIELD_P, ref.P), (A.FIELD_STATE, ref.state)):
            assert rel_err(s.get(fld), rr) < 1e-10, fld


@pytest.mark.parametrize("one_pass", [True, False], ids=["one-pass", "two-pass"])
def test_damage_rkpm_cells(gpu_available, one_pass):
    X, hc = M.hex_lattice(5, 5, 5)
    X = M.jitter(X, 0.2, 0.1, 3, 1)
    t = M.split_tets(X, hc)
    scale = 1.5
    r = O.kernel_radius(O.max_edge_length(X, t), scale)
    num, nbr, _ = O.neighbor_search(X, r)
    phi = O.rk_function_values(X, r, num, nbr)
    cells = O.rk_cells(X, t, num, nbr, phi) if one_pass else O.rk_cells_two_pass(X, t)
    lam, mu = O.lame(P_MAT["E"], P_MAT["nu"])
    mass = O.rk_mass(t, cells[3], P_MAT["rho"], num, nbr, phi)
    dt = 0.5 * M.stable_dt_estimate(X, t, P_MAT["E"], P_MAT["nu"], P_MAT["rho"])
    v0 = np.zeros_like(X)
    v0[:, 0] = 0.2 / (100 * dt) * (X[:, 0] - 0.5)
    ref = O.Solver(-1, X, None, O.NH_DAMAGE, lam, mu, mass, cells=cells,
                   two_pass=None if one_pass else (num, nbr, phi), damage=FAILING)
    ref.v[:] = v0
    ref.run(dt, 1)
    print(f"damage rkpm one_pass={one_pass}", flush=True)
    with A.Solver() as s:
        s.set_mesh(A.TET4, X, t)
        s.set_material(A.NEO_HOOKEAN_DAMAGE, P_MAT["rho"], P_MAT["E"], P_MAT["nu"], damage=FAILING)
        s.rk_setup(scale, one_pass=one_pass)
        s.compute_mass()
        s.set_field(A.FIELD_V, v0)
        s.initialize()
        s.step(dt, 1)
        for fld, rr in ((A.FIELD_U, ref.u), (A.FIELD_F, ref.f), (A.FIELD_P, ref.P), (A.FIELD_STATE, ref.state)):
            assert rel_err(s.get(fld), rr) < 1e-12, fld
        s.step(dt, 99)
        ref.run(dt, 99)
        d = ref.state[0::2]
        assert d.max() == 0.9 and not (d == 2.0).any()  # markers became the maximum damage
        for fld, rr in ((A.FIELD_U, ref.u), (A.FIELD_V, ref.v), (A.FIELD_F, ref.f), (A.FIELD_P, ref.P),
                        (A.FIELD_STATE, ref.state)):
            assert rel_err(s.get(fld), rr) < 1e-10, fld


def test_damage_parameter_checks(gpu_available):
    X, hc, conn = _mesh(A.TET4, 2, 1)
    bad = [(dict(FAILING, I1_failure=3.0), "I1_critical must be < I1_failure"),
           (dict(FAILING, alpha=-1.0), "alpha must be >= 0"),
           (dict(FAILING, maximum_damage=1.5), "max_effective_damage must be in [0, 1]"),
           (dict(FAILING, I1_critical=-1.0), "I1_critical must be >= 0")]
    for dmg, msg in bad:
        with A.Solver() as s:
            s.set_mesh(A.TET4, X, conn)
            with pytest.raises(A.AphipError, match=re.escape(msg)):
                s.set_material(A.NEO_HOOKEAN_DAMAGE, 1000.0, 1e9, 0.3, damage=dmg)
    with A.Solver() as s:
        with pytest.raises(A.AphipError, match="ensembles take"):
            s.batch_begin(A.TET4, A.NEO_HOOKEAN_DAMAGE)
