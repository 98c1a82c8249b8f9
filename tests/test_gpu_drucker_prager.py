"""GPU parity of the Drucker-Prager material (include/Materials/DruckerPrager.h), a
rate-form material with an oblique return, against the CPU oracle through the C ABI:
one evaluation, 100 steps of a stretching block that yields (Tet4, Hex8 per element and
per Gauss point, the element-partials path) and the RKPM cells (one-pass and two-pass).
Bars: 1e-12 per evaluation, 1e-10 after 100 steps.
"""
import numpy as np
import pytest

import oracle as O
import aperi_hip as A
from aperi_hip import meshgen as M
from conftest import rel_err

pytestmark = pytest.mark.gpu

E, NU, RHO = 1e9, 0.3, 1000.0
DP = dict(A1=3e6, A2=0.2, A2G=0.1)
ODP = dict(DP, bulk=E / (3 * (1 - 2 * NU)))


def _mesh(etype, n, seed):
    X, hc = M.hex_lattice(n, n, n)
    X = M.jitter(X, 1.0 / n, 0.1, seed, 1)
    return X, hc, (hc if etype == A.HEX8 else M.split_tets(X, hc))


def _gpu(etype, X, conn, state_per_qp=0, flags=0):
    s = A.Solver(flags=flags)
    s.set_mesh(etype, X, conn)
    s.set_material(A.DRUCKER_PRAGER, RHO, E, NU, state_per_qp=state_per_qp, drucker_prager=DP)
    return s


CASES = [(A.TET4, 0, 0), (A.HEX8, 0, 0), (A.HEX8, 1, 0), (A.TET4, 0, A.FLAG_TWO_PASS), (A.HEX8, 0, A.FLAG_TWO_PASS)]
IDS = ["tet4", "hex8", "hex8-qp", "tet4-partials", "hex8-partials"]


@pytest.mark.parametrize("etype,per_qp,flags", CASES, ids=IDS)
def test_dp_force_parity(gpu_available, etype, per_qp, flags):
    X, hc, conn = _mesh(etype, 5, 11)
    lam, mu = O.lame(E, NU)
    u = 2e-2 * (2.0 * M.uniform01(5, 7, X.size).reshape(X.shape) - 1.0)
    dt = 3e-5
    ne = len(conn)
    f_ref, H_ref, P_ref, s_ref = O.internal_force(etype, X, conn, O.DP, lam, mu, u, state_per_qp=per_qp,
                                                   creep=ODP, rate=(np.zeros((ne, 9)), np.zeros((ne, 9)), dt))
    assert (s_ref[0::2] > 0).sum() > 0 and (s_ref[0::2] == 0).sum() > 0  # some points yield
    print(f"dp force etype={etype} per_qp={per_qp} flags={flags}", flush=True)
    with _gpu(etype, X, conn, per_qp, flags) as s:
        s.set_field(A.FIELD_U, u)
        s.internal_force(0.0, dt)
        for fld, rr in ((A.FIELD_F, f_ref), (A.FIELD_H, H_ref), (A.FIELD_P, P_ref), (A.FIELD_STATE, s_ref)):
            assert rel_err(s.get(fld), rr) < 1e-12, fld


@pytest.mark.parametrize("etype,per_qp,flags", CASES, ids=IDS)
def test_dp_100_steps(gpu_available, etype, per_qp, flags):
    X, hc, conn = _mesh(etype, 4, 23)
    lam, mu = O.lame(E, NU)
    mass = O.lumped_mass(etype, X, conn, RHO)
    dt = 0.5 * M.stable_dt_estimate(X, M.split_tets(X, hc), E, NU, RHO)
    v0 = np.zeros_like(X)
    v0[:, 0] = 0.2 / (100 * dt) * (X[:, 0] - 0.5)
    ref = O.Solver(etype, X, conn, O.DP, lam, mu, mass, state_per_qp=per_qp, creep=ODP)
    ref.v[:] = v0
    ref.run(dt, 100)
    assert ref.state[0::2].max() > 1e-2 and np.isfinite(ref.state).all()
    print(f"dp 100 steps etype={etype} per_qp={per_qp} flags={flags}", flush=True)
    with _gpu(etype, X, conn, per_qp, flags) as s:
        s.compute_mass()
        s.set_field(A.FIELD_V, v0)
        s.initialize()
        s.step(dt, 100)
        for fld, rr in ((A.FIELD_U, ref.u), (A.FIELD_V, ref.v), (A.FIELD_A, ref.a), (A.FIELD_F, ref.f),
                        (A.FIELD_H, ref.H), (A.FIELD_P, ref.P), (A.FIELD_STATE, ref.state)):
            assert rel_err(s.get(fld), rr) < 1e-10, fld


@pytest.mark.parametrize("one_pass", [True, False], ids=["one-pass", "two-pass"])
def test_dp_rkpm_cells(gpu_available, one_pass):
    X, hc = M.hex_lattice(5, 5, 5)
    X = M.jitter(X, 0.2, 0.1, 3, 1)
    t = M.split_tets(X, hc)
    scale = 1.5
    r = O.kernel_radius(O.max_edge_length(X, t), scale)
    num, nbr, _ = O.neighbor_search(X, r)
    phi = O.rk_function_values(X, r, num, nbr)
    cells = O.rk_cells(X, t, num, nbr, phi) if one_pass else O.rk_cells_two_pass(X, t)
    lam, mu = O.lame(E, NU)
    mass = O.rk_mass(t, cells[3], RHO, num, nbr, phi)
    dt = 0.5 * M.stable_dt_estimate(X, t, E, NU, RHO)
    v0 = np.zeros_like(X)
    v0[:, 0] = 0.2 / (100 * dt) * (X[:, 0] - 0.5)
    ref = O.Solver(-1, X, None, O.DP, lam, mu, mass, cells=cells, two_pass=None if one_pass else (num, nbr, phi),
                   creep=ODP)
    ref.v[:] = v0
    ref.run(dt, 100)
    print(f"dp rkpm one_pass={one_pass}", flush=True)
    with A.Solver() as s:
        s.set_mesh(A.TET4, X, t)
        s.set_material(A.DRUCKER_PRAGER, RHO, E, NU, drucker_prager=DP)
        s.rk_setup(scale, one_pass=one_pass)
        s.compute_mass()
        s.set_field(A.FIELD_V, v0)
        s.initialize()
        s.step(dt, 100)
        assert ref.state[0::2].max() > 1e-2
        for fld, rr in ((A.FIELD_U, ref.u), (A.FIELD_V, ref.v), (A.FIELD_F, ref.f), (A.FIELD_P, ref.P),
                        (A.FIELD_STATE, ref.state)):
            assert rel_err(s.get(fld), rr) < 1e-10, fld
