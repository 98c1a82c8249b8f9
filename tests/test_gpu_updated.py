"""GPU parity of the updated Lagrangian formulation (ExplicitTimeIntegratorUpdated,
include/ExplicitTimeIntegrator.h:193-226, :466-486; ComputeInternalForceBase.h:158-166;
ComputeInternalForceGaussian.h:109-123) against the oracle's literal restatement
(B F_r, w / det F_r) through the C ABI, and against the GPU's own total formulation on
Tet4, where the two are equal up to rounding.  Bars: 1e-10 after 100 steps (oracle),
1e-9 (total vs updated).
"""
import numpy as np
import pytest

import oracle as O
import aperi_hip as A
from aperi_hip import meshgen as M
from conftest import rel_err

pytestmark = pytest.mark.gpu

E, NU, RHO = 1e9, 0.3, 1000.0
KINDS = {"nh": (A.NEO_HOOKEAN, {}), "j2": (A.J2, dict(yield_stress=2e6, hardening=1e8)),
         "creep": (A.POWER_LAW_CREEP, {})}
CREEP = dict(A=1e-18, n=3.0, m=0.0, constant_temperature=0.0)


def _block(etype, n=4, seed=4):
    X, hc = M.hex_lattice(n, n, n)
    X = M.jitter(X, 1.0 / n, 0.1, seed, 1)
    return X, hc, (hc if etype == A.HEX8 else M.split_tets(X, hc))


def _v0(X, dt):
    v0 = np.zeros_like(X)
    v0[:, 0] = 0.2 / (100 * dt) * (X[:, 0] - 0.5)
    v0[:, 1] = 0.05 / (100 * dt) * (X[:, 2] - 0.5)
    return v0


def _solver(etype, X, conn, kind, kw, formulation, state_per_qp=0):
    s = A.Solver()
    s.set_mesh(etype, X, conn)
    s.set_material(kind, RHO, E, NU, kw.get("yield_stress", 0.0), kw.get("hardening", 0.0), state_per_qp,
                   creep=CREEP if kind == A.POWER_LAW_CREEP else None)
    s.set_formulation(formulation)
    return s


@pytest.mark.parametrize("etype", [A.TET4, A.HEX8], ids=["tet4", "hex8"])
@pytest.mark.parametrize("mat", list(KINDS))
def test_updated_100_steps(gpu_available, etype, mat):
    kind, kw = KINDS[mat]
    X, hc, conn = _block(etype)
    lam, mu = O.lame(E, NU)
    mass = O.lumped_mass(etype, X, conn, RHO)
    dt = 0.5 * M.stable_dt_estimate(X, M.split_tets(X, hc), E, NU, RHO)
    v0 = _v0(X, dt)
    okind = {A.NEO_HOOKEAN: O.NEO_HOOKEAN, A.J2: O.J2, A.POWER_LAW_CREEP: O.CREEP}[kind]
    ref = O.Solver(etype, X, conn, okind, lam, mu, mass, yield_stress=kw.get("yield_stress", 0.0),
                   hardening=kw.get("hardening", 0.0),
                   creep=dict(CREEP, bulk=E / (3 * (1 - 2 * NU))) if kind == A.POWER_LAW_CREEP else None,
                   updated=True)
    ref.v[:] = v0
    ref.run(dt, 100)
    print(f"updated 100 steps etype={etype} mat={mat}", flush=True)
    with _solver(etype, X, conn, kind, kw, A.UPDATED_LAGRANGIAN) as s:
        s.compute_mass()
        s.set_field(A.FIELD_V, v0)
        s.initialize()
        s.step(dt, 100)
        fields = [(A.FIELD_U, ref.u), (A.FIELD_V, ref.v), (A.FIELD_A, ref.a), (A.FIELD_F, ref.f),
                  (A.FIELD_H, ref.H), (A.FIELD_P, ref.P)]
        if kind != A.NEO_HOOKEAN:
            fields.append((A.FIELD_STATE, ref.state))
        for fld, rr in fields:
            assert rel_err(s.get(fld), rr) < 1e-10, fld


def test_updated_equals_total_tet4(gpu_available):
    X, hc, conn = _block(A.TET4)
    dt = 0.5 * M.stable_dt_estimate(X, conn, E, NU, RHO)
    v0 = _v0(X, dt)
    out = {}
    for form in (A.TOTAL_LAGRANGIAN, A.UPDATED_LAGRANGIAN):
        with _solver(A.TET4, X, conn, A.NEO_HOOKEAN, {}, form) as s:
            s.compute_mass()
            s.set_field(A.FIELD_V, v0)
            s.initialize()
            s.step(dt, 100)
            out[form] = [s.get(f) for f in (A.FIELD_U, A.FIELD_F, A.FIELD_H, A.FIELD_P)]
    for a, b in zip(out[A.UPDATED_LAGRANGIAN], out[A.TOTAL_LAGRANGIAN]):
        assert rel_err(a, b) < 1e-9


def test_updated_rejections(gpu_available):
    X, hc, conn = _block(A.TET4, 3)
    with A.Solver() as s:
        s.set_mesh(A.TET4, X, conn)
        s.set_material(A.NEO_HOOKEAN, RHO, E, NU)
        s.rk_setup(1.5)
        with pytest.raises(A.AphipError, match="not to RKPM cells"):
            s.set_formulation(A.UPDATED_LAGRANGIAN)
    with _solver(A.TET4, X, conn, A.NEO_HOOKEAN, {}, A.UPDATED_LAGRANGIAN) as s:
        s.compute_mass()
        with pytest.raises(A.AphipError, match="total formulation"):
            s.power_method_time_increment()
        with pytest.raises(A.AphipError):
            s.set_formulation(1)  # incremental: not supported
    # the reverse call orders: a formulation first, then the cell set-ups or a rate-form material
    for form in (A.UPDATED_LAGRANGIAN, A.SEMI_LAGRANGIAN):
        with _solver(A.TET4, X, conn, A.NEO_HOOKEAN, {}, form) as s:
            with pytest.raises(A.AphipError, match="not to RKPM cells"):
                s.rk_setup(1.5)
            with pytest.raises(A.AphipError, match="not to RKPM cells"):
                s.rk_setup(1.5, one_pass=False)
            with pytest.raises(A.AphipError, match="not to smoothed cells"):
                s.strain_smoothing_setup()
    with _solver(A.TET4, X, conn, A.NEO_HOOKEAN, {}, A.SEMI_LAGRANGIAN) as s:
        with pytest.raises(A.AphipError, match="rate-form"):
            s.set_material(A.POWER_LAW_CREEP, RHO, E, NU, creep=CREEP)
        with pytest.raises(A.AphipError, match="rate-form"):
            s.set_material(A.DRUCKER_PRAGER, RHO, E, NU, drucker_prager=dict(A1=1e8, A2=0.1, A2G=0.1))


# ---- semi-Lagrangian (ExplicitTimeIntegratorSemi, include/ExplicitTimeIntegrator.h:488-546) ----

@pytest.mark.parametrize("etype", [A.TET4, A.HEX8], ids=["tet4", "hex8"])
@pytest.mark.parametrize("mat", ["nh", "j2"])
def test_semi_lagrangian_100_steps(gpu_available, etype, mat):
    kind, kw = KINDS[mat]
    X, hc, conn = _block(etype)
    lam, mu = O.lame(E, NU)
    mass = O.lumped_mass(etype, X, conn, RHO)
    dt = 0.5 * M.stable_dt_estimate(X, M.split_tets(X, hc), E, NU, RHO)
    v0 = _v0(X, dt)
    okind = {A.NEO_HOOKEAN: O.NEO_HOOKEAN, A.J2: O.J2}[kind]
    ref = O.Solver(etype, X, conn, okind, lam, mu, mass, yield_stress=kw.get("yield_stress", 0.0),
                   hardening=kw.get("hardening", 0.0), semi_interval=7)
    ref.v[:] = v0
    ref.run(dt, 100)
    print(f"semi 100 steps etype={etype} mat={mat}", flush=True)
    with A.Solver() as s:
        s.set_mesh(etype, X, conn)
        s.set_material(kind, RHO, E, NU, kw.get("yield_stress", 0.0), kw.get("hardening", 0.0))
        s.set_semi_lagrangian(7)
        s.compute_mass()
        s.set_field(A.FIELD_V, v0)
        s.initialize()
        s.step(dt, 30)
        s.step(dt, 70)
        fields = [(A.FIELD_U, ref.u), (A.FIELD_V, ref.v), (A.FIELD_A, ref.a), (A.FIELD_F, ref.f),
                  (A.FIELD_H, ref.H), (A.FIELD_P, ref.P)]
        if kind == A.J2:
            fields.append((A.FIELD_STATE, ref.state))
        for fld, rr in fields:
            assert rel_err(s.get(fld), rr) < 1e-10, fld


def test_semi_interval_one_is_updated(gpu_available):
    X, hc, conn = _block(A.HEX8)
    dt = 0.5 * M.stable_dt_estimate(X, M.split_tets(X, hc), E, NU, RHO)
    v0 = _v0(X, dt)
    out = []
    for semi in (False, True):
        with A.Solver() as s:
            s.set_mesh(A.HEX8, X, conn)
            s.set_material(A.NEO_HOOKEAN, RHO, E, NU)
            if semi:
                s.set_semi_lagrangian(1)
            else:
                s.set_formulation(A.UPDATED_LAGRANGIAN)
            s.compute_mass()
            s.set_field(A.FIELD_V, v0)
            s.initialize()
            s.step(dt, 50)
            out.append([s.get(f) for f in (A.FIELD_U, A.FIELD_F, A.FIELD_H, A.FIELD_P)])
    for a, b in zip(*out):
        assert np.array_equal(a, b)


def test_semi_refuses_rate_form(gpu_available):
    X, hc, conn = _block(A.TET4, 3)
    with A.Solver() as s:
        s.set_mesh(A.TET4, X, conn)
        s.set_material(A.POWER_LAW_CREEP, RHO, E, NU, creep=CREEP)
        with pytest.raises(A.AphipError, match="rate-form"):
            s.set_semi_lagrangian(5)
