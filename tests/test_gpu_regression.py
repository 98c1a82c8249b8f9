"""The GPU step on the reference's single-tet regression case
(test/regression_tests/single_tet/fem): the gold maxima of compare.exodiff:26-52 with
their locations to the printed 8 digits, every output against the oracle at 1e-12; and
the case's own power-method time stepper (scale 0.1, update every 10 steps) with the
eigenvector carried from update to update, against the oracle's restatement."""
import numpy as np
import pytest

import oracle as O
import aperi_hip as A
import single_tet_case as C
from conftest import rel_err

pytestmark = pytest.mark.gpu
GOLD_REL = 5e-8


def _gpu(case):
    m = case["material"]
    s = A.Solver()
    s.set_mesh(A.TET4, case["X"], case["conn"])
    s.set_material(A.NEO_HOOKEAN, m["density"], m["youngs_modulus"], m["poissons_ratio"])
    s.compute_mass()
    s.set_fixed(case["fixed_mask"])
    s.set_field(A.FIELD_V, case["v0"])
    return s


def _fields(s):
    return lambda: tuple(s.get(f) for f in (A.FIELD_U, A.FIELD_V, A.FIELD_A, A.FIELD_F))


def _oracle(case):
    m = case["material"]
    lam, mu = O.lame(m["youngs_modulus"], m["poissons_ratio"])
    mass = O.lumped_mass(0, case["X"], case["conn"], m["density"])
    ref = O.Solver(0, case["X"], case["conn"], O.NEO_HOOKEAN, lam, mu, mass, case["fixed_mask"])
    ref.v[:] = case["v0"]
    ref.run(0.0, 0)
    return ref


def _compare(outs, ref_outs):
    assert len(outs) == len(ref_outs)
    for o, r in zip(outs, ref_outs):
        assert o[0] == r[0]
        for k in range(1, 5):
            if np.abs(r[k]).max() > 0.0:
                assert rel_err(o[k], r[k]) < 1e-12, (o[0], k)
            else:
                assert np.abs(o[k]).max() == 0.0


def test_single_tet_regression_gold(gpu_available):
    case = C.load()
    print("single tet regression, fixed increment 0.086222", flush=True)
    with _gpu(case) as s:
        s.initialize()
        outs = C.run(case, lambda dt: s.step(dt, 1), _fields(s))
        assert s.get(A.FIELD_MASS)[0] == pytest.approx(case["mass"], rel=GOLD_REL)
    got = C.maxima(outs)
    for name, g in case["maxima"].items():
        value, output, node = got[name]
        assert value == pytest.approx(g["value"], rel=GOLD_REL), name
        assert (output, node) == (g["output"], g["node"]), name
    ref = _oracle(case)
    _compare(outs, C.run(case, lambda dt: ref.run(dt, 1), lambda: (ref.u, ref.v, ref.a, ref.f)))


def test_single_tet_power_method_stepper_matches_oracle(gpu_available):
    case = C.load()
    scale, interval = 0.1, 10
    print("single tet, power-method time stepper", flush=True)
    with _gpu(case) as s:
        s.initialize()
        gpu_stepper = C.PowerMethodStepper(lambda prev: s.power_method_time_increment(previous_time_increment=prev)[0],
                                           scale, interval, case["time_end"])
        outs = C.run(case, lambda dt: s.step(dt, 1), _fields(s), gpu_stepper)
    ref = _oracle(case)
    l = O.max_edge_length(case["X"], case["conn"])
    eig = {}
    ref_stepper = C.PowerMethodStepper(lambda prev: ref.power_method(l, previous_dt=prev, eigenvector=eig)[0],
                                       scale, interval, case["time_end"])
    ref_outs = C.run(case, lambda dt: ref.run(dt, 1), lambda: (ref.u, ref.v, ref.a, ref.f), ref_stepper)
    assert [n for n, _ in gpu_stepper.updates] == [n for n, _ in ref_stepper.updates]
    for (_, a), (_, b) in zip(gpu_stepper.updates, ref_stepper.updates):
        assert a == pytest.approx(b, rel=1e-9)
    assert len(gpu_stepper.updates) > 5
    assert len(outs) == len(ref_outs)
    for o, r in zip(outs, ref_outs):
        assert o[0] == pytest.approx(r[0], rel=1e-9)
        for k in range(1, 5):
            if np.abs(r[k]).max() > 0.0:
                assert rel_err(o[k], r[k]) < 1e-8, (o[0], k)
