"""The oracle against the reference's own regression gold: the single-tet explicit
dynamics case of test/regression_tests/single_tet/fem (neo-Hookean Tet4, fixed base,
apex released with v_z = -0.001, 100 increments of 0.086222, output every increment).
Every maximum the reference's exodiff summary lists (compare.exodiff:26-52, the gold's
absolute maxima with their output index and node) is reproduced, with its location, to
the 8 digits printed there.  This pins the whole restated chain at once: lumped mass
(src/MassUtils.cpp:26-39), Tet4 B and weight, neo-Hookean stress, the internal force at
free and fixed nodes (reactions), the central-difference update with velocity and
acceleration BCs, and the output schedule."""
import numpy as np
import pytest

import oracle as O
import single_tet_case as C

GOLD_REL = 5e-8  # the gold is printed with 8 significant digits


def _oracle_run(case, get_increment=None):
    m = case["material"]
    lam, mu = O.lame(m["youngs_modulus"], m["poissons_ratio"])
    mass = O.lumped_mass(0, case["X"], case["conn"], m["density"])
    ref = O.Solver(0, case["X"], case["conn"], O.NEO_HOOKEAN, lam, mu, mass, case["fixed_mask"])
    ref.v[:] = case["v0"]
    ref.run(0.0, 0)  # initial force and acceleration
    outs = C.run(case, lambda dt: ref.run(dt, 1), lambda: (ref.u, ref.v, ref.a, ref.f), get_increment)
    return ref, mass, outs


def test_single_tet_regression_gold():
    case = C.load()
    _, mass, outs = _oracle_run(case)
    assert mass == pytest.approx(case["mass"], rel=GOLD_REL)
    assert len(outs) == case["output"]["count"]
    assert outs[-1][0] == pytest.approx(case["output"]["last_time"], rel=1e-12)
    got = C.maxima(outs)
    for name, g in case["maxima"].items():
        value, output, node = got[name]
        print(f"{name}: {value:.10e} at t{output},n{node}  gold {g['value']:.8e} at t{g['output']},n{g['node']}")
        assert value == pytest.approx(g["value"], rel=GOLD_REL), name
        assert (output, node) == (g["output"], g["node"]), name
    # the components the summary lists as zero stay exactly zero
    for name in ("displacement_x", "displacement_y", "velocity_x", "velocity_y", "acceleration_x", "acceleration_y"):
        assert got[name][0] == 0.0, name


def test_single_tet_power_method_increment_closed_form():
    """The input's power-method stepper: at t = 0 the only free DOF is the apex z, so
    lambda_max = k_zz / m with k_zz = V (lambda + 2 mu) (dN_apex/dz)^2 = V (lambda + 2 mu) / h^2
    (small-strain neo-Hookean = linear elastic).  The restated power method (finite
    difference with eps = 1e-5 times the max edge length) must land on it to O(eps)."""
    case = C.load()
    m = case["material"]
    lam, mu = O.lame(m["youngs_modulus"], m["poissons_ratio"])
    mass = O.lumped_mass(0, case["X"], case["conn"], m["density"])
    ref = O.Solver(0, case["X"], case["conn"], O.NEO_HOOKEAN, lam, mu, mass, case["fixed_mask"])
    l = O.max_edge_length(case["X"], case["conn"])
    eig = {}
    dt, it, cv = ref.power_method(l, u_in=np.zeros_like(case["X"]), eigenvector=eig)
    V, h = np.sqrt(3.0) / 4.0, 1.0
    dt_exact = 2.0 / np.sqrt(V * (lam + 2.0 * mu) / (h * h) / mass[0])
    assert cv and dt == pytest.approx(dt_exact, rel=2e-5)
    # the eigenvector is kept for the next call: it starts converged (one iteration,
    # previous increment passed in) and flips sign with each application of -M^-1 K
    v1 = eig["v"].copy()
    assert abs(abs(v1[2]) - 1.0) < 1e-12 and np.count_nonzero(v1) == 1
    dt2, it2, cv2 = ref.power_method(l, u_in=np.zeros_like(case["X"]), eigenvector=eig, previous_dt=dt)
    assert cv2 and it2 == 1 and dt2 == pytest.approx(dt_exact, rel=2e-5)
    assert eig["v"][2] == -v1[2]
    # 0.1 x that increment is shorter than the gold's 0.086222 by 3.7e-4
    assert 0.1 * dt == pytest.approx(0.0861897, rel=1e-6)
