"""Pins the oracle's small dense math (oracle/or_math.c) against the reference's own
known answers and against independent mathematics."""
import math

import numpy as np
import pytest

import oracle as O

# test/unit_tests/LinearElasticPropertiesTest.cpp:18-20
E, NU, G, K, LAM = 200.0, 0.3, 76.92307692307692, 166.66666666666666, 115.38461538461539
CODES = {"E": 0, "nu": 1, "G": 2, "lambda": 3, "K": 4}
VALUES = {"E": E, "nu": NU, "G": G, "lambda": LAM, "K": K}


@pytest.mark.parametrize("pair", [("E", "nu"), ("E", "G"), ("E", "lambda"), ("E", "K"), ("nu", "G"),
                                  ("nu", "lambda"), ("nu", "K"), ("G", "lambda"), ("G", "K"), ("lambda", "K")])
def test_elastic_properties_all_pairs(pair):
    a, b = pair
    out = O.elastic_properties(CODES[a], VALUES[a], CODES[b], VALUES[b])
    for k, v in VALUES.items():
        assert out[k] == pytest.approx(v, rel=1e-9, abs=1e-3)  # reference tolerance is 1e-3 absolute


@pytest.mark.parametrize("a,b", [((0, -1.0), (1, 0.3)), ((0, 200.0), (1, 0.5)), ((0, 200.0), (1, -1.5)),
                                 ((2, -1.0), (1, 0.3))])
def test_elastic_properties_bounds(a, b):
    with pytest.raises(ValueError):
        O.elastic_properties(a[0], a[1], b[0], b[1])


def test_lame_matches_closed_form():
    lam, mu = O.lame(200e9, 0.3)
    assert lam == pytest.approx(200e9 * 0.3 / (1.3 * 0.4), rel=1e-15)
    assert mu == pytest.approx(200e9 / 2.6, rel=1e-15)


def test_det_ap_im1_identity():
    # MathUtilsTest.cpp:590-607: det(I + A) - 1 at 1e-12, random A
    rng = np.random.default_rng(3)
    for _ in range(50):
        A = rng.uniform(-1, 1, (3, 3)) * rng.choice([1e-6, 1e-2, 1.0])
        ref = np.linalg.det(np.eye(3) + A) - 1.0
        assert abs(O.det_ap_im1(A) - ref) <= 1e-12 * max(1.0, abs(ref))
    # the point of DetApIm1: accuracy at small H where det(I+H)-1 cancels
    A = np.diag([1e-12, 2e-12, -5e-13])
    assert O.det_ap_im1(A) == pytest.approx(2.5e-12, rel=1e-9)


def test_kernel_values():
    # MathUtilsTest.cpp:576-588: W(0)=1, W(R)=0, W(R/2)=0.25
    R = 1.7
    assert O.kernel_value([0, 0, 0], R) == 1.0
    assert O.kernel_value([R, 0, 0], R) == 0.0
    assert O.kernel_value([0, R / 2, 0], R) == pytest.approx(0.25, abs=1e-15)
    assert O.kernel_value([0, 0, 2 * R], R) == 0.0
    # cubic B-spline in q: both branches meet at q = 1/2 and W is C1 there
    for q in (0.1, 0.3, 0.49999, 0.5, 0.7, 0.95):
        w = O.kernel_value([q * R, 0, 0], R)
        ref = 1 - 6 * q * q + 6 * q ** 3 if q < 0.5 else 2 * (1 - q) ** 3
        assert w == pytest.approx(ref, abs=1e-14)


def test_inverse3_cofactor():
    rng = np.random.default_rng(4)
    for _ in range(50):
        A = rng.uniform(-1, 1, (3, 3)) + 2 * np.eye(3)
        Ai, det = O.inverse3_cofactor(A)
        assert np.allclose(Ai, np.linalg.inv(A), rtol=1e-13, atol=1e-13)
        assert det == pytest.approx(np.linalg.det(A), rel=1e-13)


@pytest.mark.parametrize("n", [3, 4])
def test_fullpivlu_inverse(n):
    rng = np.random.default_rng(5 + n)
    for _ in range(50):
        A = rng.uniform(-1, 1, (n, n)) + n * np.eye(n)
        Ai, rank = O.fullpivlu_inverse(A)
        assert rank == n
        assert np.allclose(Ai @ A, np.eye(n), atol=1e-13)
    # ill-conditioned moment-like matrix still inverts to working accuracy
    h = 1e-3
    pts = rng.uniform(-h, h, (12, 3))
    Hm = np.hstack([np.ones((12, 1)), pts])
    M = Hm.T @ Hm
    Mi, rank = O.fullpivlu_inverse(M)
    assert rank == 4
    assert np.allclose(Mi @ M, np.eye(4), atol=1e-8)


def test_fullpivlu_rank_deficient_zero_fill():
    # Eigen's FullPivLU::solve sets unknowns beyond the rank to zero, so the
    # "inverse" of diag(d, 0, 0) is diag(1/d, 0, 0) and of 0 is 0 (this is what
    # triggers the Shepard fallback, ShapeFunctionsFunctorReproducingKernel.h:229-237).
    Ai, rank = O.fullpivlu_inverse(np.diag([4.0, 0.0, 0.0]))
    assert rank == 1
    assert np.array_equal(Ai, np.diag([0.25, 0.0, 0.0]))
    Ai, rank = O.fullpivlu_inverse(np.zeros((4, 4)))
    assert rank == 0 and not Ai.any()
