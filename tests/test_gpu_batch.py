"""Ensembles: per-specimen results of one batched context are bitwise identical to
running every specimen alone; graph replay and autotuning do not change results."""
import numpy as np
import pytest

import oracle as O
import aperi_hip as A
from aperi_hip import meshgen as M
from conftest import rel_err

pytestmark = pytest.mark.gpu

FIELDS = (A.FIELD_U, A.FIELD_V, A.FIELD_A, A.FIELD_F, A.FIELD_H, A.FIELD_P)


def _specimens(n_spec=5, seed=11):
    specs = M.ensemble(n_spec, seed=seed, nmin=2, nmax=7)
    for s in specs:
        s["dt"] = 0.5 * M.stable_dt_estimate(s["X"], s["conn"], s["E"], s["nu"], s["rho"])
    return specs


def _alone(s, steps, graph=False):
    with A.Solver() as g:
        g.set_mesh(A.HEX8, s["X"], s["conn"])
        g.set_material(A.NEO_HOOKEAN, s["rho"], s["E"], s["nu"])
        g.compute_mass()
        g.set_field(A.FIELD_V, s["v0"])
        g.set_graph(graph)
        g.initialize()
        g.step(s["dt"], steps)
        return [g.get(f) for f in FIELDS]


def _batch(specs, steps, graph=False):
    b = A.Solver()
    b.batch_begin(A.HEX8, A.NEO_HOOKEAN)
    for s in specs:
        b.batch_add(s["X"], s["conn"], s["rho"], s["E"], s["nu"], s["dt"])
    b.batch_end()
    b.compute_mass()
    b.set_field(A.FIELD_V, np.concatenate([s["v0"] for s in specs]))
    b.set_graph(graph)
    b.initialize()
    b.step(0.0, steps)
    return b


def test_batch_bitwise_equals_per_specimen_loop(gpu_available):
    specs = _specimens()
    print("specimen edges", [s["n"] for s in specs], flush=True)
    with _batch(specs, 12) as b:
        got = [b.get(f) for f in FIELDS]
        for k, s in enumerate(specs):
            n0, nn, e0, ne = b.batch_ranges(k)
            ref = _alone(s, 12)
            for fi, (gf, rf) in enumerate(zip(got, ref)):
                part = gf[n0:n0 + nn] if fi < 4 else gf[e0:e0 + ne]
                assert np.array_equal(part, rf), (k, fi)


def test_batch_matches_oracle(gpu_available):
    specs = _specimens(3, seed=4)
    with _batch(specs, 10) as b:
        u = b.get(A.FIELD_U)
        for k, s in enumerate(specs):
            n0, nn, _, _ = b.batch_ranges(k)
            lam, mu = O.lame(s["E"], s["nu"])
            m = O.lumped_mass(1, s["X"], s["conn"], s["rho"])
            ref = O.Solver(1, s["X"], s["conn"], O.NEO_HOOKEAN, lam, mu, m)
            ref.v[:] = s["v0"]
            ref.run(s["dt"], 10)
            assert rel_err(u[n0:n0 + nn], ref.u) < 1e-12


def test_graph_replay_is_bitwise_eager(gpu_available):
    specs = _specimens(4, seed=7)
    with _batch(specs, 9, graph=False) as b1, _batch(specs, 9, graph=True) as b2:
        for f in FIELDS:
            assert np.array_equal(b1.get(f), b2.get(f))
    # several replays with eager work and synchronisation in between
    s = specs[0]
    a = _alone(s, 6, graph=False)
    with A.Solver() as g:
        g.set_mesh(A.HEX8, s["X"], s["conn"])
        g.set_material(A.NEO_HOOKEAN, s["rho"], s["E"], s["nu"])
        g.compute_mass()
        g.set_field(A.FIELD_V, s["v0"])
        g.set_graph(True)
        g.initialize()
        for _ in range(3):
            g.step(s["dt"], 2)
            g.synchronize()
            _ = g.get(A.FIELD_U)
        for f, r in zip(FIELDS, a):
            assert np.array_equal(g.get(f), r)


def test_autotune_restores_state(gpu_available):
    X, hc = M.hex_lattice(8, 8, 8)
    tets = M.split_tets(X, hc)
    v0 = M.normal(3, 3, X.size).reshape(X.shape)
    dt = 0.5 * M.stable_dt_estimate(X, tets, 200e9, 0.3, 7850.0)
    with A.Solver() as g, A.Solver() as h:
        for s in (g, h):
            s.set_mesh(A.TET4, X, tets)
            s.set_material(A.NEO_HOOKEAN, 7850.0, 200e9, 0.3)
            s.compute_mass()
            s.set_field(A.FIELD_V, v0)
            s.initialize()
            s.step(dt, 3)
        choice = g.autotune(dt, 3)
        te, threads = divmod(choice % 100000000, 10000)
        # the Tet4 candidates of aphip_autotune: one element per thread, 64 to 1024 per tile
        # (the pipelined kernel at 128 and 256)
        assert te == threads and te in (64, 128, 256, 512, 1024)
        assert choice < 100000000 or te in (128, 256)
        g.step(dt, 5)
        h.step(dt, 5)
        for f in FIELDS:
            assert rel_err(g.get(f), h.get(f)) < 1e-13


def test_autotune_keeps_each_specimens_increment(gpu_available):
    """Autotuning an ensemble renumbers its nodes; every node must keep its own specimen's
    time increment, so further steps agree with an untuned ensemble to rounding."""
    specs = _specimens(6, seed=21)
    assert len({s["dt"] for s in specs}) == len(specs)
    with _batch(specs, 3) as g, _batch(specs, 3) as h:
        for s in (g, h):
            s.set_graph(False)
        g.autotune(0.0, 2)
        g.step(0.0, 8)
        h.step(0.0, 8)
        for f in FIELDS:
            assert rel_err(g.get(f), h.get(f)) < 1e-12, f


@pytest.mark.parametrize("kind", [A.J2, A.SVK], ids=["j2", "svk"])
def test_autotune_restores_hex8_state(gpu_available, kind):
    X, hc = M.hex_lattice(9, 9, 9)
    X = M.jitter(X, 1.0 / 9, 0.1, 3, 1)
    v0 = 40.0 * M.normal(5, 3, X.size).reshape(X.shape)
    dt = 0.5 * M.stable_dt_estimate(X, hc, 200e9, 0.3, 7850.0)
    with A.Solver() as g, A.Solver() as h:
        for s in (g, h):
            s.set_mesh(A.HEX8, X, hc)
            s.set_material(kind, 7850.0, 200e9, 0.3, 250e6 if kind == A.J2 else 0.0,
                           1e9 if kind == A.J2 else 0.0)
            s.compute_mass()
            s.set_field(A.FIELD_V, v0)
            s.initialize()
            s.step(dt, 4)
        g.autotune(dt, 2)
        g.step(dt, 6)
        h.step(dt, 6)
        for f in FIELDS:
            assert rel_err(g.get(f), h.get(f)) < 1e-12, f
        if kind == A.J2:
            st = h.get(A.FIELD_STATE)
            assert st.max() > 0.0 and rel_err(g.get(A.FIELD_STATE), st) < 1e-12
