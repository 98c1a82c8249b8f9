"""GPU parity of nodal strain smoothing on thex meshes, with subcells and F-bar
(aphip_rk_nodal_setup; the reference's nodal_smoothing_cell integration, include/
MeshLabelerProcessor.h:227-330, :690-827, include/SmoothedCellDataProcessor.h:407-675,
include/ComputeInternalForceSmoothedCell.h:120-413) against the oracle through the C ABI.

Bars: active nodes, neighbor lists, cells and subcell node lists identical; RK functions,
subcell derivatives and volumes, mass, force, H, P and state within 1e-12 relative after one
evaluation; every field within 1e-10 after 100 steps; nodes without degrees of freedom
untouched by the steps.
"""
import numpy as np
import pytest

import oracle as O
import aperi_hip as A
from aperi_hip import meshgen as M
from conftest import rel_err

pytestmark = pytest.mark.gpu

RHO, E, NU = 2700.0, 70e9, 0.33
J2 = dict(yield_stress=300e6, hardening=1e9)
SCALE = 0.9


def _thex(n=3, seed=7):
    return M.thex_block(n, seed)


def _elem_subcell(d, ne):
    es = np.empty(ne, dtype=np.int64)
    ss, se = d["sub_start"], d["sub_elems"]
    for s in range(len(ss) - 1):
        es[se[ss[s]:ss[s + 1]]] = s
    return es


def _gpu(Xh, hx, kind, subdomains, f_bar):
    s = A.Solver()
    s.set_mesh(A.HEX8, Xh, hx)
    s.set_material(kind, RHO, E, NU, J2["yield_stress"] if kind == A.J2 else 0.0,
                   J2["hardening"] if kind == A.J2 else 0.0)
    s.rk_nodal_setup(SCALE, subdomains, f_bar)
    return s


@pytest.mark.parametrize("subdomains", [0, 1])
def test_nodal_setup_matches_oracle(gpu_available, subdomains):
    Xh, hx = _thex()
    d = O.nodal_setup(Xh, hx, SCALE, subdomains, density=RHO)
    print(f"nodal set-up subdomains={subdomains}: {len(Xh)} nodes, {len(hx)} hexes, "
          f"{len(d['cell_sub']) - 1} cells", flush=True)
    with _gpu(Xh, hx, A.J2, subdomains, True) as s:
        cs, es, act, fbar = s.rk_nodal_cells()
        assert np.array_equal(act, d["active"])
        assert np.array_equal(cs, d["cell_sub"]) and np.array_equal(es, _elem_subcell(d, len(hx)))
        assert fbar == (subdomains == 0)  # F-bar only where cells hold several subcells
        num, nbr, tr = s.rk_neighbors()
        valid = np.arange(40)[None, :] < d["num"][:, None]
        assert np.array_equal(num, d["num"]) and tr == d["trunc"]
        assert np.array_equal(nbr, np.where(valid, d["nbr"], -1))
        assert np.abs(s.rk_function_values() - d["phi"]).max() < 1e-12
        start, nodes, dphi, vol = s.rk_cells()
        assert np.array_equal(start, d["cells"][0]) and np.array_equal(nodes, d["cells"][1])
        assert rel_err(dphi, d["cells"][2]) < 1e-12 and rel_err(vol, d["cells"][3]) < 1e-12
        assert rel_err(s.get(A.FIELD_VOLUME), d["elem_volume"]) < 1e-12
        s.compute_mass()
        m = s.get(A.FIELD_MASS)
        assert rel_err(m, d["mass"]) < 1e-13 and np.all(m[~act.astype(bool)] == 0.0)


@pytest.mark.parametrize("kind,subdomains,f_bar", [(A.NEO_HOOKEAN, 0, True), (A.J2, 0, True),
                                                    (A.LINEAR_ELASTIC, 0, True), (A.SVK, 0, True),
                                                    (A.NEO_HOOKEAN, 0, False), (A.J2, 1, False)],
                         ids=["nh-fbar", "j2-fbar", "le-fbar", "svk-fbar", "nh-subcells", "j2-cells"])
def test_nodal_force_matches_oracle(gpu_available, kind, subdomains, f_bar):
    Xh, hx = _thex(3, 11)
    d = O.nodal_setup(Xh, hx, SCALE, subdomains)
    lam, mu = O.lame(E, NU)
    okind = {A.NEO_HOOKEAN: O.NEO_HOOKEAN, A.J2: O.J2, A.LINEAR_ELASTIC: O.LINEAR_ELASTIC, A.SVK: O.SVK}[kind]
    kw = dict(yield_stress=J2["yield_stress"], hardening=J2["hardening"]) if kind == A.J2 else {}
    u = 4e-3 * (2 * M.uniform01(8, 3, Xh.size).reshape(Xh.shape) - 1)
    if f_bar:
        f, H, _, P, _, st = O.fbar_internal_force(len(Xh), d["cells"], (d["cell_sub"], d["cell_volume"]), okind,
                                                  lam, mu, u, **kw)
    else:
        f, H, P, st = O.cell_internal_force(len(Xh), d["cells"], okind, lam, mu, u, **kw)
    es = _elem_subcell(d, len(hx))
    print(f"nodal force kind={kind} subdomains={subdomains} f_bar={f_bar}", flush=True)
    with _gpu(Xh, hx, kind, subdomains, f_bar) as s:
        s.set_field(A.FIELD_U, u)
        s.internal_force(0.0, 1e-7)
        assert rel_err(s.get(A.FIELD_F), f) < 1e-12
        assert rel_err(s.get(A.FIELD_H), H[es]) < 1e-12
        assert rel_err(s.get(A.FIELD_P), P[es]) < 1e-12
        if kind == A.J2:
            assert st.max() > 0.0
            assert rel_err(s.get(A.FIELD_STATE), st[es]) < 1e-12


@pytest.mark.parametrize("kind", [A.NEO_HOOKEAN, A.J2], ids=["nh", "j2"])
def test_nodal_fbar_100_steps(gpu_available, kind):
    Xh, hx = _thex(3, 13)
    d = O.nodal_setup(Xh, hx, SCALE, 0, density=RHO)
    lam, mu = O.lame(E, NU)
    act = d["active"].astype(bool)
    okind = O.NEO_HOOKEAN if kind == A.NEO_HOOKEAN else O.J2
    kw = dict(yield_stress=J2["yield_stress"], hardening=J2["hardening"]) if kind == A.J2 else {}
    ref = O.Solver(-1, Xh, None, okind, lam, mu, d["mass"], None, cells=d["cells"],
                   fbar_cells=(d["cell_sub"], d["cell_volume"]), active=d["active"], **kw)
    dt = 0.5 * M.stable_dt_estimate(Xh, hx, E, NU, RHO)
    c = Xh - 0.5
    v0 = (0.02 / (100 * dt)) * np.stack([-0.5 * c[:, 0], -0.5 * c[:, 1], c[:, 2]], axis=1)
    ref.v[:] = v0
    ref.run(dt, 100)
    es = _elem_subcell(d, len(hx))
    print(f"nodal F-bar 100 steps kind={kind}", flush=True)
    with _gpu(Xh, hx, kind, 0, True) as s:
        s.compute_mass()
        s.set_field(A.FIELD_V, v0)
        s.initialize()
        s.step(dt, 100)
        for fld, r in ((A.FIELD_U, ref.u), (A.FIELD_V, ref.v), (A.FIELD_A, ref.a), (A.FIELD_F, ref.f)):
            assert rel_err(s.get(fld), r) < 1e-10, fld
        for fld, r in ((A.FIELD_H, ref.H), (A.FIELD_P, ref.P)):
            assert rel_err(s.get(fld), r[es]) < 1e-10, fld
        if kind == A.J2:
            assert ref.state.max() > 0.0
            assert rel_err(s.get(A.FIELD_STATE), ref.state[es]) < 1e-10
        # nodes without degrees of freedom keep their initial values
        assert np.array_equal(s.get(A.FIELD_V)[~act], v0[~act])
        assert np.all(s.get(A.FIELD_U)[~act] == 0.0)


def test_nodal_refusals(gpu_available):
    X, hc = M.hex_lattice(3, 3, 3)
    with A.Solver() as s:
        s.set_mesh(A.HEX8, X, hc)
        s.set_material(A.NEO_HOOKEAN, RHO, E, NU)
        with pytest.raises(A.AphipError, match="exactly one active node"):
            s.rk_nodal_setup(SCALE, 0, True)
    Xh, hx = _thex(2, 3)
    with A.Solver() as s:
        s.set_mesh(A.HEX8, Xh, hx)
        s.set_material(A.POWER_LAW_CREEP, RHO, E, NU, creep=dict(A=1e-18, n=3.0, m=0.0, constant_temperature=0.0))
        with pytest.raises(A.AphipError, match="F-bar takes"):
            s.rk_nodal_setup(SCALE, 0, True)
    with _gpu(Xh, hx, A.NEO_HOOKEAN, 0, True) as s:
        s.compute_mass()
        with pytest.raises(A.AphipError, match="nodal cells"):
            s.power_method_time_increment()
        with pytest.raises(A.AphipError, match="nodal cells"):
            s.set_material_fields(np.full(len(hx), E), np.full(len(hx), NU))


@pytest.mark.parametrize("tile_cells", ["8", "64"])
def test_fbar_tiles_are_deterministic_and_size_independent(gpu_available, monkeypatch, tile_cells):
    """The F-bar force runs one tile kernel per tile of whole nodal cells (at most
    APHIP_RK_TILE_CELLS subcells, a larger cell alone).  Two runs give the same bits, and
    the tile size changes nothing beyond rounding: 30 J2 steps match the oracle at 1e-10."""
    Xh, hx = _thex(3, 23)
    d = O.nodal_setup(Xh, hx, SCALE, 0, density=RHO)
    lam, mu = O.lame(E, NU)
    ref = O.Solver(-1, Xh, None, O.J2, lam, mu, d["mass"], None, cells=d["cells"],
                   fbar_cells=(d["cell_sub"], d["cell_volume"]), active=d["active"], **J2)
    dt = 0.5 * M.stable_dt_estimate(Xh, hx, E, NU, RHO)
    c = Xh - 0.5
    v0 = (0.03 / (30 * dt)) * np.stack([-0.5 * c[:, 0], -0.5 * c[:, 1], c[:, 2]], axis=1)
    ref.v[:] = v0
    ref.run(dt, 30)
    es = _elem_subcell(d, len(hx))
    monkeypatch.setenv("APHIP_RK_TILE_CELLS", tile_cells)
    print(f"F-bar tiles of at most {tile_cells} subcells", flush=True)
    out = []
    for _ in range(2):
        with _gpu(Xh, hx, A.J2, 0, True) as s:
            s.compute_mass()
            s.set_field(A.FIELD_V, v0)
            s.initialize()
            s.step(dt, 30)
            out.append([s.get(f) for f in (A.FIELD_U, A.FIELD_V, A.FIELD_F, A.FIELD_P, A.FIELD_STATE)])
    for a, b in zip(*out):
        assert np.array_equal(a, b)
    u, v, f, P, st = out[0]
    assert ref.state.max() > 0.0
    for g, r in ((u, ref.u), (v, ref.v), (f, ref.f)):
        assert rel_err(g, r) < 1e-10
    assert rel_err(P, ref.P[es]) < 1e-10 and rel_err(st, ref.state[es]) < 1e-10
