"""The allocation-free power method (aphip_power_method_time_increment_noalloc) against the
allocating entry point through the C ABI: same iteration counts and convergence flags,
the same increments (to 1e-12), the same eigenvector carried to the next call and the
simulation state restored, over two calls in a row on two meshes, the second mesh
smaller than the first; a mesh whose work buffers exceed the scratch is refused."""
import numpy as np
import pytest

import aperi_hip as A
from aperi_hip import meshgen as M

pytestmark = pytest.mark.gpu

FIELDS = (A.FIELD_U, A.FIELD_F, A.FIELD_H, A.FIELD_P)


def _case(etype, n, kind, seed):
    X, hc = M.hex_lattice(n, n, n)
    X = M.jitter(X, 1.0 / n, 0.15, seed, 1)
    conn = hc if etype == A.HEX8 else M.split_tets(X, hc)
    fixed = np.zeros(X.shape, dtype=np.uint8)
    fixed[X[:, 2] < 1e-12] = 1
    u = 1e-3 * (2.0 * M.uniform01(seed, 5, X.size).reshape(X.shape) - 1.0)
    u[fixed.astype(bool)] = 0.0
    return dict(etype=etype, X=X, conn=conn, kind=kind, fixed=fixed, u=u)


def _solver(case):
    s = A.Solver()
    s.set_mesh(case["etype"], case["X"], case["conn"])
    s.set_material(case["kind"], 7850.0, 200e9, 0.3)
    s.compute_mass()
    s.set_fixed(case["fixed"])
    s.set_field(A.FIELD_U, case["u"])
    s.internal_force()
    return s


CALLS = (dict(), dict(num_iterations=200, tolerance=1e-10, variation_tolerance=0.0))


def test_no_alloc_power_method_matches_allocating_entry_point(gpu_available):
    big = _case(A.TET4, 5, A.NEO_HOOKEAN, 31)     # 216 nodes, 750 Tet4
    small = _case(A.HEX8, 3, A.LINEAR_ELASTIC, 37)  # 64 nodes, 27 Hex8
    print(f"no-alloc power method: Tet4 {len(big['X'])} nodes, then Hex8 {len(small['X'])} nodes", flush=True)
    ref = {name: _solver(c) for name, c in (("big", big), ("small", small))}
    na = {name: _solver(c) for name, c in (("big", big), ("small", small))}
    try:
        before = {name: {f: s.get(f) for f in FIELDS} for name, s in na.items()}
        # call k on the big mesh, then call k on the small one, for k = 1, 2 (the second
        # call starts from the eigenvector the first one left in its context)
        for kw in CALLS:
            for name in ("big", "small"):
                dt_r, it_r, cv_r = ref[name].power_method_time_increment(**kw)
                dt_n, it_n, cv_n = na[name].power_method_time_increment(no_alloc=True, **kw)
                print(f"  {name} {kw or 'defaults'}: dt {dt_n:.17g} vs {dt_r:.17g}, iterations {it_n} vs {it_r}",
                      flush=True)
                assert (it_n, cv_n) == (it_r, cv_r)
                assert dt_n == pytest.approx(dt_r, rel=1e-12, abs=0.0)
                for f in FIELDS:
                    assert np.array_equal(na[name].get(f), before[name][f]), (name, f)
    finally:
        for s in list(ref.values()) + list(na.values()):
            s.close()


def test_no_alloc_power_method_refuses_a_mesh_beyond_its_scratch(gpu_available):
    # 50^3 hexes x 6 = 750,000 Tet4: the H and P work buffers alone need 108 MB, more than the 64 MiB
    X, hc = M.hex_lattice(50, 50, 50)
    with A.Solver() as s:
        s.set_mesh(A.TET4, X, M.split_tets(X, hc))
        s.set_material(A.LINEAR_ELASTIC, 7850.0, 200e9, 0.3)
        s.compute_mass()
        with pytest.raises(A.AphipError, match="MiB of scratch"):
            s.power_method_time_increment(no_alloc=True)
