"""Exodus II results files (the reference's field output, src/IoMesh.cpp:190-240),
read back with scipy's independent netCDF reader.

CPU: the file encoder itself (netCDF 64-bit-offset header, non-record data, records,
Exodus II names and 1-based connectivity) on a known two-Hex8 sample written by the
C++ test program.
GPU: the library's results file of the reference's single-tet regression run
(test/regression_tests/single_tet/fem): structure, every stored value against the
fields read through the C ABI at the same output, and the maxima the reference's
exodiff summary lists (compare.exodiff:26-52) computed from the file itself."""
import os
import subprocess

import numpy as np
import pytest
from scipy.io import netcdf_file

import single_tet_case as C

HERE = os.path.dirname(os.path.abspath(__file__))
BIN = os.path.join(HERE, "cpp", "test_adapter")


def _names(var):
    return [b"".join(row).decode().rstrip("\x00") for row in var.data]


def test_exodus_encoder_sample(tmp_path):
    if not os.path.exists(BIN):
        pytest.fail(f"{BIN} is not built; run __graft_entry__.build() or make -C tests/cpp")
    path = str(tmp_path / "sample.exo")
    subprocess.run([BIN, "exodus", path], check=True, timeout=60)
    with netcdf_file(path, "r", mmap=False) as f:
        assert f.version_byte == 2  # 64-bit offset
        d = f.dimensions
        assert (d["num_nodes"], d["num_elem"], d["num_el_blk"], d["num_nod_per_el1"], d["num_dim"]) == (12, 2, 1, 8, 3)
        assert d["time_step"] is None  # the record dimension
        assert f.title == b"two hexes" and f.floating_point_word_size == 8 and f.maximum_name_length == 32
        v = f.variables
        assert v["connect1"].elem_type == b"HEX8"
        np.testing.assert_array_equal(v["connect1"].data, [[1, 2, 5, 4, 7, 8, 11, 10], [2, 3, 6, 5, 8, 9, 12, 11]])
        i = np.arange(12)
        np.testing.assert_array_equal(v["coordx"].data, i % 3 * 0.5)
        np.testing.assert_array_equal(v["coordz"].data, i // 6 * 0.5)
        assert _names(v["name_nod_var"]) == ["displacement_x", "force_z"]
        assert _names(v["name_elem_var"]) == ["pk1_stress_xy", "volume"]
        assert _names(v["coor_names"]) == ["x", "y", "z"] and _names(v["eb_names"]) == ["block_1"]
        np.testing.assert_array_equal(v["time_whole"].data, [0.0, 0.25, 0.5])
        for s in range(3):
            np.testing.assert_array_equal(v["vals_nod_var1"].data[s], 1e-3 * s * i)
            np.testing.assert_array_equal(v["vals_nod_var2"].data[s], -0.5 * s + i)
            np.testing.assert_array_equal(v["vals_elem_var1eb1"].data[s], [1.0 + s, -2.0 - s])
            np.testing.assert_array_equal(v["vals_elem_var2eb1"].data[s], [0.125, 0.125])
        np.testing.assert_array_equal(v["elem_var_tab"].data, [[1, 1]])


@pytest.mark.gpu
def test_results_file_of_single_tet_regression(gpu_available, tmp_path):
    import aperi_hip as A

    case = C.load()
    m = case["material"]
    path = str(tmp_path / "results.exo")
    print("results file of the single-tet regression", flush=True)
    with A.Solver() as s:
        s.set_mesh(A.TET4, case["X"], case["conn"])
        s.set_material(A.NEO_HOOKEAN, m["density"], m["youngs_modulus"], m["poissons_ratio"])
        s.compute_mass()
        s.set_fixed(case["fixed_mask"])
        s.set_field(A.FIELD_V, case["v0"])
        s.initialize()
        s.results_open(path, "single tet")
        clock = {"t": 0.0}
        seen = []

        def step(dt):
            s.step(dt, 1)
            clock["t"] += dt  # the same additions as the loop's own clock

        def output():
            s.results_write(clock["t"])
            got = tuple(s.get(f) for f in (A.FIELD_U, A.FIELD_V, A.FIELD_A, A.FIELD_F))
            seen.append(got + (s.get(A.FIELD_H), s.get(A.FIELD_P)))
            return got

        outs = C.run(case, step, output)
        s.results_close()
    with netcdf_file(path, "r", mmap=False) as f:
        v = f.variables
        assert f.dimensions["num_nodes"] == 4 and f.dimensions["num_elem"] == 1 and f.title == b"single tet"
        assert v["connect1"].elem_type == b"TETRA4"
        np.testing.assert_array_equal(v["connect1"].data, case["conn"] + 1)
        np.testing.assert_array_equal(np.stack([v["coordx"].data, v["coordy"].data, v["coordz"].data], 1), case["X"])
        names = _names(v["name_nod_var"])
        enames = _names(v["name_elem_var"])
        T = v["time_whole"].data
        assert len(T) == case["output"]["count"] and T[-1] == pytest.approx(case["output"]["last_time"], rel=1e-12)
        np.testing.assert_array_equal(T, [o[0] for o in outs])

        def nod(name):
            return v[f"vals_nod_var{names.index(name) + 1}"].data

        def ele(name):
            return v[f"vals_elem_var{enames.index(name) + 1}eb1"].data

        for k, fld in enumerate(("displacement", "velocity", "acceleration", "force")):
            for d, c in enumerate("xyz"):
                np.testing.assert_array_equal(nod(f"{fld}_{c}"), np.stack([o[1 + k][:, d] for o in outs]))
        for c in "xyz":  # one scalar per node, the reference's three equal components
            assert nod(f"mass_{c}") == pytest.approx(np.sqrt(3.0) / 16.0, rel=1e-14)
        np.testing.assert_array_equal(nod("essential_boundary_x")[0], [1, 1, 1, 1])
        np.testing.assert_array_equal(nod("essential_boundary_z")[0], [0, 1, 1, 1])
        for q, fld in enumerate(("displacement_gradient", "pk1_stress")):
            for k, sfx in enumerate(("xx", "xy", "xz", "yx", "yy", "yz", "zx", "zy", "zz")):
                np.testing.assert_array_equal(ele(f"{fld}_{sfx}")[:, 0], [o[4 + q][0, k] for o in seen])
        assert ele("volume")[0, 0] == pytest.approx(np.sqrt(3.0) / 4.0, rel=1e-14)
        # what exodiff compares against the gold: its maxima, recomputed from the file
        for name, g in case["maxima"].items():
            a = np.abs(nod(name))
            i, j = np.unravel_index(np.argmax(a), a.shape)
            assert a[i, j] == pytest.approx(g["value"], rel=5e-8), name
            assert (i + 1, j + 1) == (g["output"], g["node"]), name
