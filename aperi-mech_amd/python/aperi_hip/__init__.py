"""Python host bindings of the aperi-hip C ABI (include/aperi_hip.h).

The compute lives in libaperi_hip.so (hand-written gfx950 HIP kernels); this
module only marshals numpy arrays through ctypes.  There is no CPU fallback:
if the library is missing, loading fails with an error.

The `Solver` class mirrors the calling sequence of aperi-mech's ExplicitSolver
(src/Solvers/Explicit.cpp:308-452): mesh + material (element block and
Material), mass (BuildMassMatrix), initial force/acceleration, then fixed-dt
steps; `internal_force` mirrors ElementBase::ComputeInternalForceAllElements.
"""
import ctypes
import os

import numpy as np

from . import meshgen  # noqa: F401

_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.normpath(os.path.join(_PKG, "..", "..", "lib", "libaperi_hip.so"))

TET4, HEX8 = 0, 1
LINEAR_ELASTIC, NEO_HOOKEAN, J2, SVK, NEO_HOOKEAN_DAMAGE, POWER_LAW_CREEP, DRUCKER_PRAGER = 0, 1, 2, 3, 4, 5, 6
# optional neo-Hookean-with-damage keys and the reference's defaults (Materials/Factory.h:78-95)
DAMAGE_DEFAULTS = dict(maximum_damage=0.99, maximum_damage_increment=1.0, stop_failing_at_time=-1.0,
                       fail_all_at_time=-1.0)
FIELD_U, FIELD_V, FIELD_A, FIELD_F, FIELD_MASS, FIELD_H, FIELD_P, FIELD_STATE, FIELD_VOLUME = range(9)
FIELD_RK_DISPLACEMENT, FIELD_RK_FORCE_LOCAL = 9, 10
TOTAL_LAGRANGIAN, UPDATED_LAGRANGIAN, SEMI_LAGRANGIAN = 0, 2, 3
MAX_NODE_NUM_NEIGHBORS = 40
FLAG_TWO_PASS = 1
FLAG_PIPELINED = 2

_d = ctypes.POINTER(ctypes.c_double)
_i32 = ctypes.POINTER(ctypes.c_int32)
_i64 = ctypes.POINTER(ctypes.c_int64)
_u8 = ctypes.POINTER(ctypes.c_uint8)

EXPORTED = [
    "aphip_create", "aphip_destroy", "aphip_last_error", "aphip_device_count", "aphip_set_mesh", "aphip_get_node_permutation",
    "aphip_set_material", "aphip_set_material_fields", "aphip_set_formulation", "aphip_set_formulation_semi", "aphip_compute_mass", "aphip_set_fixed",
    "aphip_set_field", "aphip_get_field", "aphip_set_field_device", "aphip_get_field_device",
    "aphip_internal_force", "aphip_initialize", "aphip_step",
    "aphip_synchronize", "aphip_set_tiling", "aphip_set_graph", "aphip_autotune", "aphip_batch_begin",
    "aphip_batch_add", "aphip_batch_end", "aphip_batch_ranges", "aphip_internal_force_vjp",
    "aphip_adjoint_forward", "aphip_adjoint_backward", "aphip_rk_setup", "aphip_rk_setup_method", "aphip_strain_smoothing_setup", "aphip_rk_nodal_setup", "aphip_rk_get_nodal_cells", "aphip_rk_get_neighbors", "aphip_rk_get_function_values",
    "aphip_rk_get_cells", "aphip_set_profiling", "aphip_get_profile", "aphip_reset_profile",
    "aphip_bandwidth_probe", "aphip_power_method_time_increment", "aphip_power_method_time_increment_noalloc",
    "aphip_results_open", "aphip_results_write", "aphip_results_close",
]


class Config(ctypes.Structure):
    _fields_ = [("device", ctypes.c_int32), ("flags", ctypes.c_int32)]


class MaterialParams(ctypes.Structure):
    _fields_ = [("density", ctypes.c_double), ("youngs_modulus", ctypes.c_double),
                ("poissons_ratio", ctypes.c_double), ("yield_stress", ctypes.c_double),
                ("hardening_modulus", ctypes.c_double), ("state_per_qp", ctypes.c_int32),
                ("reserved", ctypes.c_int32),
                ("I1_critical", ctypes.c_double), ("I1_failure", ctypes.c_double), ("alpha", ctypes.c_double),
                ("maximum_damage", ctypes.c_double), ("maximum_damage_increment", ctypes.c_double),
                ("stop_failing_at_time", ctypes.c_double), ("fail_all_at_time", ctypes.c_double),
                ("creep_A", ctypes.c_double), ("creep_n", ctypes.c_double), ("creep_m", ctypes.c_double),
                ("constant_temperature", ctypes.c_double), ("use_constant_temperature", ctypes.c_int32),
                ("reserved2", ctypes.c_int32),
                ("dp_A1", ctypes.c_double), ("dp_A2", ctypes.c_double), ("dp_A2G", ctypes.c_double)]


def material_params(density, E, nu, yield_stress=0.0, hardening=0.0, state_per_qp=0, damage=None, creep=None,
                    drucker_prager=None):
    """aphip_material_params; damage: dict with I1_critical, I1_failure, alpha and optionally
    maximum_damage, maximum_damage_increment, stop_failing_at_time, fail_all_at_time;
    creep: dict with A, n, m and constant_temperature (the reference requires one);
    drucker_prager: dict with A1, A2, A2G."""
    p = MaterialParams(density, E, nu, yield_stress, hardening, state_per_qp, 0)
    if damage is not None:
        for k, v in dict(DAMAGE_DEFAULTS, **damage).items():
            setattr(p, k, v)
    if creep is not None:
        p.creep_A, p.creep_n, p.creep_m = creep["A"], creep["n"], creep["m"]
        if "constant_temperature" in creep:
            p.constant_temperature, p.use_constant_temperature = creep["constant_temperature"], 1
    if drucker_prager is not None:
        p.dp_A1, p.dp_A2, p.dp_A2G = drucker_prager["A1"], drucker_prager["A2"], drucker_prager["A2G"]
    return p


class AphipError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"aperi-hip error {code}: {msg}")
        self.code = code


_LIB = None


def load_library(path=LIB_PATH):
    """Load libaperi_hip.so (raises if it has not been built)."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(path):
        raise FileNotFoundError(f"{path} not built; run __graft_entry__.build() or make -C aperi-mech_amd")
    L = ctypes.CDLL(path)
    L.aphip_last_error.restype = ctypes.c_char_p
    for name in EXPORTED:
        fn = getattr(L, name)
        if name != "aphip_last_error":
            fn.restype = ctypes.c_int
    _LIB = L
    return L


def device_count():
    """Visible HIP devices, as reported by the library's own HIP runtime."""
    L = load_library()
    n = ctypes.c_int32()
    _check(L.aphip_device_count(ctypes.byref(n)))
    return n.value


def bandwidth_probe(nbytes, reps=20, device=0):
    """Device milliseconds per device-to-device copy of `nbytes` (multiple of 16)."""
    L = load_library()
    ms = ctypes.c_double()
    _check(L.aphip_bandwidth_probe(ctypes.c_int32(device), ctypes.c_int64(nbytes), ctypes.c_int32(reps),
                                   ctypes.byref(ms)))
    return ms.value


def _check(rc):
    if rc != 0:
        raise AphipError(rc, _LIB.aphip_last_error().decode())


def _p(a, t=_d):
    return None if a is None else a.ctypes.data_as(t)


class Solver:
    """One element block on one GPU behind the C ABI."""

    def __init__(self, device=0, flags=0):
        self.L = load_library()
        self.ctx = ctypes.c_void_p()
        cfg = Config(device, flags)
        _check(self.L.aphip_create(ctypes.byref(self.ctx), ctypes.byref(cfg)))
        self.nn = self.ne = 0
        self.etype = None
        self.kind = None
        self.state_per_qp = 0

    def close(self):
        if self.ctx:
            self.L.aphip_destroy(self.ctx)
            self.ctx = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # ---- set-up ----
    def set_mesh(self, etype, X, conn):
        X = np.ascontiguousarray(X, dtype=np.float64)
        conn = np.ascontiguousarray(conn, dtype=np.int32)
        self.nn, self.ne, self.etype = X.shape[0], conn.shape[0], etype
        _check(self.L.aphip_set_mesh(self.ctx, etype, self.nn, _p(X), self.ne, _p(conn, _i32)))
        self._nodal = False
        return self

    def node_permutation(self):
        out = np.zeros(self.nn, dtype=np.int32)
        _check(self.L.aphip_get_node_permutation(self.ctx, _p(out, _i32)))
        return out

    def set_material(self, kind, density, E, nu, yield_stress=0.0, hardening=0.0, state_per_qp=0, damage=None,
                     creep=None, drucker_prager=None):
        p = material_params(density, E, nu, yield_stress, hardening, state_per_qp, damage, creep, drucker_prager)
        _check(self.L.aphip_set_material(self.ctx, kind, ctypes.byref(p)))
        self.kind, self.state_per_qp = kind, state_per_qp
        return self

    def set_formulation(self, formulation):
        """TOTAL_LAGRANGIAN or UPDATED_LAGRANGIAN (after set_mesh and set_material)."""
        _check(self.L.aphip_set_formulation(self.ctx, int(formulation)))
        return self

    def set_semi_lagrangian(self, update_interval):
        """Semi-Lagrangian with a reference-configuration update every update_interval steps."""
        _check(self.L.aphip_set_formulation_semi(self.ctx, int(update_interval)))
        return self

    def set_material_fields(self, E=None, nu=None):
        if E is None:
            _check(self.L.aphip_set_material_fields(self.ctx, None, None))
            return self
        E = np.ascontiguousarray(E, dtype=np.float64)
        nu = np.ascontiguousarray(nu, dtype=np.float64)
        _check(self.L.aphip_set_material_fields(self.ctx, _p(E), _p(nu)))
        return self

    def compute_mass(self):
        tot = ctypes.c_double()
        _check(self.L.aphip_compute_mass(self.ctx, ctypes.byref(tot)))
        return tot.value

    def set_fixed(self, mask):
        if mask is None:
            _check(self.L.aphip_set_fixed(self.ctx, None))
            return self
        m = np.ascontiguousarray(mask, dtype=np.uint8)
        _check(self.L.aphip_set_fixed(self.ctx, _p(m, _u8)))
        return self

    def set_field(self, field, data):
        d = np.ascontiguousarray(data, dtype=np.float64)
        _check(self.L.aphip_set_field(self.ctx, field, _p(d)))
        return self

    def set_field_device(self, field, ptr):
        """Field from a device array (integer device address, caller numbering), queued on
        the context's stream."""
        _check(self.L.aphip_set_field_device(self.ctx, field, ctypes.c_void_p(int(ptr))))
        return self

    def get_field_device(self, field, ptr):
        """Field into a device array (integer device address, caller numbering), queued on
        the context's stream; complete after synchronize()."""
        _check(self.L.aphip_get_field_device(self.ctx, field, ctypes.c_void_p(int(ptr))))
        return self

    def field_shape(self, field):
        return self._shape(field)

    def _shape(self, field):
        if field in (FIELD_U, FIELD_V, FIELD_A, FIELD_F, FIELD_RK_DISPLACEMENT, FIELD_RK_FORCE_LOCAL):
            return (self.nn, 3)
        if field == FIELD_MASS:
            return (self.nn,)
        if field in (FIELD_H, FIELD_P):
            return (self.ne, 9)
        if field == FIELD_STATE:
            width = 2 if self.kind in (NEO_HOOKEAN_DAMAGE, POWER_LAW_CREEP, DRUCKER_PRAGER) else 1
            return (width * self.ne * (8 if (self.etype == HEX8 and self.state_per_qp) else 1),)
        return (self.ne,)

    def get(self, field):
        out = np.zeros(self._shape(field))
        _check(self.L.aphip_get_field(self.ctx, field, _p(out)))
        return out

    def power_method_time_increment(self, num_iterations=50, epsilon=1e-5, tolerance=1e-3,
                                    variation_tolerance=1e-7, seed=20261202, previous_time_increment=0.0,
                                    no_alloc=False):
        """(dt, iterations, converged) of the power-method stable increment; the start
        vector is the eigenvector of the previous call (random on the first call).
        no_alloc: the entry point that allocates no device memory (meshes up to its
        scratch size)."""
        dt, it, cv = ctypes.c_double(), ctypes.c_int32(), ctypes.c_int32()
        fn = (self.L.aphip_power_method_time_increment_noalloc if no_alloc
              else self.L.aphip_power_method_time_increment)
        _check(fn(self.ctx, ctypes.c_double(previous_time_increment),
                                                        int(num_iterations), ctypes.c_double(epsilon),
                                                        ctypes.c_double(tolerance),
                                                        ctypes.c_double(variation_tolerance),
                                                        ctypes.c_uint64(seed), ctypes.byref(dt), ctypes.byref(it),
                                                        ctypes.byref(cv)))
        return dt.value, it.value, bool(cv.value)

    # ---- Exodus II results file ----
    def results_open(self, path, title=""):
        """Create an Exodus II results file of this mesh (caller numbering)."""
        _check(self.L.aphip_results_open(self.ctx, os.fsencode(os.fspath(path)), title.encode()))
        return self

    def results_write(self, time):
        """Append the current fields as one time step."""
        _check(self.L.aphip_results_write(self.ctx, ctypes.c_double(time)))
        return self

    def results_close(self):
        _check(self.L.aphip_results_close(self.ctx))
        return self

    # ---- hot path ----
    def internal_force(self, time=0.0, dt=0.0):
        _check(self.L.aphip_internal_force(self.ctx, ctypes.c_double(time), ctypes.c_double(dt)))
        return self

    def initialize(self):
        _check(self.L.aphip_initialize(self.ctx))
        return self

    def step(self, dt, n_steps=1):
        _check(self.L.aphip_step(self.ctx, ctypes.c_double(dt), int(n_steps)))
        return self

    def set_tiling(self, elems_per_tile, threads):
        _check(self.L.aphip_set_tiling(self.ctx, int(elems_per_tile), int(threads)))
        return self

    def synchronize(self):
        _check(self.L.aphip_synchronize(self.ctx))
        return self

    def set_graph(self, on=True):
        _check(self.L.aphip_set_graph(self.ctx, 1 if on else 0))
        return self

    def autotune(self, dt, trial_steps=5):
        chosen = ctypes.c_int32()
        _check(self.L.aphip_autotune(self.ctx, ctypes.c_double(dt), int(trial_steps), ctypes.byref(chosen)))
        return chosen.value

    # ---- adjoints ----
    def internal_force_vjp(self, f_bar):
        fb = np.ascontiguousarray(f_bar, dtype=np.float64)
        ub = np.zeros((self.nn, 3))
        Eb = np.zeros(self.ne)
        nb = np.zeros(self.ne)
        _check(self.L.aphip_internal_force_vjp(self.ctx, _p(fb), _p(ub), _p(Eb), _p(nb)))
        return ub, Eb, nb

    def adjoint_forward(self, dt, n_steps, checkpoint_interval):
        _check(self.L.aphip_adjoint_forward(self.ctx, ctypes.c_double(dt), int(n_steps), int(checkpoint_interval)))
        return self

    def adjoint_backward(self, u_bar_T, v_bar_T=None):
        ub = np.ascontiguousarray(u_bar_T, dtype=np.float64)
        vb = None if v_bar_T is None else np.ascontiguousarray(v_bar_T, dtype=np.float64)
        Eb = np.zeros(self.ne)
        nb = np.zeros(self.ne)
        v0b = np.zeros((self.nn, 3))
        _check(self.L.aphip_adjoint_backward(self.ctx, _p(ub), _p(vb), _p(Eb), _p(nb), _p(v0b)))
        return Eb, nb, v0b

    # ---- ensembles ----
    def batch_begin(self, etype, kind):
        _check(self.L.aphip_batch_begin(self.ctx, etype, kind))
        self.etype, self.kind, self._nn, self._ne = etype, kind, 0, 0
        return self

    def batch_add(self, X, conn, density, E, nu, dt, yield_stress=0.0, hardening=0.0):
        X = np.ascontiguousarray(X, dtype=np.float64)
        conn = np.ascontiguousarray(conn, dtype=np.int32)
        p = MaterialParams(density, E, nu, yield_stress, hardening, 0, 0)
        sp = ctypes.c_int32()
        _check(self.L.aphip_batch_add(self.ctx, X.shape[0], _p(X), conn.shape[0], _p(conn, _i32), ctypes.byref(p),
                                      ctypes.c_double(dt), ctypes.byref(sp)))
        self._nn += X.shape[0]
        self._ne += conn.shape[0]
        return sp.value

    def batch_end(self):
        _check(self.L.aphip_batch_end(self.ctx))
        self.nn, self.ne = self._nn, self._ne
        return self

    def batch_ranges(self, specimen):
        a, b, c, d = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
        _check(self.L.aphip_batch_ranges(self.ctx, specimen, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c),
                                         ctypes.byref(d)))
        return a.value, b.value, c.value, d.value

    # ---- RKPM ----
    def strain_smoothing_setup(self):
        """Finite-element strain smoothing, one cell per Tet4 (ElementSmoothedTetrahedron4)."""
        _check(self.L.aphip_strain_smoothing_setup(self.ctx))
        self._nodal = False
        return self

    def rk_setup(self, kernel_radius_scale, one_pass=True):
        """RK functions and smoothed cells; one_pass=False selects the two-pass force
        (physical displacement gathered before, force scattered after the cells)."""
        _check(self.L.aphip_rk_setup_method(self.ctx, ctypes.c_double(kernel_radius_scale), 1 if one_pass else 0))
        self._nodal = False
        return self

    def rk_nodal_setup(self, kernel_radius_scale, subdomains=0, f_bar=False):
        """Nodal smoothing cells on a thex mesh (0: every hex a subcell, 1: one per cell)."""
        _check(self.L.aphip_rk_nodal_setup(self.ctx, ctypes.c_double(kernel_radius_scale), int(subdomains),
                                           1 if f_bar else 0))
        self._nodal = True
        return self

    def rk_nodal_cells(self):
        """(cell_subcells, element_subcell, active, f_bar) of the nodal cells."""
        nc, ns, fb = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int32()
        _check(self.L.aphip_rk_get_nodal_cells(self.ctx, ctypes.byref(nc), ctypes.byref(ns), None, None, None,
                                               ctypes.byref(fb)))
        cs = np.zeros(nc.value + 1, dtype=np.int64)
        es = np.zeros(self.ne, dtype=np.int32)
        act = np.zeros(self.nn, dtype=np.uint8)
        _check(self.L.aphip_rk_get_nodal_cells(self.ctx, None, None, _p(cs, _i64), _p(es, _i32), _p(act, _u8), None))
        return cs, es, act, bool(fb.value)

    def rk_neighbors(self):
        num = np.zeros(self.nn, dtype=np.int32)
        nbr = np.zeros((self.nn, MAX_NODE_NUM_NEIGHBORS), dtype=np.int32)
        tr = ctypes.c_int64()
        _check(self.L.aphip_rk_get_neighbors(self.ctx, _p(num, _i32), _p(nbr, _i32), ctypes.byref(tr)))
        return num, nbr, tr.value

    def rk_function_values(self):
        phi = np.zeros((self.nn, MAX_NODE_NUM_NEIGHBORS))
        _check(self.L.aphip_rk_get_function_values(self.ctx, _p(phi)))
        return phi

    def rk_nnz(self):
        """Entries of the smoothed-cell CSR (sum of the cell node-list lengths)."""
        nnz = ctypes.c_int64()
        _check(self.L.aphip_rk_get_cells(self.ctx, ctypes.byref(nnz), None, None, None, None))
        return nnz.value

    def rk_cells(self):
        """Cell CSR (start, nodes, dphi, volume): per element in caller order, or, after
        rk_nodal_setup, per subcell in subcell order."""
        nnz = ctypes.c_int64()
        _check(self.L.aphip_rk_get_cells(self.ctx, ctypes.byref(nnz), None, None, None, None))
        ncells = self.ne
        if getattr(self, "_nodal", False):
            ns = ctypes.c_int64()
            _check(self.L.aphip_rk_get_nodal_cells(self.ctx, None, ctypes.byref(ns), None, None, None, None))
            ncells = ns.value
        start = np.zeros(ncells + 1, dtype=np.int64)
        nodes = np.zeros(nnz.value, dtype=np.int32)
        dphi = np.zeros((nnz.value, 3))
        vol = np.zeros(ncells)
        _check(self.L.aphip_rk_get_cells(self.ctx, ctypes.byref(nnz), _p(start, _i64), _p(nodes, _i32), _p(dphi),
                                         _p(vol)))
        return start, nodes, dphi, vol

    # ---- profiling ----
    def set_profiling(self, on=True):
        _check(self.L.aphip_set_profiling(self.ctx, 1 if on else 0))
        return self

    def profile(self, which):
        n = ctypes.c_int64()
        ms = ctypes.c_double()
        _check(self.L.aphip_get_profile(self.ctx, which, ctypes.byref(n), ctypes.byref(ms)))
        return n.value, ms.value

    def reset_profile(self):
        _check(self.L.aphip_reset_profile(self.ctx))
        return self
