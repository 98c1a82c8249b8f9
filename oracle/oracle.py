"""ctypes front end of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py.  The product (aperi-mech_amd/) never imports it.
Every function here is a thin marshalling layer over oracle/*.c, which cites
the reference file:line it restates.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

TET4, HEX8 = 0, 1
LINEAR_ELASTIC, NEO_HOOKEAN, J2, SVK, NH_DAMAGE, CREEP, DP = 0, 1, 2, 3, 4, 5, 6
MAX_NODE_NUM_NEIGHBORS = 40  # include/Constants.h:15

_d = ctypes.POINTER(ctypes.c_double)
_i32 = ctypes.POINTER(ctypes.c_int32)
_i64 = ctypes.POINTER(ctypes.c_int64)
_u8 = ctypes.POINTER(ctypes.c_uint8)


class Material(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int32), ("state_per_qp", ctypes.c_int32),
                ("lam", ctypes.c_double), ("mu", ctypes.c_double),
                ("yield_stress", ctypes.c_double), ("hardening_modulus", ctypes.c_double),
                ("elem_lambda", _d), ("elem_mu", _d),
                ("I1_critical", ctypes.c_double), ("I1_failure", ctypes.c_double), ("alpha", ctypes.c_double),
                ("maximum_damage", ctypes.c_double), ("maximum_damage_increment", ctypes.c_double),
                ("bulk", ctypes.c_double), ("creep_A", ctypes.c_double), ("creep_n", ctypes.c_double),
                ("creep_m", ctypes.c_double), ("temperature", ctypes.c_double),
                ("dp_A1", ctypes.c_double), ("dp_A2", ctypes.c_double), ("dp_A2G", ctypes.c_double),
                ("H_n", _d), ("P_n", _d), ("dt", ctypes.c_double),
                ("updated", ctypes.c_int32), ("pad_", ctypes.c_int32)]


class Problem(ctypes.Structure):
    _fields_ = [("etype", ctypes.c_int32), ("n_nodes", ctypes.c_int64), ("X", _d),
                ("n_elem", ctypes.c_int64), ("conn", _i32), ("mat", Material),
                ("mass", _d), ("fixed", _u8), ("node_dt", _d),
                ("cell_start", _i64), ("cell_nodes", _i32), ("cell_dphi", _d), ("cell_volume", _d),
                ("rk_two_pass", ctypes.c_int32), ("max_neighbors", ctypes.c_int32), ("num_neighbors", _i32),
                ("neighbors", _i32), ("phi", _d), ("x_cur", _d), ("d_inc", _d),
                ("semi_interval", ctypes.c_int64), ("semi_steps", _i64),
                ("fbar", ctypes.c_int32), ("pad2_", ctypes.c_int32), ("n_fbar_cells", ctypes.c_int64),
                ("fbar_cell_start", _i64), ("fbar_cell_volume", _d), ("active", _u8)]


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = ctypes.CDLL(path)
        L.or_det_ap_im1.restype = ctypes.c_double
        L.or_kernel_value.restype = ctypes.c_double
        L.or_power_method_dt.restype = ctypes.c_double
        L.or_rk_cells_count.restype = ctypes.c_int64
        L.or_nodal_active.restype = ctypes.c_int64
        L.or_nodal_cells.restype = ctypes.c_int64
        L.or_rk_subcells_count.restype = ctypes.c_int64
        _LIB = L
    return _LIB


def _p(a, t=_d):
    return None if a is None else a.ctypes.data_as(t)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _i32a(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def lame(E, nu):
    """(lambda, mu) from Young's modulus and Poisson's ratio (Materials/Base.h:97-107)."""
    lam, mu, K = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
    if lib().or_lame_from_E_nu(ctypes.c_double(E), ctypes.c_double(nu), ctypes.byref(lam), ctypes.byref(mu),
                               ctypes.byref(K)):
        raise ValueError("invalid E/nu")
    return lam.value, mu.value


def elastic_properties(code_a, a, code_b, b):
    out = np.zeros(5)
    rc = lib().or_elastic_properties(code_a, ctypes.c_double(a), code_b, ctypes.c_double(b), _p(out))
    if rc:
        raise ValueError("invalid property pair or bounds")
    return dict(zip(["E", "nu", "G", "lambda", "K"], out))


def det_ap_im1(A):
    return lib().or_det_ap_im1(_p(_f64(A).reshape(9)))


def kernel_value(d, R):
    return lib().or_kernel_value(_p(_f64(d)), ctypes.c_double(R))


def inverse3_cofactor(A):
    out = np.zeros(9)
    det = ctypes.c_double()
    lib().or_inverse3_cofactor(_p(_f64(A).reshape(9)), _p(out), ctypes.byref(det))
    return out.reshape(3, 3), det.value


def fullpivlu_inverse(A):
    A = _f64(A)
    n = A.shape[0]
    out = np.zeros((n, n))
    rank = lib().or_fullpivlu_inverse(n, _p(A), _p(out))
    return out, rank


def stress(kind, lam, mu, H, state_n=0.0, yield_stress=0.0, hardening=0.0):
    H = _f64(H).reshape(9)
    P = np.zeros(9)
    sn = np.array([state_n], dtype=np.float64)
    snp1 = np.zeros(1)
    lib().or_stress(kind, ctypes.c_double(lam), ctypes.c_double(mu), ctypes.c_double(yield_stress),
                    ctypes.c_double(hardening), _p(H), _p(sn), _p(snp1), _p(P))
    return P.reshape(3, 3), snp1[0]


def tet4_bmatrix(X):
    B = np.zeros(12)
    w = ctypes.c_double()
    lib().or_tet4_bmatrix(_p(_f64(X).reshape(12)), _p(B), ctypes.byref(w))
    return B.reshape(4, 3), w.value


def hex8_bmatrix(X, gauss_id):
    B = np.zeros(24)
    w = ctypes.c_double()
    lib().or_hex8_bmatrix(_p(_f64(X).reshape(24)), gauss_id, _p(B), ctypes.byref(w))
    return B.reshape(8, 3), w.value


def smoothed_tet4_bmatrix(X):
    B = np.zeros(12)
    V = ctypes.c_double()
    lib().or_smoothed_tet4_bmatrix(_p(_f64(X).reshape(12)), _p(B), ctypes.byref(V))
    return B.reshape(4, 3), V.value


def smoothed_hex8_bmatrix(X):
    B = np.zeros(24)
    V = ctypes.c_double()
    lib().or_smoothed_hex8_bmatrix(_p(_f64(X).reshape(24)), _p(B), ctypes.byref(V))
    return B.reshape(8, 3), V.value


DAMAGE_DEFAULTS = dict(maximum_damage=0.99, maximum_damage_increment=1.0)  # Materials/Factory.h:78-81


def _material(kind, lam, mu, yield_stress=0.0, hardening=0.0, elem_lambda=None, elem_mu=None, state_per_qp=0,
              damage=None, creep=None, rate=None):
    """damage: dict of I1_critical, I1_failure, alpha[, maximum_damage, maximum_damage_increment]
    for NH_DAMAGE; creep: dict of A, n, m, constant_temperature and bulk (the bulk modulus)
    for CREEP; rate: (H_n, P_n, dt) with H_n, P_n float64 [n_elem, 9] arrays (kept alive by
    the caller)."""
    m = Material()
    if creep is not None and "A1" in creep:  # Drucker-Prager parameters (A1, A2, A2G, bulk)
        m.bulk, m.dp_A1, m.dp_A2, m.dp_A2G = creep["bulk"], creep["A1"], creep["A2"], creep["A2G"]
    elif creep is not None:
        m.bulk, m.creep_A, m.creep_n, m.creep_m = creep["bulk"], creep["A"], creep["n"], creep["m"]
        m.temperature = creep.get("constant_temperature", 0.0)
    if rate is not None:
        m.H_n, m.P_n, m.dt = _p(rate[0]), _p(rate[1]), rate[2]
    m.kind, m.state_per_qp = kind, state_per_qp
    m.lam, m.mu, m.yield_stress, m.hardening_modulus = lam, mu, yield_stress, hardening
    m.elem_lambda = _p(elem_lambda)
    m.elem_mu = _p(elem_mu)
    if damage is not None:
        d = dict(DAMAGE_DEFAULTS, **damage)
        for k in ("I1_critical", "I1_failure", "alpha", "maximum_damage", "maximum_damage_increment"):
            setattr(m, k, d[k])
    return m


def state_width(kind):
    return {J2: 1, NH_DAMAGE: 2, CREEP: 2, DP: 2}.get(kind, 0)


def state_size(etype, kind, n_elem, state_per_qp=0):
    w = state_width(kind)
    if w == 0:
        return 1
    return w * n_elem * (8 if (etype == HEX8 and state_per_qp) else 1)


def stress_creep(lam, mu, H, Hn, Pn, dt, creep, state_n=(0.0, 0.0), kind=None):
    """power-law creep: (P, (effective creep strain, effective stress rate)); with
    kind=DP and creep = dict(A1, A2, A2G, bulk), Drucker-Prager: (P, (plastic strain,
    volumetric plastic strain))."""
    mat = _material(CREEP if kind is None else kind, lam, mu, creep=creep)
    P = np.zeros(9)
    snp1 = np.zeros(2)
    sn = _f64(np.asarray(state_n, dtype=np.float64))
    lib().or_stress_rate(ctypes.byref(mat), ctypes.c_double(lam), ctypes.c_double(mu), _p(_f64(H).reshape(9)),
                         _p(_f64(Hn).reshape(9)), _p(_f64(Pn).reshape(9)), ctypes.c_double(dt), _p(sn), _p(snp1), _p(P))
    return P.reshape(3, 3), snp1


def stress_damage(lam, mu, H, damage, state_n=(0.0, 0.0)):
    """neo-Hookean with damage: (P, (damage, I1)) from H and the previous damage."""
    mat = _material(NH_DAMAGE, lam, mu, damage=damage)
    H = _f64(H).reshape(9)
    P = np.zeros(9)
    sn = _f64(np.asarray(state_n, dtype=np.float64))
    snp1 = np.zeros(2)
    lib().or_stress_mat(ctypes.byref(mat), ctypes.c_double(lam), ctypes.c_double(mu), _p(H), _p(sn), _p(snp1), _p(P))
    return P.reshape(3, 3), snp1


def internal_force(etype, X, conn, kind, lam, mu, u, state_n=None, yield_stress=0.0, hardening=0.0,
                   elem_lambda=None, elem_mu=None, state_per_qp=0, nthreads=1, damage=None, creep=None, rate=None):
    """f = -sum_e int B^T P dV, with H, P (last Gauss point) and the new state."""
    X, conn, u = _f64(X), _i32a(conn), _f64(u)
    ne, nn = conn.shape[0], X.shape[0]
    el = None if elem_lambda is None else _f64(elem_lambda)
    em = None if elem_mu is None else _f64(elem_mu)
    if kind in (CREEP, DP) and rate is None:
        rate = (np.zeros((ne, 9)), np.zeros((ne, 9)), 0.0)
    if rate is not None:
        rate = (_f64(rate[0]), _f64(rate[1]), rate[2])
    mat = _material(kind, lam, mu, yield_stress, hardening, el, em, state_per_qp, damage, creep, rate)
    ns = state_size(etype, kind, ne, state_per_qp)
    sn = np.zeros(ns) if state_n is None else _f64(state_n).copy()
    snp1 = np.zeros(ns)
    H = np.zeros((ne, 9))
    P = np.zeros((ne, 9))
    f = np.zeros((nn, 3))
    rc = lib().or_internal_force(etype, nn, _p(X), ne, _p(conn, _i32), ctypes.byref(mat), _p(u), _p(sn), _p(snp1),
                                 _p(H), _p(P), _p(f), nthreads)
    if rc:
        raise RuntimeError("or_internal_force failed")
    return f, H, P, snp1


def element_volumes(etype, X, conn):
    X, conn = _f64(X), _i32a(conn)
    vol = np.zeros(conn.shape[0])
    lib().or_element_volumes(etype, conn.shape[0], _p(conn, _i32), _p(X), _p(vol))
    return vol


def lumped_mass(etype, X, conn, density, elem_density=None):
    X, conn = _f64(X), _i32a(conn)
    m = np.zeros(X.shape[0])
    ed = None if elem_density is None else _f64(elem_density)
    lib().or_lumped_mass(etype, X.shape[0], _p(X), conn.shape[0], _p(conn, _i32), ctypes.c_double(density),
                         _p(ed), _p(m))
    return m


class Solver:
    """The explicit loop of src/Solvers/Explicit.cpp:308-452 on the CPU oracle."""

    def __init__(self, etype, X, conn, kind, lam, mu, mass, fixed=None, yield_stress=0.0, hardening=0.0,
                 elem_lambda=None, elem_mu=None, node_dt=None, state_per_qp=0, cells=None, two_pass=None,
                 damage=None, creep=None, updated=False, semi_interval=0, fbar_cells=None, active=None):
        """updated: the updated Lagrangian formulation (Gaussian elements); semi_interval > 0:
        the semi-Lagrangian one with that reference update interval."""
        """cells: RK smoothed cells (start, nodes, dphi, volume); two_pass: (num, nbr, phi)
        of the RK functions, with `cells` from rk_cells_two_pass, for the two-pass force.
        fbar_cells: (cell_start, cell_volume) grouping `cells` (the subcells) into nodal
        cells for the F-bar force; active: nodes with degrees of freedom (others are not
        integrated)."""
        self.X = _f64(X)
        self.conn = None if conn is None else _i32a(conn)
        self.mass = _f64(mass)
        self.fixed = None if fixed is None else np.ascontiguousarray(fixed, dtype=np.uint8)
        if self.fixed is not None and self.fixed.size != self.X.size:
            raise ValueError("fixed must hold n_nodes x 3 flags")
        self.el = None if elem_lambda is None else _f64(elem_lambda)
        self.em = None if elem_mu is None else _f64(elem_mu)
        self.node_dt = None if node_dt is None else _f64(node_dt)
        self.nn = self.X.shape[0]
        p = Problem()
        p.etype = etype
        p.n_nodes = self.nn
        p.X = _p(self.X)
        if cells is not None:
            self.cs, self.cn, self.cd, self.cv = (np.ascontiguousarray(cells[0], dtype=np.int64), _i32a(cells[1]),
                                                  _f64(cells[2]), _f64(cells[3]))
            p.etype = -1
            p.n_elem = self.cv.shape[0]
            p.cell_start, p.cell_nodes = _p(self.cs, _i64), _p(self.cn, _i32)
            p.cell_dphi, p.cell_volume = _p(self.cd), _p(self.cv)
            p.conn = None
            if two_pass is not None:
                self.tp = (_i32a(two_pass[0]), _i32a(two_pass[1]), _f64(two_pass[2]))
                p.rk_two_pass = 1
                p.max_neighbors = self.tp[1].shape[1]
                p.num_neighbors, p.neighbors, p.phi = _p(self.tp[0], _i32), _p(self.tp[1], _i32), _p(self.tp[2])
        else:
            p.n_elem = self.conn.shape[0]
            p.conn = _p(self.conn, _i32)
        ne_ = self.cv.shape[0] if cells is not None else self.conn.shape[0]
        self.Hn, self.Pn = np.zeros((ne_, 9)), np.zeros((ne_, 9))  # N halves of the H, P pairs
        p.mat = _material(kind, lam, mu, yield_stress, hardening, self.el, self.em, state_per_qp, damage, creep,
                          (self.Hn, self.Pn, 0.0))
        if updated or semi_interval:
            self.xc, self.dinc = self.X.copy(), np.zeros_like(self.X)
            p.mat.updated = 2 if semi_interval else 1
            p.x_cur, p.d_inc = _p(self.xc), _p(self.dinc)
            self.semi_steps = np.zeros(1, dtype=np.int64)
            p.semi_interval, p.semi_steps = int(semi_interval), _p(self.semi_steps, _i64)
        if fbar_cells is not None:
            self.fcs = np.ascontiguousarray(fbar_cells[0], dtype=np.int64)
            self.fcv = _f64(fbar_cells[1])
            p.fbar, p.n_fbar_cells = 1, self.fcs.shape[0] - 1
            p.fbar_cell_start, p.fbar_cell_volume = _p(self.fcs, _i64), _p(self.fcv)
        self.active = None if active is None else np.ascontiguousarray(active, dtype=np.uint8)
        p.active = _p(self.active, _u8)
        p.mass = _p(self.mass)
        p.fixed = _p(self.fixed, _u8)
        p.node_dt = _p(self.node_dt)
        self.p = p
        self.ne = p.n_elem
        ns = state_size(etype, kind, self.ne, state_per_qp)
        self.sa, self.sb = np.zeros(ns), np.zeros(ns)
        self.u = np.zeros((self.nn, 3))
        self.v = np.zeros((self.nn, 3))
        self.a = np.zeros((self.nn, 3))
        self.f = np.zeros((self.nn, 3))
        self.H = np.zeros((self.ne, 9))
        self.P = np.zeros((self.ne, 9))
        self.state = self.sb
        self.initialized = False

    def run(self, dt, n_steps, nthreads=1):
        flag = ctypes.c_int()
        rc = lib().or_solve(ctypes.byref(self.p), ctypes.c_double(dt), n_steps, 0 if self.initialized else 1,
                            _p(self.u), _p(self.v), _p(self.a), _p(self.f), _p(self.H), _p(self.P),
                            _p(self.sa), _p(self.sb), ctypes.byref(flag), nthreads)
        if rc:
            raise RuntimeError("or_solve failed")
        self.initialized = True
        if flag.value == 0:  # keep "sb" as the NP1 buffer across calls
            self.sa, self.sb = self.sb, self.sa
        self.state = self.sb
        return self

    def power_method_dt(self, n_iter=50, epsilon=1e-5, seed=20261202, nthreads=1):
        return lib().or_power_method_dt(ctypes.byref(self.p), n_iter, ctypes.c_double(epsilon),
                                        ctypes.c_uint64(seed), nthreads)

    def power_method(self, max_edge, u_in=None, n_iter=50, epsilon=1e-5, tolerance=1e-3,
                     variation_tolerance=1e-7, seed=20261202, nthreads=1, previous_dt=0.0, eigenvector=None):
        """The reference's power-method increment about u_in (default: the solver's u):
        returns (dt, iterations, converged).  `eigenvector`, a dict, carries the vector
        between calls like the reference's processor does (empty dict: random start)."""
        u = _f64(self.u if u_in is None else u_in)
        l = _f64(max_edge)
        dt, it, cv = ctypes.c_double(), ctypes.c_int(), ctypes.c_int()
        st = None if self.state.size == 0 else _f64(self.state)
        eig, valid = None, 0
        if eigenvector is not None:
            valid = 1 if "v" in eigenvector else 0
            eig = eigenvector.setdefault("v", np.zeros(u.size))
        rc = lib().or_power_method(ctypes.byref(self.p), _p(u), _p(st), _p(l), n_iter, ctypes.c_double(epsilon),
                                   ctypes.c_double(tolerance), ctypes.c_double(variation_tolerance),
                                   ctypes.c_uint64(seed), ctypes.c_double(previous_dt), _p(eig), valid, nthreads,
                                   ctypes.byref(dt), ctypes.byref(it), ctypes.byref(cv))
        if rc != 0:
            raise RuntimeError("or_power_method failed")
        return dt.value, it.value, bool(cv.value)


# ---------------- RKPM ----------------

def max_edge_length(X, conn, active=None):
    X, conn = _f64(X), _i32a(conn)
    out = np.zeros(X.shape[0])
    act = None if active is None else np.ascontiguousarray(active, dtype=np.uint8)
    lib().or_max_edge_length(X.shape[0], _p(X), conn.shape[0], _p(conn, _i32), conn.shape[1], _p(act, _u8), _p(out))
    return out


def kernel_radius(max_edge, scale):
    """src/NeighborSearchProcessor.cpp:154-185: r = scale * l, scale 1.0 -> 1.0 + 1e-6."""
    if scale < 0:
        raise ValueError("scale must be non-negative")
    if scale == 1.0:
        scale += 1.0e-6
    return np.asarray(max_edge) * scale


def neighbor_search(X, radius, max_neighbors=MAX_NODE_NUM_NEIGHBORS, brute_force=False, range_mask=None):
    """range_mask: only nodes with a nonzero entry are candidates (the active nodes)."""
    X, radius = _f64(X), _f64(radius)
    n = X.shape[0]
    num = np.zeros(n, dtype=np.int32)
    nbr = np.zeros((n, max_neighbors), dtype=np.int32)
    trunc = ctypes.c_int64()
    mask = None if range_mask is None else np.ascontiguousarray(range_mask, dtype=np.uint8)
    rc = lib().or_neighbor_search_masked(n, _p(X), _p(radius), max_neighbors, _p(mask, _u8),
                                         1 if brute_force else 0, _p(num, _i32), _p(nbr, _i32), ctypes.byref(trunc))
    if rc:
        raise RuntimeError("or_neighbor_search failed")
    return num, nbr, trunc.value


def rk_function_values(X, radius, num, nbr):
    X, radius = _f64(X), _f64(radius)
    n, mx = nbr.shape
    phi = np.zeros((n, mx))
    lib().or_rk_function_values(n, _p(X), _p(radius), mx, _p(_i32a(num), _i32), _p(_i32a(nbr), _i32), _p(phi))
    return phi


def rk_cells(X, conn, num, nbr, phi):
    X, conn, num, nbr, phi = _f64(X), _i32a(conn), _i32a(num), _i32a(nbr), _f64(phi)
    ne, npe, mx = conn.shape[0], conn.shape[1], nbr.shape[1]
    start = np.zeros(ne + 1, dtype=np.int64)
    nnz = lib().or_rk_cells_count(ne, _p(conn, _i32), npe, mx, _p(num, _i32), _p(nbr, _i32), _p(start, _i64))
    if nnz < 0:
        raise RuntimeError("or_rk_cells_count failed")
    nodes = np.zeros(nnz, dtype=np.int32)
    dphi = np.zeros((nnz, 3))
    vol = np.zeros(ne)
    rc = lib().or_rk_cells_build(X.shape[0], _p(X), ne, _p(conn, _i32), npe, mx, _p(num, _i32), _p(nbr, _i32), _p(phi),
                                 _p(start, _i64), _p(nodes, _i32), _p(dphi), _p(vol))
    if rc:
        raise RuntimeError("or_rk_cells_build failed")
    return start, nodes, dphi, vol


def cell_internal_force(n_nodes, cells, kind, lam, mu, u, state_n=None, yield_stress=0.0, hardening=0.0,
                        nthreads=1, damage=None, creep=None, rate=None):
    start, nodes, dphi, vol = cells
    nc = vol.shape[0]
    if kind in (CREEP, DP) and rate is None:
        rate = (np.zeros((nc, 9)), np.zeros((nc, 9)), 0.0)
    if rate is not None:
        rate = (_f64(rate[0]), _f64(rate[1]), rate[2])
    mat = _material(kind, lam, mu, yield_stress, hardening, damage=damage, creep=creep, rate=rate)
    w = max(state_width(kind), 1)
    sn = np.zeros(nc * w) if state_n is None else _f64(state_n).copy()
    snp1 = np.zeros(nc * w)
    H, P, f = np.zeros((nc, 9)), np.zeros((nc, 9)), np.zeros((n_nodes, 3))
    lib().or_cell_internal_force(n_nodes, nc, _p(np.ascontiguousarray(start, dtype=np.int64), _i64),
                                 _p(_i32a(nodes), _i32), _p(_f64(dphi)), _p(_f64(vol)), ctypes.byref(mat),
                                 _p(_f64(u)), _p(sn), _p(snp1), _p(H), _p(P), _p(f), nthreads)
    return f, H, P, snp1


def rk_cells_two_pass(X, conn):
    """Cells of the two-pass method: each element's own nodes and (B_k V)/V."""
    X, conn = _f64(X), _i32a(conn)
    ne, npe = conn.shape
    start = np.zeros(ne + 1, dtype=np.int64)
    nodes = np.zeros(npe * ne, dtype=np.int32)
    dphi = np.zeros((npe * ne, 3))
    vol = np.zeros(ne)
    rc = lib().or_rk_cells_two_pass(ne, _p(X), _p(conn, _i32), npe, _p(start, _i64), _p(nodes, _i32), _p(dphi),
                                    _p(vol))
    if rc:
        raise RuntimeError("or_rk_cells_two_pass failed")
    return start, nodes, dphi, vol


def rk_gather(num, nbr, phi, src):
    """Physical values from coefficients: sum_k phi_ik src_(nbr_ik)."""
    num, nbr, phi, src = _i32a(num), _i32a(nbr), _f64(phi), _f64(src)
    out = np.zeros_like(src)
    lib().or_rk_gather(nbr.shape[0], nbr.shape[1], _p(num, _i32), _p(nbr, _i32), _p(phi), _p(src), _p(out))
    return out


def rk_scatter(num, nbr, phi, src):
    """Coefficient values from physical ones: the transpose of rk_gather."""
    num, nbr, phi, src = _i32a(num), _i32a(nbr), _f64(phi), _f64(src)
    out = np.zeros_like(src)
    lib().or_rk_scatter(nbr.shape[0], nbr.shape[1], _p(num, _i32), _p(nbr, _i32), _p(phi), _p(src), _p(out))
    return out


def rk_two_pass_force(n_nodes, cells2, num, nbr, phi, kind, lam, mu, d, state_n=None, yield_stress=0.0,
                      hardening=0.0, damage=None):
    """Two-pass internal force of coefficient displacements d: (f, H, P, state, u_phys, f_local)."""
    up = rk_gather(num, nbr, phi, d)
    fl, H, P, s = cell_internal_force(n_nodes, cells2, kind, lam, mu, up, state_n, yield_stress, hardening,
                                      damage=damage)
    return rk_scatter(num, nbr, phi, fl), H, P, s, up, fl


def rk_mass(conn, elem_volume, density, num, nbr, phi):
    conn, num, nbr, phi = _i32a(conn), _i32a(num), _i32a(nbr), _f64(phi)
    n, mx = nbr.shape
    m = np.zeros(n)
    lib().or_rk_mass(n, conn.shape[0], _p(conn, _i32), conn.shape[1], _p(_f64(elem_volume)), ctypes.c_double(density), mx,
                     _p(num, _i32), _p(nbr, _i32), _p(phi), _p(m))
    return m


# ---------------- nodal strain smoothing (thex meshes), subcells, F-bar ----------------

def nodal_active(hexes, n_nodes, node_id=None):
    """(active mask, number of hexes without exactly one active node)."""
    hexes = _i32a(hexes)
    act = np.zeros(n_nodes, dtype=np.uint8)
    ids = None if node_id is None else np.ascontiguousarray(node_id, dtype=np.int64)
    bad = lib().or_nodal_active(n_nodes, hexes.shape[0], _p(hexes, _i32), _p(ids, _i64), _p(act, _u8))
    return act, int(bad)


def nodal_cells(hexes, active, subdomains):
    """Cells (one per active node, ascending) and subcells: returns (hex_cell, cell_start,
    cell_hexes, sub_start, sub_elems, cell_sub), where the hexes of a cell are listed in
    ascending order, subdomains 0 makes every hex a subcell and 1 every cell one subcell;
    cell_sub[c]..cell_sub[c+1] are the cell's subcells."""
    hexes = _i32a(hexes)
    ne = hexes.shape[0]
    hc = np.zeros(ne, dtype=np.int32)
    nc = lib().or_nodal_cells(len(active), ne, _p(hexes, _i32), _p(np.ascontiguousarray(active, dtype=np.uint8), _u8),
                              _p(hc, _i32))
    order = np.lexsort((np.arange(ne), hc)).astype(np.int32)
    cell_start = np.zeros(nc + 1, dtype=np.int64)
    np.add.at(cell_start, hc + 1, 1)
    cell_start = np.cumsum(cell_start)
    if subdomains == 0:
        sub_start = np.arange(ne + 1, dtype=np.int64)
        cell_sub = cell_start.copy()
    else:
        sub_start = cell_start.copy()
        cell_sub = np.arange(nc + 1, dtype=np.int64)
    return hc, cell_start, order, sub_start, order.copy(), cell_sub


def rk_subcells(X, conn, sub_start, sub_elems, num, nbr, phi):
    """Subcell CSR (start, nodes, dphi, sub_volume) and every element's smoothed volume."""
    X, conn, num, nbr, phi = _f64(X), _i32a(conn), _i32a(num), _i32a(nbr), _f64(phi)
    sub_start = np.ascontiguousarray(sub_start, dtype=np.int64)
    sub_elems = _i32a(sub_elems)
    ns, npe, mx = sub_start.shape[0] - 1, conn.shape[1], nbr.shape[1]
    start = np.zeros(ns + 1, dtype=np.int64)
    nnz = lib().or_rk_subcells_count(ns, _p(sub_start, _i64), _p(sub_elems, _i32), _p(conn, _i32), npe, mx,
                                     _p(num, _i32), _p(nbr, _i32), _p(start, _i64))
    if nnz < 0:
        raise RuntimeError("or_rk_subcells_count failed")
    nodes = np.zeros(nnz, dtype=np.int32)
    dphi = np.zeros((nnz, 3))
    vol = np.zeros(ns)
    evol = np.zeros(conn.shape[0])
    rc = lib().or_rk_subcells_build(X.shape[0], _p(X), ns, _p(sub_start, _i64), _p(sub_elems, _i32), _p(conn, _i32),
                                    npe, mx, _p(num, _i32), _p(nbr, _i32), _p(phi), _p(start, _i64), _p(nodes, _i32),
                                    _p(dphi), _p(vol), _p(evol))
    if rc:
        raise RuntimeError("or_rk_subcells_build failed")
    return (start, nodes, dphi, vol), evol


def fbar_internal_force(n_nodes, cells, fbar_cells, kind, lam, mu, u, state_n=None, yield_stress=0.0,
                        hardening=0.0, nthreads=1):
    """F-bar force of subcells `cells` grouped by fbar_cells = (cell_sub, cell_volume):
    returns (f, H, Hbar, P, Pbar, state_np1)."""
    start, nodes, dphi, vol = cells
    cs, cv = np.ascontiguousarray(fbar_cells[0], dtype=np.int64), _f64(fbar_cells[1])
    ns = vol.shape[0]
    mat = _material(kind, lam, mu, yield_stress, hardening)
    w = max(state_width(kind), 1)
    sn = np.zeros(ns * w) if state_n is None else _f64(state_n).copy()
    snp1 = np.zeros(ns * w)
    H, Hb, P, Pb, f = np.zeros((ns, 9)), np.zeros((ns, 9)), np.zeros((ns, 9)), np.zeros((ns, 9)), np.zeros((n_nodes, 3))
    rc = lib().or_fbar_internal_force(n_nodes, cs.shape[0] - 1, _p(cs, _i64), _p(cv),
                                      _p(np.ascontiguousarray(start, dtype=np.int64), _i64), _p(_i32a(nodes), _i32),
                                      _p(_f64(dphi)), _p(_f64(vol)), ctypes.byref(mat), _p(_f64(u)), _p(sn), _p(snp1),
                                      _p(H), _p(Hb), _p(P), _p(Pb), _p(f), nthreads)
    if rc:
        raise RuntimeError("or_fbar_internal_force failed")
    return f, H, Hb, P, Pb, snp1


def nodal_setup(X, hexes, scale, subdomains, density=None, node_id=None):
    """The whole nodal-smoothing set-up on a thex mesh: active nodes, max edge lengths
    (edges to inactive nodes doubled), radii of the active nodes, search over the active
    nodes, RK functions at every node, cells and subcells, subcell CSR, RK mass.  Returns a
    dict."""
    X, hexes = _f64(X), _i32a(hexes)
    act, bad = nodal_active(hexes, len(X), node_id)
    if bad:
        raise ValueError(f"{bad} hexes without exactly one active node")
    edge = max_edge_length(X, hexes, active=act)
    radius = np.where(act.astype(bool), kernel_radius(edge, scale), 0.0)
    num, nbr, trunc = neighbor_search(X, radius, range_mask=act)
    phi = rk_function_values(X, radius, num, nbr)
    hc, cell_start, cell_hexes, sub_start, sub_elems, cell_sub = nodal_cells(hexes, act, subdomains)
    cells, evol = rk_subcells(X, hexes, sub_start, sub_elems, num, nbr, phi)
    cell_volume = np.array([evol[cell_hexes[cell_start[c]:cell_start[c + 1]]].sum()
                            for c in range(len(cell_start) - 1)])
    out = dict(active=act, edge=edge, radius=radius, num=num, nbr=nbr, trunc=trunc, phi=phi, hex_cell=hc,
               cell_start=cell_start, cell_hexes=cell_hexes, sub_start=sub_start, sub_elems=sub_elems,
               cell_sub=cell_sub, cells=cells, elem_volume=evol, cell_volume=cell_volume)
    if density is not None:
        out["mass"] = rk_mass(hexes, evol, density, num, nbr, phi)
    return out
