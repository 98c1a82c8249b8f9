"""Device-resident field transfer (aphip_set_field_device / aphip_get_field_device) with
hipMalloc'd caller arrays: every field equals the host-array path bitwise, a force
evaluation fed and read through device arrays equals the host-fed one bitwise, and
arrays that are not device memory of the context's device, or too short, are refused."""
import ctypes

import numpy as np
import pytest

import aperi_hip as A
from aperi_hip import meshgen as M

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


class DeviceArray:
    """A hipMalloc'd buffer of float64 values, through the HIP runtime the library uses."""

    _hip = None

    def __init__(self, n):
        if DeviceArray._hip is None:
            A.load_library()  # loads libamdhip64 first, so the name below resolves to that runtime
            DeviceArray._hip = ctypes.CDLL("libamdhip64.so.7")
        self.n = int(n)
        self.ptr = ctypes.c_void_p()
        assert self._hip.hipMalloc(ctypes.byref(self.ptr), ctypes.c_size_t(8 * max(self.n, 1))) == 0

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64).reshape(-1)
        assert a.size == self.n
        assert self._hip.hipMemcpy(self.ptr, a.ctypes.data_as(ctypes.c_void_p), ctypes.c_size_t(8 * a.size), H2D) == 0
        return self

    def download(self):
        out = np.empty(self.n)
        assert self._hip.hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), self.ptr, ctypes.c_size_t(8 * self.n), D2H) == 0
        return out

    def free(self):
        if self.ptr:
            self._hip.hipFree(self.ptr)
            self.ptr = ctypes.c_void_p()

    @property
    def address(self):
        return self.ptr.value


def _mesh(n=5, seed=21):
    X, hc = M.hex_lattice(n, n, n)
    X = M.jitter(X, 1.0 / n, 0.1, seed, 1)
    return X, M.split_tets(X, hc)


def _solver(X, conn):
    s = A.Solver()
    s.set_mesh(A.TET4, X, conn)
    s.set_material(A.J2, 2700.0, 70e9, 0.33, 300e6, 1e9)
    return s


def _get_device(s, field):
    d = DeviceArray(int(np.prod(s.field_shape(field))))
    try:
        s.get_field_device(field, d.address)
        s.synchronize()
        return d.download().reshape(s.field_shape(field))
    finally:
        d.free()


def test_device_fields_equal_host_fields(gpu_available):
    X, conn = _mesh()
    rng = np.random.default_rng(3)
    u = 5e-3 * rng.uniform(-1, 1, X.shape)
    v = rng.normal(0, 1, X.shape)
    mass = 1.0 + rng.uniform(0, 1, len(X))
    state = 1e-3 * rng.uniform(0, 1, len(conn))
    with _solver(X, conn) as h, _solver(X, conn) as d:
        # host-fed reference context
        h.set_field(A.FIELD_U, u)
        h.set_field(A.FIELD_V, v)
        h.set_field(A.FIELD_MASS, mass)
        h.set_field(A.FIELD_STATE, state)
        # device-fed context
        bufs = {f: DeviceArray(a.size).upload(a) for f, a in
                ((A.FIELD_U, u), (A.FIELD_V, v), (A.FIELD_MASS, mass), (A.FIELD_STATE, state))}
        try:
            for f, b in bufs.items():
                d.set_field_device(f, b.address)
            d.synchronize()
        finally:
            for b in bufs.values():
                b.free()
        for s in (h, d):
            s.internal_force(0.0, 1e-7)
        for f in (A.FIELD_U, A.FIELD_V, A.FIELD_F, A.FIELD_MASS, A.FIELD_H, A.FIELD_P, A.FIELD_STATE,
                  A.FIELD_VOLUME):
            host = h.get(f)
            assert np.array_equal(d.get(f), host), f
            assert np.array_equal(_get_device(d, f), host), f
            assert np.array_equal(_get_device(h, f), host), f
        # a few explicit steps after device-side set-up, read back on the device
        for s in (h, d):
            s.initialize()
            s.step(2e-8, 6)
        for f in (A.FIELD_U, A.FIELD_V, A.FIELD_A, A.FIELD_F, A.FIELD_H, A.FIELD_P, A.FIELD_STATE):
            assert np.array_equal(_get_device(d, f), h.get(f)), f


def test_force_call_through_device_arrays(gpu_available):
    """The binding of ComputeInternalForceAllElements with device-resident fields: u in and
    f out through caller device arrays, no host array on the way; equal to the host path."""
    X, conn = _mesh(6, 4)
    rng = np.random.default_rng(9)
    u_dev, f_dev = DeviceArray(X.size), DeviceArray(X.size)
    try:
        with _solver(X, conn) as h, _solver(X, conn) as d:
            for k in range(3):
                u = 4e-3 * rng.uniform(-1, 1, X.shape)
                u_dev.upload(u)
                d.set_field_device(A.FIELD_U, u_dev.address)
                d.internal_force(0.0, 1e-7)
                d.get_field_device(A.FIELD_F, f_dev.address)
                d.synchronize()
                h.set_field(A.FIELD_U, u)
                h.internal_force(0.0, 1e-7)
                assert np.array_equal(f_dev.download().reshape(X.shape), h.get(A.FIELD_F)), k
    finally:
        u_dev.free()
        f_dev.free()


def test_device_array_checks(gpu_available):
    X, conn = _mesh(20, 2)
    with _solver(X, conn) as s:
        host = np.zeros(X.shape)
        with pytest.raises(A.AphipError, match="device memory"):
            s.set_field_device(A.FIELD_U, host.ctypes.data)
        with pytest.raises(A.AphipError, match="device memory"):
            s.get_field_device(A.FIELD_F, host.ctypes.data)
        # 1 MiB against the 3.5 MB of H (48,000 tets x 9 doubles): short whatever the
        # runtime's allocation granularity
        short = DeviceArray(1 << 17)
        try:
            with pytest.raises(A.AphipError, match="shorter"):
                s.get_field_device(A.FIELD_H, short.address)
            with pytest.raises(A.AphipError, match="shorter"):
                s.get_field_device(A.FIELD_P, short.address)
        finally:
            short.free()
        bad = DeviceArray(len(X)).upload(np.r_[np.ones(len(X) - 1), 0.0])
        try:
            with pytest.raises(A.AphipError, match="positive"):
                s.set_field_device(A.FIELD_MASS, bad.address)
            with pytest.raises(A.AphipError, match="mass not computed"):
                s.get(A.FIELD_MASS)
        finally:
            bad.free()
