"""Seeded synthetic meshes and initial conditions for the explicit force step.

The reference builds its test meshes with STK's generated-mesh strings such as
"generated:NxMxK|tets" (src/IoMesh.cpp:145-147, test/unit_tests/UnitTestMeshUtils.cpp:54).
This module produces the same family of meshes as plain arrays:

* a structured lattice of Hex8 elements with the node ordering of
  include/ShapeFunctionsFunctorHex8.h (0:(-,-,-) 1:(+,-,-) 2:(+,+,-) 3:(-,+,-), 4-7 at +z);
* an optional split of every hex into 6 Tet4 sharing the body diagonal 0-6
  (the split the reference's kernel-radius test describes: one pair of tets per
  hex across the diagonal, the rest across faces; NeighborSearchProcessorTest.cpp:193-203);
* counter-based random fields: value i of stream s under seed k is
  splitmix64(k ^ (s << 40) ^ i), mapped to [0, 1) with 53 bits.

Everything is numpy; nothing here runs on the GPU.
"""
import math

import numpy as np

DEFAULT_SEED = 20261202

_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def splitmix64(x):
    x = np.asarray(x, dtype=np.uint64)
    with np.errstate(over="ignore"):
        x = x + np.uint64(0x9E3779B97F4A7C15)
        x = (x ^ (x >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        x = (x ^ (x >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        x = x ^ (x >> np.uint64(31))
    return x


def uniform01(seed, stream, n):
    """n uniform doubles in [0, 1) from counter-based splitmix64."""
    idx = np.arange(n, dtype=np.uint64)
    key = np.uint64(seed) ^ (np.uint64(stream) << np.uint64(40))
    r = splitmix64(key ^ idx)
    return (r >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def normal(seed, stream, n):
    """n standard normal doubles (Box-Muller on two uniform streams)."""
    u1 = uniform01(seed, 2 * stream + 1000, n)
    u2 = uniform01(seed, 2 * stream + 1001, n)
    u1 = np.maximum(u1, 1e-300)
    return np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * math.pi * u2)


def hex_lattice(nx, ny, nz, lx=1.0, ly=1.0, lz=1.0, origin=(0.0, 0.0, 0.0)):
    """Structured (nx x ny x nz) Hex8 block.  Returns X (n_nodes x 3), conn (n_elem x 8)."""
    ii, jj, kk = np.meshgrid(np.arange(nx + 1), np.arange(ny + 1), np.arange(nz + 1), indexing="ij")
    nid = lambda i, j, k: i + (nx + 1) * (j + (ny + 1) * k)
    X = np.zeros(((nx + 1) * (ny + 1) * (nz + 1), 3))
    ids = nid(ii, jj, kk).ravel()
    X[ids, 0] = origin[0] + ii.ravel() * (lx / nx)
    X[ids, 1] = origin[1] + jj.ravel() * (ly / ny)
    X[ids, 2] = origin[2] + kk.ravel() * (lz / nz)
    e_i, e_j, e_k = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    # element order: x fastest, then y, then z
    e_i = e_i.transpose(2, 1, 0).ravel()
    e_j = e_j.transpose(2, 1, 0).ravel()
    e_k = e_k.transpose(2, 1, 0).ravel()
    conn = np.stack([nid(e_i, e_j, e_k), nid(e_i + 1, e_j, e_k), nid(e_i + 1, e_j + 1, e_k), nid(e_i, e_j + 1, e_k),
                     nid(e_i, e_j, e_k + 1), nid(e_i + 1, e_j, e_k + 1), nid(e_i + 1, e_j + 1, e_k + 1),
                     nid(e_i, e_j + 1, e_k + 1)], axis=1).astype(np.int32)
    return X, conn


# The 6 tets of a hex that share the diagonal 0-6, one per axis ordering of the
# path 0 -> 6; each listed as (0, a, b, 6) with positive orientation.
_KUHN = [(0, 1, 2, 6), (0, 5, 1, 6), (0, 2, 3, 6), (0, 3, 7, 6), (0, 4, 5, 6), (0, 7, 4, 6)]


def split_tets(X, hex_conn):
    """Split every Hex8 into 6 positively oriented Tet4 around the 0-6 diagonal."""
    hc = np.asarray(hex_conn)
    tets = np.concatenate([hc[:, list(t)][:, None, :] for t in _KUHN], axis=1).reshape(-1, 4)
    v = signed_volumes(X, tets)
    neg = v < 0
    tets[neg] = tets[neg][:, [0, 2, 1, 3]]
    return tets.astype(np.int32)


def signed_volumes(X, tets):
    a = X[tets[:, 1]] - X[tets[:, 0]]
    b = X[tets[:, 2]] - X[tets[:, 0]]
    c = X[tets[:, 3]] - X[tets[:, 0]]
    return np.einsum("ij,ij->i", a, np.cross(b, c)) / 6.0


def jitter(X, h, frac, seed=DEFAULT_SEED, stream=1):
    """X + frac*h*U[-1,1] per coordinate (counter-based)."""
    r = uniform01(seed, stream, X.size).reshape(X.shape)
    return X + frac * h * (2.0 * r - 1.0)


def lame(E, nu):
    """(lambda, mu) from E and nu, include/Materials/Base.h:102-107."""
    return (E * nu) / ((1 + nu) * (1 - 2 * nu)), E / (2 * (1 + nu))


def stable_dt_estimate(X, conn, E, nu, rho):
    """Conservative element wave-transit bound on the stable increment: min_e(l_e) / c_d,
    c_d = sqrt((lambda+2mu)/rho).  Tet4: l_e = (3V/A_max)/sqrt2 (smallest altitude,
    scaled for the lumped-mass Kuhn tets); Hex8: l_e = 0.8 x the shortest edge.  Both sit
    below the power-method value 2/sqrt(lambda_max(M^-1 K)) on lattice meshes."""
    lam, mu = lame(E, nu)
    c = math.sqrt((lam + 2 * mu) / rho)
    conn = np.asarray(conn)
    npe = conn.shape[1]
    if npe == 4:
        x = [X[conn[:, k]] for k in range(4)]
        vol = np.abs(np.einsum("ij,ij->i", x[1] - x[0], np.cross(x[2] - x[0], x[3] - x[0]))) / 6.0
        amax = np.zeros(len(conn))
        for (a, b, d) in ((0, 1, 2), (0, 1, 3), (0, 2, 3), (1, 2, 3)):
            amax = np.maximum(amax, 0.5 * np.linalg.norm(np.cross(x[b] - x[a], x[d] - x[a]), axis=1))
        return float((3.0 * vol / amax).min()) / math.sqrt(2.0) / c
    lmin = np.inf
    for a in range(npe):
        for b in range(a + 1, npe):
            d = np.linalg.norm(X[conn[:, a]] - X[conn[:, b]], axis=1)
            lmin = min(lmin, float(d.min()))
    return 0.8 * lmin / c


def thex(X, tets):
    """Split every tetrahedron into four hexahedra ("thex"): one hex per tet vertex v, with
    v, the midpoints of the three edges at v, the centroids of the three faces at v and the
    tet's centroid.  Nodes: the tet vertices keep their ids 0..n-1 (the lowest ids, so each
    hex's lowest id is its tet vertex), then the edge midpoints, the face centroids and the
    tet centroids.  Hex node order: v, m(v,p), c(v,p,q), m(v,q), m(v,r), c(v,p,r), centroid,
    c(v,q,r) with (p - v, q - v, r - v) right-handed, so every hex has positive volume."""
    X = np.asarray(X, dtype=np.float64)
    tets = np.asarray(tets, dtype=np.int64)
    nv, nt = len(X), len(tets)
    edges = np.sort(tets[:, [[0, 1], [0, 2], [0, 3], [1, 2], [1, 3], [2, 3]]].reshape(-1, 2), axis=1)
    uedges, einv = np.unique(edges, axis=0, return_inverse=True)
    faces = np.sort(tets[:, [[0, 1, 2], [0, 1, 3], [0, 2, 3], [1, 2, 3]]].reshape(-1, 3), axis=1)
    ufaces, finv = np.unique(faces, axis=0, return_inverse=True)
    ne_, nf = len(uedges), len(ufaces)
    Xh = np.concatenate([X, 0.5 * (X[uedges[:, 0]] + X[uedges[:, 1]]),
                         (X[ufaces[:, 0]] + X[ufaces[:, 1]] + X[ufaces[:, 2]]) / 3.0,
                         X[tets].mean(axis=1)])
    einv = einv.reshape(nt, 6) + nv
    finv = finv.reshape(nt, 4) + nv + ne_
    cid = nv + ne_ + nf + np.arange(nt)
    eslot = {(0, 1): 0, (0, 2): 1, (0, 3): 2, (1, 2): 3, (1, 3): 4, (2, 3): 5}
    fslot = {(0, 1, 2): 0, (0, 1, 3): 1, (0, 2, 3): 2, (1, 2, 3): 3}
    hexes = np.zeros((4 * nt, 8), dtype=np.int64)
    for lv in range(4):
        others = [k for k in range(4) if k != lv]
        v = tets[:, lv]
        p, q, r = (tets[:, k] for k in others)
        det = np.einsum("ij,ij->i", X[p] - X[v], np.cross(X[q] - X[v], X[r] - X[v]))
        for t in range(nt):
            a, b, c = others if det[t] > 0 else (others[0], others[2], others[1])
            E = lambda i, j: einv[t, eslot[tuple(sorted((i, j)))]]
            F = lambda i, j, k: finv[t, fslot[tuple(sorted((i, j, k)))]]
            hexes[4 * t + lv] = [tets[t, lv], E(lv, a), F(lv, a, b), E(lv, b), E(lv, c), F(lv, a, c), cid[t],
                                 F(lv, b, c)]
    return Xh, hexes.astype(np.int32)


def thex_block(n=3, seed=DEFAULT_SEED, jitter_frac=0.1):
    """A jittered n^3 hex lattice on [0,1]^3, split into Kuhn tets, each tet into four hexes."""
    X, hc = hex_lattice(n, n, n)
    X = jitter(X, 1.0 / n, jitter_frac, seed, 1)
    return thex(X, split_tets(X, hc))


# ---------------- named workloads ----------------

def plain_hex_cube(n=20, seed=DEFAULT_SEED):
    """Unit cube of n^3 Hex8, v0 ~ U[-1,1]^3, z=0 face fixed; NH rho=7850, E=200 GPa, nu=0.3."""
    X, conn = hex_lattice(n, n, n)
    v0 = 2.0 * uniform01(seed, 2, X.size).reshape(X.shape) - 1.0
    fixed = np.zeros(X.shape, dtype=np.uint8)
    fixed[np.abs(X[:, 2]) < 1e-12, :] = 1
    v0[fixed.astype(bool)] = 0.0
    return dict(X=X, conn=conn, etype=1, v0=v0, fixed=fixed, rho=7850.0, E=200e9, nu=0.3)


def large_tet_block(n=100, seed=DEFAULT_SEED, jitter_frac=0.1, sigma=0.5):
    """[0,1]^3 with n^3 hexes split into 6 Kuhn tets; nodes jittered by 10% h; v0 ~ N(0, 0.5^2)."""
    X, hc = hex_lattice(n, n, n)
    tets = split_tets(X, hc)
    X = jitter(X, 1.0 / n, jitter_frac, seed, 1)
    v0 = sigma * normal(seed, 3, X.size).reshape(X.shape)
    return dict(X=X, conn=tets, etype=0, v0=v0, fixed=None, rho=7850.0, E=200e9, nu=0.3)


def rkpm_lattice(n_nodes_edge=64, seed=DEFAULT_SEED, jitter_frac=0.1):
    """n^3-node jittered lattice on [0,1]^3 with a Kuhn Tet4 background mesh; J2 aluminium-like
    material; linear velocity ramp about the centre with seeded amplitude s ~ U[0.9, 1.1]."""
    m = n_nodes_edge - 1
    X, hc = hex_lattice(m, m, m)
    tets = split_tets(X, hc)
    X = jitter(X, 1.0 / m, jitter_frac, seed, 1)
    s = 0.9 + 0.2 * uniform01(seed, 4, 1)[0]
    return dict(X=X, conn=tets, etype=0, rho=2700.0, E=70e9, nu=0.33, yield_stress=300e6, hardening=1e9,
                ramp_scale=s)


def rkpm_velocity(X, vstar, s):
    c = X - 0.5
    return s * vstar * np.stack([-0.5 * c[:, 0], -0.5 * c[:, 1], c[:, 2]], axis=1)


def ensemble(n_specimens=256, seed=DEFAULT_SEED, nmin=8, nmax=48):
    """Uneven Hex8 cube specimens: edge counts log-uniform in [nmin, nmax], E ~ U[50,250] GPa,
    nu ~ U[0.2, 0.45], rho = 7850, seeded initial velocity U[-1,1]^3 per node."""
    u = uniform01(seed, 10, n_specimens)
    edges = np.rint(np.exp(math.log(nmin) + u * (math.log(nmax) - math.log(nmin)))).astype(int)
    Es = 50e9 + 200e9 * uniform01(seed, 11, n_specimens)
    nus = 0.2 + 0.25 * uniform01(seed, 12, n_specimens)
    specs = []
    for s in range(n_specimens):
        n = int(edges[s])
        X, conn = hex_lattice(n, n, n)
        v0 = 2.0 * uniform01(seed + 7919 * (s + 1), 2, X.size).reshape(X.shape) - 1.0
        specs.append(dict(X=X, conn=conn, etype=1, v0=v0, E=float(Es[s]), nu=float(nus[s]), rho=7850.0, n=n))
    return specs


def fourier_field(points, seed, stream, n_modes=64, length=0.2):
    """Smooth seeded random field of about unit variance: sqrt(2/M) sum_m cos(k_m . x + phi_m),
    k_m ~ N(0, I / length^2), phi_m ~ U[0, 2 pi)  (a random-Fourier-feature Gaussian field
    with squared-exponential correlation of length `length`)."""
    k = normal(seed, stream, 3 * n_modes).reshape(n_modes, 3) / length
    phi = 2.0 * math.pi * uniform01(seed, stream + 1, n_modes)
    return math.sqrt(2.0 / n_modes) * np.cos(points @ k.T + phi).sum(axis=1)


def adjoint_block(n=64, seed=DEFAULT_SEED):
    """n^3 Hex8 on [0,1]^3 with per-element E, nu from smooth random fields around 200 GPa
    (+-20 % at two sigma) and 0.3 (+-0.05); v0 ~ U[-1,1]^3 m/s; the target displacement
    comes from the same forward run with E scaled by (1 + 0.05 g), g a third field."""
    X, conn = hex_lattice(n, n, n)
    cen = X[conn].mean(axis=1)
    E = 200e9 * (1.0 + 0.1 * fourier_field(cen, seed, 30))
    nu = 0.3 + 0.025 * fourier_field(cen, seed, 32)
    E_target = E * (1.0 + 0.05 * fourier_field(cen, seed, 34))
    v0 = 2.0 * uniform01(seed, 36, X.size).reshape(X.shape) - 1.0
    return dict(X=X, conn=conn, etype=1, E=E, nu=nu, E_target=E_target, v0=v0, rho=7850.0)
