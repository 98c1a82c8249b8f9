"""Adjoint (VJP) kernels against central finite differences of the CPU oracle.

Bar: gradients agree with central differences (h = 1e-6 relative) to 1e-6 relative.
"""
import numpy as np
import pytest

import oracle as O
import aperi_hip as A
from aperi_hip import meshgen as M

pytestmark = pytest.mark.gpu

RHO = 7850.0


def _mesh(etype, n, seed):
    X, hc = M.hex_lattice(n, n, n)
    X = M.jitter(X, 1.0 / n, 0.1, seed, 1)
    conn = hc if etype == A.HEX8 else M.split_tets(X, hc)
    return X, conn


def _fields(ne, seed):
    E = 200e9 * (0.8 + 0.4 * M.uniform01(seed, 20, ne))
    nu = 0.25 + 0.1 * M.uniform01(seed, 21, ne)
    return E, nu


def _oracle_force(etype, X, conn, kind, E, nu, u):
    lam, mu = M.lame(E, nu)
    f, *_ = O.internal_force(etype, X, conn, kind, 0.0, 0.0, u, elem_lambda=lam, elem_mu=mu)
    return f


@pytest.mark.parametrize("etype", [A.TET4, A.HEX8], ids=["tet4", "hex8"])
@pytest.mark.parametrize("kind", [A.LINEAR_ELASTIC, A.NEO_HOOKEAN], ids=["le", "nh"])
def test_internal_force_vjp_vs_fd(gpu_available, etype, kind):
    X, conn = _mesh(etype, 3, 4)
    E, nu = _fields(len(conn), 2)
    u = 2e-2 * (2 * M.uniform01(6, 7, X.size).reshape(X.shape) - 1)
    fb = M.normal(8, 1, X.size).reshape(X.shape)
    with A.Solver() as s:
        s.set_mesh(etype, X, conn)
        s.set_material(kind, RHO, 200e9, 0.3)
        s.set_material_fields(E, nu)
        s.set_field(A.FIELD_U, u)
        ub, Eb, nb = s.internal_force_vjp(fb)
    # directional derivatives: <f_bar, df/du du> = <u_bar, du>, likewise for E and nu
    du = M.normal(9, 1, X.size).reshape(X.shape)
    h = 1e-6 * np.abs(u).max() / np.abs(du).max()
    fd = (np.sum(fb * _oracle_force(etype, X, conn, kind, E, nu, u + h * du)) -
          np.sum(fb * _oracle_force(etype, X, conn, kind, E, nu, u - h * du))) / (2 * h)
    assert np.sum(ub * du) == pytest.approx(fd, rel=1e-6)
    dE = M.normal(10, 1, len(conn)) * E
    hE = 1e-6
    fdE = (np.sum(fb * _oracle_force(etype, X, conn, kind, E + hE * dE, nu, u)) -
           np.sum(fb * _oracle_force(etype, X, conn, kind, E - hE * dE, nu, u))) / (2 * hE)
    assert np.sum(Eb * dE) == pytest.approx(fdE, rel=1e-6)
    dn = 0.01 * M.normal(11, 1, len(conn))
    hn = 1e-6
    fdn = (np.sum(fb * _oracle_force(etype, X, conn, kind, E, nu + hn * dn, u)) -
           np.sum(fb * _oracle_force(etype, X, conn, kind, E, nu - hn * dn, u))) / (2 * hn)
    assert np.sum(nb * dn) == pytest.approx(fdn, rel=1e-6)


def _oracle_final_u(etype, X, conn, kind, E, nu, v0, fixed, dt, n_steps):
    lam, mu = M.lame(E, nu)
    m = O.lumped_mass(etype, X, conn, RHO)
    s = O.Solver(etype, X, conn, kind, 0.0, 0.0, m, fixed, elem_lambda=lam, elem_mu=mu)
    s.v[:] = v0
    s.run(dt, n_steps)
    return s.u.copy()


@pytest.mark.parametrize("etype", [A.HEX8, A.TET4], ids=["hex8", "tet4"])
def test_adjoint_gradients_vs_central_fd(gpu_available, etype):
    X, conn = _mesh(etype, 3, 12)
    ne = len(conn)
    E, nu = _fields(ne, 5)
    v0 = 20.0 * M.normal(3, 3, X.size).reshape(X.shape)
    fixed = np.zeros(X.shape, dtype=np.uint8)
    fixed[X[:, 2] < 1e-3] = 1
    v0[fixed.astype(bool)] = 0.0
    dt = 0.5 * M.stable_dt_estimate(X, conn, 200e9, 0.3, RHO)
    n_steps, K = 40, 7
    E_t, nu_t = E * (1 + 0.05 * M.normal(7, 1, ne)), nu
    # target from the forward model with a perturbed field
    lam_t, mu_t = M.lame(E_t, nu_t)
    m = O.lumped_mass(etype, X, conn, RHO)
    tgt = O.Solver(etype, X, conn, O.NEO_HOOKEAN, 0.0, 0.0, m, fixed, elem_lambda=lam_t, elem_mu=mu_t)
    tgt.v[:] = v0
    tgt.run(dt, n_steps)
    target = tgt.u.copy()
    with A.Solver() as s:
        s.set_mesh(etype, X, conn)
        s.set_material(A.NEO_HOOKEAN, RHO, 200e9, 0.3)
        s.set_material_fields(E, nu)
        s.compute_mass()
        s.set_fixed(fixed)
        s.set_field(A.FIELD_V, v0)
        s.initialize()
        s.adjoint_forward(dt, n_steps, K)
        uT = s.get(A.FIELD_U)
        Eb, nb, v0b = s.adjoint_backward(uT - target)
        assert np.array_equal(s.get(A.FIELD_U), uT)  # the final state is restored
    uT_of = lambda E_, nu_, v_: _oracle_final_u(etype, X, conn, A.NEO_HOOKEAN, E_, nu_, v_, fixed, dt, n_steps)

    def central_difference(plus, minus, h):
        # L+ - L- = (u+ - u-) . ((u+ + u-)/2 - u*) exactly, without the cancellation of
        # subtracting two losses
        up, um = uT_of(*plus), uT_of(*minus)
        return np.sum((up - um) * (0.5 * (up + um) - target)) / (2 * h)

    rng = np.random.default_rng(1)
    checked = 0
    for e in rng.choice(ne, 4, replace=False):
        h = 1e-6 * E[e]
        Ep, Em = E.copy(), E.copy()
        Ep[e] += h
        Em[e] -= h
        fd = central_difference((Ep, nu, v0), (Em, nu, v0), h)
        assert Eb[e] == pytest.approx(fd, rel=1e-6)
        h = 1e-6 * nu[e]
        np_, nm = nu.copy(), nu.copy()
        np_[e] += h
        nm[e] -= h
        fd = central_difference((E, np_, v0), (E, nm, v0), h)
        assert nb[e] == pytest.approx(fd, rel=1e-6)
        checked += 2
    free = np.argwhere(~fixed.astype(bool))
    for i, d in free[rng.choice(len(free), 4, replace=False)]:
        # a velocity component can sit near zero: the step is relative to the field's amplitude
        h = 1e-6 * np.abs(v0).max()
        vp, vm = v0.copy(), v0.copy()
        vp[i, d] += h
        vm[i, d] -= h
        fd = central_difference((E, nu, vp), (E, nu, vm), h)
        assert v0b[i, d] == pytest.approx(fd, rel=1e-6)
        checked += 1
    assert checked == 12
    # fixed DOFs carry no velocity sensitivity
    assert np.all(v0b[fixed.astype(bool)] == 0.0)


def test_torch_autograd_function(gpu_available):
    import torch
    from aperi_hip.autograd import ExplicitDynamics, Problem
    X, conn = _mesh(A.HEX8, 2, 3)
    ne = len(conn)
    E, nu = _fields(ne, 9)
    v0 = 10.0 * M.normal(2, 3, X.size).reshape(X.shape)
    dt = 0.5 * M.stable_dt_estimate(X, conn, 200e9, 0.3, RHO)
    prob = Problem(X=X, conn=conn, dt=dt, n_steps=12, checkpoint_interval=5)
    Et = torch.tensor(E, requires_grad=True)
    nt = torch.tensor(nu, requires_grad=True)
    vt = torch.tensor(v0, requires_grad=True)
    uT = ExplicitDynamics.apply(Et, nt, vt, prob)
    loss = 0.5 * (uT ** 2).sum()
    loss.backward()
    with A.Solver() as s:
        s.set_mesh(A.HEX8, X, conn)
        s.set_material(A.NEO_HOOKEAN, RHO, 200e9, 0.3)
        s.set_material_fields(E, nu)
        s.compute_mass()
        s.set_field(A.FIELD_V, v0)
        s.initialize()
        s.adjoint_forward(dt, 12, 5)
        u = s.get(A.FIELD_U)
        Eb, nb, v0b = s.adjoint_backward(u)
    assert np.allclose(uT.detach().numpy(), u, rtol=0, atol=0)
    assert np.array_equal(Et.grad.numpy(), Eb) and np.array_equal(nt.grad.numpy(), nb)
    assert np.array_equal(vt.grad.numpy(), v0b)


@pytest.mark.parametrize("etype,big", [(A.TET4, (1024, 1024)), (A.HEX8, (256, 256))], ids=["tet4", "hex8"])
def test_vjp_independent_of_tile_shape(gpu_available, etype, big):
    """The VJP evaluates one element per thread in workgroups of at most 512 Tet4 or 128
    Hex8: a context planned with larger tiles is re-planned for it, and the result
    matches a context with a small plan to rounding.  The meshes (9^3 Hex8 = 729 Hex8,
    4374 Tet4) span several tiles of every shape."""
    X, conn = _mesh(etype, 9, 12)
    E, nu = _fields(len(conn), 3)
    u = 2e-2 * (2 * M.uniform01(16, 7, X.size).reshape(X.shape) - 1)
    fb = M.normal(18, 1, X.size).reshape(X.shape)
    out = []
    for tiling in ((64, 64), big):
        with A.Solver() as s:
            s.set_mesh(etype, X, conn)
            s.set_material(A.NEO_HOOKEAN, RHO, 200e9, 0.3)
            s.set_material_fields(E, nu)
            if tiling:
                s.set_tiling(*tiling)
            s.set_field(A.FIELD_U, u)
            out.append(s.internal_force_vjp(fb))
    for a, b in zip(*out):
        assert np.abs(a - b).max() <= 1e-12 * np.abs(b).max()


def test_renumbering_invalidates_the_recording(gpu_available):
    """A new tile plan renumbers the nodes; checkpoints recorded in the old numbering are
    dropped, and the backward pass asks for a new forward run instead of using them."""
    X, conn = _mesh(A.HEX8, 4, 2)
    v0 = 10.0 * M.normal(4, 3, X.size).reshape(X.shape)
    dt = 0.5 * M.stable_dt_estimate(X, conn, 200e9, 0.3, RHO)
    with A.Solver() as s:
        s.set_mesh(A.HEX8, X, conn)
        s.set_material(A.NEO_HOOKEAN, RHO, 200e9, 0.3)
        s.compute_mass()
        s.set_field(A.FIELD_V, v0)
        s.initialize()
        s.adjoint_forward(dt, 6, 3)
        g1 = s.adjoint_backward(s.get(A.FIELD_U))
        s.set_tiling(16, 128)
        with pytest.raises(A.AphipError, match="renumbered"):
            s.adjoint_backward(s.get(A.FIELD_U))
        s.set_tiling(64, 64)
        s.set_field(A.FIELD_V, v0)
        s.set_field(A.FIELD_U, np.zeros_like(X))
        s.initialize()
        s.adjoint_forward(dt, 6, 3)
        g2 = s.adjoint_backward(s.get(A.FIELD_U))
    for a, b in zip(g1, g2):
        assert np.abs(a - b).max() <= 1e-12 * np.abs(b).max()
