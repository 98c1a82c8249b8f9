"""torch.autograd binding of the explicit-dynamics adjoint.

`ExplicitDynamics.apply(E, nu, v0, problem)` maps per-element Young's moduli and
Poisson ratios and the initial nodal velocity to the displacement after
`problem.n_steps` central-difference steps, computed on the GPU by
libaperi_hip.so.  Its backward pass is the library's hand-written adjoint
(aphip_adjoint_backward: stress and force VJPs, reverse central-difference sweep,
checkpoint-and-recompute).  Tensors are float64 host tensors; the library owns
the device copies (it does not share torch's allocator or streams).
"""
from dataclasses import dataclass

import numpy as np
import torch

from . import FIELD_U, FIELD_V, HEX8, NEO_HOOKEAN, Solver


@dataclass
class Problem:
    X: np.ndarray
    conn: np.ndarray
    etype: int = HEX8
    kind: int = NEO_HOOKEAN
    density: float = 7850.0
    dt: float = 1e-7
    n_steps: int = 200
    checkpoint_interval: int = 20
    fixed: np.ndarray = None


class ExplicitDynamics(torch.autograd.Function):
    @staticmethod
    def forward(ctx, E, nu, v0, problem):
        s = Solver()
        s.set_mesh(problem.etype, problem.X, problem.conn)
        Em, num = E.detach().cpu().numpy(), nu.detach().cpu().numpy()
        s.set_material(problem.kind, problem.density, float(Em.mean()), float(num.mean()))
        s.set_material_fields(Em, num)
        s.compute_mass()
        if problem.fixed is not None:
            s.set_fixed(problem.fixed)
        s.set_field(FIELD_V, v0.detach().cpu().numpy())
        s.initialize()
        s.adjoint_forward(problem.dt, problem.n_steps, problem.checkpoint_interval)
        uT = torch.from_numpy(s.get(FIELD_U))
        ctx.solver = s
        return uT

    @staticmethod
    def backward(ctx, grad_uT):
        s = ctx.solver
        Eb, nb, v0b = s.adjoint_backward(grad_uT.detach().cpu().contiguous().numpy())
        s.close()
        ctx.solver = None
        return torch.from_numpy(Eb), torch.from_numpy(nb), torch.from_numpy(v0b), None
