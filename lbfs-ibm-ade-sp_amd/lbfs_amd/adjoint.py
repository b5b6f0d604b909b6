"""torch.autograd binding of the discrete adjoint (lbfs_adjoint_* of the C ABI).

`DeformationRun(solver, n_steps, every)` is a differentiable map from
(shear modulus, local bending modulus, body force x) of membrane `index` to
the node positions after `n_steps` coupled steps, always started from the
state the solver held at its first call.  The forward pass records the run
with checkpoints every `every` steps; the backward pass is the hand-written
adjoint on the device (segments replayed from their checkpoints), so its cost
does not grow with the number of parameters.  One solver records one run:
call backward for the latest forward before running the next.
"""
import numpy as np
import torch

from . import abi


class _Run(torch.autograd.Function):
    @staticmethod
    def forward(ctx, shear, local_bending, force_x, runner):
        s = runner.solver
        if runner.started:
            s.adjoint_rewind()
        s.set_membrane_moduli(runner.index, float(shear), float(local_bending), runner.global_bending)
        s.set_body_force([float(force_x), runner.force_yz[0], runner.force_yz[1]])
        s.adjoint_forward(runner.n_steps, runner.every)
        runner.started = True
        ctx.runner = runner
        X, _, _ = s.membrane_nodes()
        return torch.from_numpy(np.ascontiguousarray(X, dtype=np.float64))

    @staticmethod
    def backward(ctx, grad_X):
        r = ctx.runner
        g = r.solver.adjoint_backward(lam_nodes=grad_X.detach().to(torch.float64).cpu().numpy())
        gm = g["moduli"][r.index]
        t = lambda v: torch.tensor(float(v), dtype=torch.float64)
        return t(gm[0]), t(gm[1]), t(g["body_force"][0]), None


class DeformationRun:
    """Differentiable node positions after n_steps coupled steps."""

    def __init__(self, solver, n_steps, every, index=0, global_bending=None, force_yz=(0.0, 0.0)):
        self.solver = solver
        self.n_steps = int(n_steps)
        self.every = int(every)
        self.index = int(index)
        m = solver._membranes[self.index]
        self.global_bending = float(m.params["global_bending_modulus"] if global_bending is None else global_bending)
        self.force_yz = (float(force_yz[0]), float(force_yz[1]))
        self.started = False

    def __call__(self, shear, local_bending, force_x):
        return _Run.apply(shear, local_bending, force_x, self)
