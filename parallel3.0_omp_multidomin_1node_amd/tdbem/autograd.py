"""torch.autograd.Function over the reverse-time adjoint of the C-ABI.

The misfit J(materials) = sum over steps K >= 1 and receiver elements r of
|u_K[r] - u_obs[K][r]|^2 (global-frame displacements) of one multi-domain
case; its gradient with respect to the per-domain (E, v, rho) comes from
tdbem_adjoint (reverse-time adjoint on the device).  torch only carries the
material tensor (CPU, float64) and the chain rule; the solves run through the
same C-ABI as every other entry point.
"""
from __future__ import annotations

import numpy as np
import torch

from . import Case


class MisfitProblem:
    """Everything but the materials: mesh, boundary ids, interfaces, domain
    ids, time step, load history, receivers and observations."""

    def __init__(self, points, ele, bcid, dt, gap, domain_ids, itf_rows, known, nstep, receivers, obs,
                 tol=1e-12, fd_step=1e-6, restart=200):
        self.points, self.ele, self.bcid = points, ele, bcid
        self.dt, self.gap = float(dt), float(gap)
        self.domain_ids = np.asarray(domain_ids, np.float64)
        self.itf_rows = itf_rows
        self.known = np.asarray(known, np.float64)
        self.nstep = int(nstep)
        self.receivers = np.asarray(receivers, np.int32)
        self.obs = np.asarray(obs, np.float64)
        self.tol, self.fd_step, self.restart = tol, fd_step, restart

    def evaluate(self, evr):
        mats = np.column_stack([self.domain_ids, np.asarray(evr, np.float64).reshape(-1, 3)])
        c = Case(self.points, self.ele, self.bcid, self.dt, self.gap, domain_count=len(self.domain_ids),
                 mat_rows=mats, itf_rows=self.itf_rows)
        try:
            c.assemble_step0(10)
            c.assemble_history(self.nstep)
            J, g, _ = c.adjoint(self.known, self.nstep, self.receivers, self.obs, restart=self.restart, tol=self.tol,
                                fd_step=self.fd_step)
        finally:
            c.close()
        return J, g


class MaterialMisfit(torch.autograd.Function):
    @staticmethod
    def forward(ctx, materials, problem):
        J, g = problem.evaluate(materials.detach().cpu().numpy())
        ctx.save_for_backward(torch.from_numpy(np.ascontiguousarray(g)).to(materials.dtype))
        return torch.tensor(J, dtype=materials.dtype)

    @staticmethod
    def backward(ctx, grad_out):
        (g,) = ctx.saved_tensors
        return grad_out * g, None


def misfit(materials: torch.Tensor, problem: MisfitProblem) -> torch.Tensor:
    """J as a differentiable function of the (domains, 3) material tensor."""
    return MaterialMisfit.apply(materials, problem)
