"""torch.autograd view of the trajectory loss, for gradient checks.

TrajectoryLoss.apply(x0, eps, sigma, problem) runs the forward trajectory
and the hand-written adjoint in libtinymd_amd.so (tmd_traj_adjoint) and
hands the adjoint's dL/dx0, dL/deps, dL/dsigma to autograd, so a torch
graph built on top of the loss (a fit of eps and sigma, a check against
finite differences) sees the library's gradient.  Tensors are FP64 on the
host; the trajectory runs on the device.
"""
import numpy as np
import torch

from .sim import Simulation


class TrajectoryProblem:
    """Everything but (x0, eps, sigma): masses, v0, box, run and loss settings."""

    def __init__(self, masses, v0, aabb, steps, reneigh=20, dt=0.005, cutoff=2.5, skin=0.3,
                 u_target=0.0, p_target=0.0, precision="fp64", nb_cap=100):
        self.masses = np.asarray(masses, dtype=np.float64)
        self.v0 = np.asarray(v0, dtype=np.float64).reshape(-1, 3)
        self.aabb = np.asarray(aabb, dtype=np.float64)
        self.steps, self.reneigh, self.dt = steps, reneigh, dt
        self.cutoff, self.skin = cutoff, skin
        self.u_target, self.p_target = u_target, p_target
        self.precision = precision
        self.nb_cap = nb_cap

    def simulation(self, x0):
        return Simulation(self.masses, x0, self.v0, self.aabb, self.cutoff + self.skin, nb_cap=self.nb_cap,
                          precision=self.precision, lj_shift=True)

    def loss(self, x0, eps, sigma):
        """L only (tmd_traj_loss), for finite differences."""
        sim = self.simulation(x0)
        ub, pb = sim.traj_loss(self.steps, self.reneigh, self.dt, self.cutoff, self.skin, eps, sigma)
        return (ub - self.u_target) ** 2 + (pb - self.p_target) ** 2

    def adjoint(self, x0, eps, sigma):
        sim = self.simulation(x0)
        return sim.traj_adjoint(self.steps, self.reneigh, self.dt, self.cutoff, self.skin, eps, sigma,
                                self.u_target, self.p_target)


class TrajectoryLoss(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x0, eps, sigma, problem):
        r = problem.adjoint(x0.detach().cpu().numpy(), float(eps), float(sigma))
        ctx.save_for_backward(torch.from_numpy(r["grad_x0"]).to(x0.dtype),
                              torch.tensor(r["grad_eps"], dtype=torch.float64),
                              torch.tensor(r["grad_sigma"], dtype=torch.float64))
        return torch.tensor(r["loss"], dtype=torch.float64)

    @staticmethod
    def backward(ctx, g):
        gx, ge, gs = ctx.saved_tensors
        return g * gx, g * ge, g * gs, None
