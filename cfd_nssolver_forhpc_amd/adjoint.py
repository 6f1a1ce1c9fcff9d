"""The time-averaged DFIB drag of a cylinder run as a differentiable function
for PyTorch (an addition: the reference has no adjoint).

    problem = DragProblem(sim, nsteps=400, avg_from=200, ckpt_every=20)
    J = DragObjective.apply(u_in, gain, problem)
    J.backward()            # u_in.grad, gain.grad

``u_in`` is the inflow profile (ny_cal values, the face-0 u Dirichlet value
of BC_staggered_main per interior row) and ``gain`` the DFIB forcing gain
(u_new = (1 - gain eta_f) T2 in update_UandF).  The forward pass restarts the
run from the problem's initial state and runs the device timestep with
checkpoints (cfd_adjoint_forward); the backward pass is the discrete adjoint
sweep with recomputation (cfd_adjoint_backward).  Tensors carry only the
parameters and the gradients: the arithmetic runs in the HIP library.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np
import torch

from .solver import Simulation


@dataclass
class DragProblem:
    sim: Simulation
    nsteps: int
    avg_from: int
    ckpt_every: int
    noise: float = 1e-3
    seed: int = 1
    last_stats: dict = field(default_factory=dict)

    def run(self, u_in: np.ndarray, gain: float) -> float:
        """J at (u_in, gain) from the initial state (no gradient record kept)."""
        self.sim.set_inflow_profile(u_in)
        self.sim.set_forcing_gain(gain)
        self.sim.init_state(self.noise, self.seed)
        return self.sim.adjoint_forward(self.nsteps, self.avg_from, self.ckpt_every)


class DragObjective(torch.autograd.Function):
    """J(u_in, gain): mean corrector-pass cD over steps [avg_from, nsteps)."""

    @staticmethod
    def forward(ctx, u_in: torch.Tensor, gain: torch.Tensor, problem: DragProblem) -> torch.Tensor:
        if u_in.dim() != 1 or u_in.numel() != problem.sim.grid.ncal[1]:
            raise ValueError("u_in must be a 1-D tensor of ny_cal values")
        J = problem.run(u_in.detach().to("cpu", torch.float64).numpy(), float(gain.detach().cpu()))
        ctx.problem = problem
        ctx.generation = problem.sim.adjoint_generation()  # the run this graph node differentiates
        ctx.device = u_in.device
        return torch.tensor(J, dtype=torch.float64, device=u_in.device)

    @staticmethod
    def backward(ctx, grad_J: torch.Tensor):
        if ctx.problem.sim.adjoint_generation() != ctx.generation:
            raise RuntimeError("DragObjective.backward: the simulation has run another objective (or changed its "
                               "state or parameters) since this forward; evaluate J again before backward")
        g_uin, g_gain, stats = ctx.problem.sim.adjoint_backward()
        ctx.problem.last_stats = stats
        gJ = grad_J.to(torch.float64)
        gu = torch.from_numpy(g_uin).to(ctx.device) * gJ
        gg = torch.tensor(g_gain, dtype=torch.float64, device=ctx.device) * gJ
        return gu, gg, None


def smoothed_inflow(ny_cal: int, seed: int = 3, amplitude: float = 0.05, passes: int = 4) -> np.ndarray:
    """Seeded inflow profile 1 + amplitude N(0, 1), smoothed by `passes`
    applications of the [1/4, 1/2, 1/4] filter with edge values repeated."""
    base = 1.0 + amplitude * np.random.default_rng(seed).standard_normal(ny_cal)
    for _ in range(passes):
        base = np.convolve(np.pad(base, 1, mode="edge"), [0.25, 0.5, 0.25], mode="valid")
    return base
