"""torch.autograd over the gfx950 rollout: positions as a differentiable function of the
initial velocities and of each scene's viscosity and surface-tension coefficient.

The forward pass steps the library (mps_adj_rollout_traj, checkpoints every `checkpoint_every`
steps) and returns the positions after the observed steps; the backward pass calls the
library's hand-written adjoint (mps_adj_backward_traj), which recomputes each segment from its
checkpoint and adds each observation's cotangent where that state was produced. Tensors cross
the C ABI as host arrays.

The library keeps one recorded rollout per context. A backward pass whose rollout has since been
replaced by another forward call raises instead of returning the later rollout's gradients.
"""
from __future__ import annotations

import numpy as np
import torch

from . import scenes as _scenes
from ._capi import HipSim


class _Rollout(torch.autograd.Function):
    @staticmethod
    def forward(ctx, v0, viscosity, surface_tension, owner, obs_steps):
        pos = owner._run(v0.detach().cpu().numpy(), viscosity.detach().cpu().numpy(),
                         surface_tension.detach().cpu().numpy(), obs_steps)
        ctx.owner = owner
        ctx.obs_steps = obs_steps
        ctx.generation = owner.sim.rollout_generation
        ctx.like = [(t.device, t.dtype) for t in (v0, viscosity, surface_tension)]
        return torch.from_numpy(pos).to(device=v0.device)

    @staticmethod
    def backward(ctx, gx):
        owner = ctx.owner
        if owner.sim.rollout_generation != ctx.generation:
            raise RuntimeError("this rollout's recorded state was replaced by a later forward call; "
                               "call backward before running the rollout again")
        g = gx.detach().to("cpu", torch.float32).numpy()
        gv0, gmu, gsig = owner.sim.adj_backward_traj(ctx.obs_steps, g)
        grads = [torch.from_numpy(a).to(device=d, dtype=t) for a, (d, t) in zip((gv0, gmu, gsig), ctx.like)]
        return grads[0], grads[1], grads[2], None, None


class DifferentiableRollout:
    """One or several scenes (a batch) stepped `steps` times from their initial state.

    __call__(v0 (n, 3), viscosity (scenes,), surface_tension (scenes,)) -> final positions
    (n, 3); trajectory(..., obs_steps) -> positions after each observed step (k, n, 3). All in
    reference order (scene after scene); differentiable with torch.autograd.
    """

    def __init__(self, scene_list: list, steps: int, checkpoint_every: int = 10):
        self.scene_list = list(scene_list)
        self.steps = int(steps)
        self.every = int(checkpoint_every)
        self.sim = HipSim(_scenes.make_config(self.scene_list[0]))
        self.batch = len(self.scene_list) > 1
        if self.batch:
            self.walls = [_scenes.sample_walls(HipSim, s) for s in self.scene_list]
        self.n = sum(s.n for s in self.scene_list)
        self.loop_counts = None

    def reset(self):
        if self.batch:
            _scenes.load_batch(self.sim, self.scene_list, self.walls)
        else:
            _scenes.load_into(self.sim, self.scene_list[0])

    def _run(self, v0, mu, sigma, obs_steps):
        self.reset()
        for q, sc in enumerate(self.scene_list):
            m = dict(sc.material, viscosity=float(mu[q]), surface_tension=float(sigma[q]))
            self.sim.scene_set_material(q, **m)
        self.sim.set("FLUID_VEL", np.asarray(v0, np.float32).reshape(self.n, 3))
        pos, _, self.loop_counts = self.sim.adj_rollout_traj(self.steps, obs_steps, self.every)
        return pos

    def __call__(self, v0, viscosity, surface_tension):
        return _Rollout.apply(v0, viscosity, surface_tension, self, (self.steps,))[0]

    def trajectory(self, v0, viscosity, surface_tension, obs_steps):
        steps = tuple(int(t) for t in obs_steps)
        return _Rollout.apply(v0, viscosity, surface_tension, self, steps)
