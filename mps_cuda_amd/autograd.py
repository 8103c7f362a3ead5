"""torch.autograd over the gfx950 rollout: final positions as a differentiable function of the
initial velocities and of each scene's viscosity and surface-tension coefficient.

The forward pass steps the library (mps_adj_rollout, checkpoints every `checkpoint_every`
steps); the backward pass calls the library's hand-written adjoint (mps_adj_backward), which
recomputes each segment from its checkpoint. Tensors cross the C ABI as host arrays.
"""
from __future__ import annotations

import numpy as np
import torch

from . import scenes as _scenes
from ._capi import HipSim


class _Rollout(torch.autograd.Function):
    @staticmethod
    def forward(ctx, v0, viscosity, surface_tension, owner):
        owner._run(v0.detach().cpu().numpy(), viscosity.detach().cpu().numpy(), surface_tension.detach().cpu().numpy())
        ctx.owner = owner
        x = torch.from_numpy(owner.sim.get("FLUID_POS").astype(np.float32))
        return x.to(v0.device)

    @staticmethod
    def backward(ctx, gx):
        gv0, gmu, gsig = ctx.owner.sim.adj_backward(gx.detach().cpu().numpy())
        return torch.from_numpy(gv0), torch.from_numpy(gmu), torch.from_numpy(gsig), None


class DifferentiableRollout:
    """One or several scenes (a batch) stepped `steps` times from their initial state.

    __call__(v0 (n, 3), viscosity (scenes,), surface_tension (scenes,)) -> final positions
    (n, 3), all in reference order (scene after scene); differentiable with torch.autograd.
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

    def _run(self, v0, mu, sigma):
        self.reset()
        for q, sc in enumerate(self.scene_list):
            m = dict(sc.material, viscosity=float(mu[q]), surface_tension=float(sigma[q]))
            self.sim.scene_set_material(q, **m)
        self.sim.set("FLUID_VEL", np.asarray(v0, np.float32).reshape(self.n, 3))
        self.loop_counts = self.sim.adj_rollout(self.steps, self.every)

    def __call__(self, v0, viscosity, surface_tension):
        return _Rollout.apply(v0, viscosity, surface_tension, self)
