"""The gfx950 adjoint against the FP64 reference: gradients of a linear functional of the final
positions with respect to the initial velocities, the viscosity and the surface-tension
coefficient, through a short rollout of a falling droplet, compared with torch autograd on the
FP64 replay of the same frozen map (tests/adjoint_ref.py, itself pinned by finite differences
and by the FP64 oracle). Checkpointing must not change the result."""
import sys

import numpy as np
import pytest
import torch

from mps_cuda_amd import HipSim, scenes
from tests.adjoint_ref import Replay

pytestmark = pytest.mark.gpu
STEPS = 6


def say(m):
    print(m, flush=True)
    sys.stdout.flush()


def _rel(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


def _load(sc):
    g = HipSim(scenes.make_config(sc))
    scenes.load_into(g, sc)
    g.set("FLUID_VEL", sc.extra["v0"])
    return g


def test_adjoint_matches_fp64_reference():
    sc = scenes.diff_droplet(n=300, radius=0.08, seed=3, steps=STEPS, h=0.04)
    g = _load(sc)
    counts = g.adj_rollout(STEPS, 4)
    xT = g.get("FLUID_POS")
    say(f"[adjoint] loop counts per step (l_df, l_cg, l_cd): {counts[:, 0].tolist()}")
    rng = np.random.default_rng(0)
    w = rng.normal(size=(sc.n, 3))
    gv0, gmu, gsig = g.adj_backward(w.astype(np.float32))
    assert np.array_equal(g.get("FLUID_POS"), xT)  # left at the final state

    rep = Replay(sc.pos, sc.radius[0], g.get("FLUID_MASS")[0], g.get("BND_POS"), g.get("BND_RADIUS")[0], sc.physics,
                 sc.material)
    v0 = torch.as_tensor(sc.extra["v0"]).clone().requires_grad_(True)
    mu = torch.tensor(sc.material["viscosity"], dtype=torch.float64, requires_grad=True)
    sig = torch.tensor(sc.material["surface_tension"], dtype=torch.float64, requires_grad=True)
    x, _ = rep.rollout(v0, mu, sig, counts[:, 0, :])
    assert _rel(x.detach().numpy(), xT) < 1e-5
    L = (torch.as_tensor(w) * x).sum()
    rv, rm, rs = torch.autograd.grad(L, (v0, mu, sig))
    e_v, e_mu, e_sig = _rel(gv0, rv.numpy()), _rel([gmu[0]], [rm.item()]), _rel([gsig[0]], [rs.item()])
    say(f"[adjoint] rel err vs FP64: v0 {e_v:.2e}  mu {e_mu:.2e} ({gmu[0]:.6e} vs {rm.item():.6e})  "
        f"sigma {e_sig:.2e} ({gsig[0]:.6e} vs {rs.item():.6e})")
    assert e_v < 1e-3 and e_mu < 1e-3 and e_sig < 1e-3


def test_checkpoint_interval_does_not_change_gradients():
    sc = scenes.diff_droplet(n=300, radius=0.08, seed=4, steps=STEPS, h=0.04)
    w = np.random.default_rng(1).normal(size=(sc.n, 3)).astype(np.float32)
    out = []
    for every in (1, 3, STEPS):
        g = _load(sc)
        g.adj_rollout(STEPS, every)
        out.append(g.adj_backward(w))
    for o in out[1:]:
        assert np.array_equal(o[0], out[0][0]) and np.array_equal(o[1], out[0][1]) and np.array_equal(o[2], out[0][2])
