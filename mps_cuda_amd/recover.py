"""Batched parameter recovery with the adjoint.

`scenes` independent scenes (sizes log-uniform in [n_min, n_max], seeded; blocks or spheres
at 0.02 m spacing with a seeded smooth initial velocity) each hide a viscosity in
U(0.005, 0.05) and a surface-tension coefficient in U(0.01, 0.5). Their `steps`-step
trajectories are the targets. Starting from (0.025, 0.1) in every scene, Adam on the
logarithms of the coefficients minimises the sum over scenes of the mean squared distance of
the final positions to the target, with gradients from the batched adjoint (all scenes in one
context, one launch per phase). Prints one progress line per iteration and a JSON summary.

    python -m mps_cuda_amd.recover [--scenes 32] [--steps 80] [--iters 50] [--out result.json]
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np
import torch

from . import scenes as S
from .autograd import DifferentiableRollout


def make_scenes(n_scenes: int, seed: int, n_min: int, n_max: int, steps: int):
    sc = S.batch_scenes(n_scenes=n_scenes, seed=seed, n_min=n_min, n_max=n_max, steps=steps)
    rng = np.random.default_rng(seed + 100)
    hidden_mu = rng.uniform(0.005, 0.05, n_scenes)
    hidden_sig = rng.uniform(0.01, 0.5, n_scenes)
    for q, s in enumerate(sc):
        s.config.update(zsort_interval=0, cg_tol=-1.0e-8, cg_max_iter=200)
        s.vel = S.fourier_velocity(s.pos, seed + 1000 + q)
    return sc, hidden_mu, hidden_sig


def run(n_scenes=32, steps=80, iters=50, seed=4, n_min=16384, n_max=131072, lr=0.1, decay=0.94, every=10,
        log=sys.stderr):
    sc, mu_true, sig_true = make_scenes(n_scenes, seed, n_min, n_max, steps)
    roll = DifferentiableRollout(sc, steps, checkpoint_every=every)
    v0 = torch.as_tensor(np.concatenate([s.vel for s in sc]), dtype=torch.float32)
    n_total = roll.n
    with torch.no_grad():
        t0 = time.perf_counter()
        target = roll(v0, torch.as_tensor(mu_true), torch.as_tensor(sig_true))
        t_target = time.perf_counter() - t0
    offs = np.cumsum([0] + [s.n for s in sc])
    theta = torch.tensor(np.log(np.stack([np.full(n_scenes, 0.025), np.full(n_scenes, 0.1)])), requires_grad=True)
    # eps far below the gradients (~1e-10 for the surface tension), so every log-coefficient
    # moves by about lr per step whatever the scale of its gradient; the rate decays by `decay`
    opt = torch.optim.Adam([theta], lr=lr, eps=1e-30)
    sched = torch.optim.lr_scheduler.ExponentialLR(opt, gamma=decay)
    history = []
    t_fb = 0.0
    for it in range(iters):
        opt.zero_grad()
        t0 = time.perf_counter()
        x = roll(v0, torch.exp(theta[0]), torch.exp(theta[1]))
        loss = sum(((x[offs[q]:offs[q + 1]] - target[offs[q]:offs[q + 1]]) ** 2).mean() for q in range(n_scenes))
        loss.backward()
        t_fb += time.perf_counter() - t0
        opt.step()
        sched.step()
        mu, sig = np.exp(theta.detach().numpy())
        err_mu = np.abs(mu - mu_true) / mu_true
        err_sig = np.abs(sig - sig_true) / sig_true
        lv = loss.item()
        history.append(dict(iter=it, loss=lv, mu_err_median=float(np.median(err_mu)),
                            sigma_err_median=float(np.median(err_sig))))
        print(f"iter {it:3d} loss {lv:.4e}  median rel err viscosity {np.median(err_mu):.3e} "
              f"surface tension {np.median(err_sig):.3e}  ({t_fb / (it + 1):.2f} s/iter)", file=log, flush=True)
    mu, sig = np.exp(theta.detach().numpy())
    return {
        "scenes": n_scenes, "particles": int(n_total), "steps": steps, "adam_iterations": iters, "lr": lr,
        "lr_decay": decay,
        "checkpoint_every": every, "sizes": f"log-uniform {n_min}..{n_max}, seed {seed}",
        "target_rollout_s": t_target,
        "forward_adjoint_particle_steps_per_s": n_total * steps * iters / t_fb,
        "seconds_per_iteration": t_fb / iters,
        "viscosity": {"true": mu_true.tolist(), "recovered": mu.tolist(),
                      "rel_err_median": float(np.median(np.abs(mu - mu_true) / mu_true)),
                      "rel_err_max": float(np.max(np.abs(mu - mu_true) / mu_true))},
        "surface_tension": {"true": sig_true.tolist(), "recovered": sig.tolist(),
                            "rel_err_median": float(np.median(np.abs(sig - sig_true) / sig_true)),
                            "rel_err_max": float(np.max(np.abs(sig - sig_true) / sig_true))},
        "history": history,
    }


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--scenes", type=int, default=32)
    ap.add_argument("--steps", type=int, default=80)
    ap.add_argument("--iters", type=int, default=50)
    ap.add_argument("--seed", type=int, default=4)
    ap.add_argument("--n-min", type=int, default=16384)
    ap.add_argument("--n-max", type=int, default=131072)
    ap.add_argument("--lr", type=float, default=0.1)
    ap.add_argument("--decay", type=float, default=0.94)
    ap.add_argument("--out", default=None)
    a = ap.parse_args()
    res = run(a.scenes, a.steps, a.iters, a.seed, a.n_min, a.n_max, a.lr, a.decay)
    s = json.dumps({k: v for k, v in res.items() if k != "history"})
    print(s, flush=True)
    if a.out:
        with open(a.out, "w") as f:
            json.dump(res, f, indent=1)
    return 0


if __name__ == "__main__":
    sys.exit(main())
