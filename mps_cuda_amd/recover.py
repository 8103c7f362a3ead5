"""Batched parameter recovery with the adjoint.

`scenes` independent scenes (sizes log-uniform in [n_min, n_max], seeded; blocks or spheres at
0.02 m spacing with a seeded smooth initial velocity) each hide a viscosity in U(0.005, 0.05), a
surface-tension coefficient in U(0.01, 0.5) and their initial velocity field. Their `steps`-step
trajectories, observed every `obs_every` steps, are the targets. Starting from (0.025, 0.1) in
every scene and from zero initial velocities, each scene's sum of squared distances of the
observed positions to its target is minimised (the scenes share every launch but not their
unknowns).

Method: Gauss-Newton with per-scene backtracking. The residual r(p) (observed positions minus
target) is linearised in the unknowns p = (v0, log viscosity, log surface tension) of each scene,
and the step d minimises |r + J d|^2:
- J_v d_v, the response of the observed positions to an initial-velocity change, is a forward
  difference of two rollouts (the change scaled to `fd_v` m/s per scene);
- J_v^T u is one pass of the hand-written adjoint over the recorded rollout with the cotangent u
  on the observed positions;
- the two coefficient columns J_mu and J_sigma are forward differences in the logarithms.
The linear least-squares problem is solved by `inner` iterations of CGLS (conjugate gradients on
the normal equations; every scene has its own step lengths), with the coefficient columns scaled
to the initial velocities' response. With the initial velocities known the problem is 2x2 per
scene and is solved directly. Each scene then takes the largest of 1, 1/4, 1/16, 1/64 of its
step (after clipping the log steps to `max_log_step`) that lowers its loss, or keeps its values.

The initial velocities are sought as trilinear interpolation of nodal values on a lattice of
spacing `v0_grid` per scene (0: one unknown per particle). The initial velocity is only partly
observable: the first step's divergence-free solve removes its compressive part and viscosity
damps its fine scales. Besides the raw error, the summary reports the error of its response:
the velocity after the first step, and at the first observation, at the hidden coefficients,
from the recovered initial velocity against that from the hidden one, relative to the part of
it that the hidden initial velocity contributes; and the trajectory fit (rms distance to the
observed positions against their rms displacement from the start).

    python -m mps_cuda_amd.recover [--scenes 32] [--steps 80] [--iters 10] [--inner 6] [--out result.json]
    python -m mps_cuda_amd.recover --known-v0        # coefficients only, hidden v0 given
    python -m mps_cuda_amd.recover --scan 13         # loss against sigma at the hidden mu and v0
"""
from __future__ import annotations

import argparse
import json
import math
import sys
import time

import numpy as np
import torch

from . import scenes as S
from .autograd import DifferentiableRollout

MU0, SIGMA0 = 0.025, 0.1


def make_scenes(n_scenes: int, seed: int, n_min: int, n_max: int, steps: int):
    sc = S.batch_scenes(n_scenes=n_scenes, seed=seed, n_min=n_min, n_max=n_max, steps=steps)
    rng = np.random.default_rng(seed + 100)
    hidden_mu = rng.uniform(0.005, 0.05, n_scenes)
    hidden_sig = rng.uniform(0.01, 0.5, n_scenes)
    for q, s in enumerate(sc):
        s.config.update(zsort_interval=0, cg_tol=-1.0e-8, cg_max_iter=200)
        s.vel = S.fourier_velocity(s.pos, seed + 1000 + q)
    return sc, hidden_mu, hidden_sig


def _per_scene_loss(x, target, offs):
    """Mean squared distance per scene over the observed states: x, target (k, n, 3)."""
    return torch.stack([((x[:, offs[q]:offs[q + 1]] - target[:, offs[q]:offs[q + 1]]) ** 2).mean()
                        for q in range(len(offs) - 1)])


class _Batch:
    """Rollouts of one batch of scenes in two contexts: `tape` records the rollout whose adjoint
    gives J_v^T u; `probe` runs every other forward rollout (differences, line searches)."""

    def __init__(self, scene_list, steps, every, obs):
        self.tape = DifferentiableRollout(scene_list, steps, checkpoint_every=every)
        self.probe = DifferentiableRollout(scene_list, steps, checkpoint_every=steps)
        self.obs = tuple(int(t) for t in obs)
        self.sizes = torch.tensor([s.n for s in scene_list])
        self.S = len(scene_list)
        self.n = int(self.sizes.sum())
        self.sid = torch.repeat_interleave(torch.arange(self.S), self.sizes)
        self.fwd = [0, 0.0]
        self.bwd = [0, 0.0]

    def forward(self, v0, mu, sigma, record=False, obs=None):
        """Observed positions (k, n, 3) as float64; `record` keeps the rollout for vjp()."""
        ctx = self.tape if record else self.probe
        t0 = time.perf_counter()
        x = ctx._run(v0.to(torch.float32).numpy(), np.asarray(mu, np.float64), np.asarray(sigma, np.float64),
                     self.obs if obs is None else tuple(obs))
        self.fwd[0] += 1
        self.fwd[1] += time.perf_counter() - t0
        return torch.from_numpy(np.asarray(x)).to(torch.float64)

    def vjp(self, u):
        """J_v^T u for the recorded rollout: u (k, n, 3) on the observed positions."""
        t0 = time.perf_counter()
        gv0, _, _ = self.tape.sim.adj_backward_traj(self.obs, u.to(torch.float32).numpy())
        self.bwd[0] += 1
        self.bwd[1] += time.perf_counter() - t0
        return torch.from_numpy(gv0).to(torch.float64)

    def ssum(self, a):
        """Per-scene sums of a (k, n, 3) or (n, 3) array."""
        pp = a.sum(dim=(0, 2)) if a.dim() == 3 else a.sum(dim=1)
        return torch.zeros(self.S, dtype=torch.float64).index_add_(0, self.sid, pp)

    def smax(self, a):
        """Per-scene maximum of |a| for an (n, 3) array."""
        return torch.zeros(self.S, dtype=torch.float64).scatter_reduce_(0, self.sid, a.abs().amax(dim=1), "amax")

    def spread(self, s):
        """Per-scene values (S,) as an (n, 1) column."""
        return s[self.sid][:, None]


def _jv(b: _Batch, x, v0, mu, sig, pv, fd_v):
    """J_v pv by a forward difference, pv scaled per scene to a largest component of fd_v m/s."""
    m = b.smax(pv)
    eps = torch.where(m > 0, fd_v / torch.where(m > 0, m, torch.ones_like(m)), torch.ones_like(m))
    xp = b.forward(v0 + b.spread(eps) * pv, mu, sig)
    return (xp - x) / b.spread(eps)[None]


class _Grid:
    """Initial velocities as trilinear interpolation of nodal values on a per-scene lattice of
    spacing h (v0 = B c; the adjoint's gradient maps back as B^T g). Keeps the unknowns to the
    smooth fields the observations can determine: the first step's divergence-free solve and the
    viscosity remove most of what a finer field would add. h = 0: one unknown per particle."""

    def __init__(self, scene_list, h):
        self.h = float(h)
        if self.h <= 0.0:
            self.B = self.T = lambda a: a
            return
        idx, wts, nsid, base = [], [], [], 0
        for q, sc in enumerate(scene_list):
            p = np.asarray(sc.pos, np.float64)
            lo = p.min(axis=0) - 0.5 * self.h
            dims = np.floor((p.max(axis=0) - lo) / self.h).astype(np.int64) + 2
            f = (p - lo) / self.h
            i0 = np.floor(f).astype(np.int64)
            t = f - i0
            for c in range(8):
                o = np.array([(c >> 0) & 1, (c >> 1) & 1, (c >> 2) & 1])
                k = i0 + o
                node = base + (k[:, 0] * dims[1] + k[:, 1]) * dims[2] + k[:, 2]
                w = np.prod(np.where(o[None, :] == 1, t, 1.0 - t), axis=1)
                idx.append(node)
                wts.append(w)
            nsid.append(np.full(int(np.prod(dims)), q))
            base += int(np.prod(dims))
        n = sum(s.n for s in scene_list)
        self.idx = torch.as_tensor(np.stack([np.concatenate(idx[c::8]) for c in range(8)], axis=1))  # (n, 8)
        self.w = torch.as_tensor(np.stack([np.concatenate(wts[c::8]) for c in range(8)], axis=1))    # (n, 8)
        assert self.idx.shape == (n, 8)
        self.M = base
        self.node_sid = torch.as_tensor(np.concatenate(nsid))

        def B(c):  # nodes (M, 3) -> particles (n, 3)
            return (self.w[:, :, None] * c[self.idx]).sum(dim=1)

        def T(g):  # particles (n, 3) -> nodes (M, 3)
            return torch.zeros(self.M, 3, dtype=torch.float64).index_add_(
                0, self.idx.reshape(-1), (self.w[:, :, None] * g[:, None, :]).reshape(-1, 3))

        self.B, self.T = B, T

    def ssum(self, b: "_Batch", a):
        """Per-scene sums of an array in the unknowns' space."""
        if self.h <= 0.0:
            return b.ssum(a)
        return torch.zeros(b.S, dtype=torch.float64).index_add_(0, self.node_sid, a.sum(dim=1))

    def spread(self, b: "_Batch", s):
        return b.spread(s) if self.h <= 0.0 else s[self.node_sid][:, None]


def _cgls_v(b: _Batch, grid: _Grid, x, r, v0, mu, sig, inner, fd_v):
    """min |r + J_v B d|^2 per scene by `inner` CGLS iterations in the grid's unknowns (the
    coefficients held). Returns (B d, relative residual per scene)."""
    rho = -r
    r0 = b.ssum(rho * rho)
    s = grid.T(b.vjp(rho))
    p = s
    gamma = grid.ssum(b, s * s)
    d = torch.zeros_like(s)
    for i in range(inner):
        q = _jv(b, x, v0, mu, sig, grid.B(p), fd_v)
        delta = b.ssum(q * q)
        alpha = torch.where(delta > 0, gamma / delta.clamp_min(1e-300), torch.zeros_like(delta))
        d = d + grid.spread(b, alpha) * p
        rho = rho - b.spread(alpha)[None] * q
        if i == inner - 1:
            break
        s = grid.T(b.vjp(rho))
        gnew = grid.ssum(b, s * s)
        beta = torch.where(gamma > 0, gnew / gamma.clamp_min(1e-300), torch.zeros_like(gnew))
        p = s + grid.spread(b, beta) * p
        gamma = gnew
    return grid.B(d), (b.ssum(rho * rho) / r0.clamp_min(1e-300)).sqrt()


def _gn_small(b: _Batch, r, cols):
    """The Gauss-Newton step of each scene in the span of a few response columns (k, n, 3):
    min |r + sum_j c_j cols_j|^2 per scene (a relative 1e-9 ridge keeps it solvable).
    Returns c (S, len(cols))."""
    m = len(cols)
    G = torch.zeros(b.S, m, m, dtype=torch.float64)
    g = torch.zeros(b.S, m, dtype=torch.float64)
    for i in range(m):
        g[:, i] = -b.ssum(cols[i] * r)
        for j in range(i, m):
            G[:, i, j] = G[:, j, i] = b.ssum(cols[i] * cols[j])
    dg = torch.diagonal(G, dim1=1, dim2=2).clamp_min(1e-300)
    G = G + torch.diag_embed(1e-9 * dg)
    return torch.linalg.solve(G, g)


def _backtrack(b: _Batch, target, per, v0, lm, ls, dv, dm, ds, tries=4):
    """Each scene takes the largest of 1, 1/4, 1/16, ... (`tries` values) of its step that
    lowers its loss, or keeps its values. Returns (v0, log mu, log sigma, accepted per scene)."""
    lam = torch.ones(b.S, dtype=torch.float64)
    done = torch.zeros(b.S, dtype=torch.bool)
    nv, nm, ns = v0.clone(), lm.clone(), ls.clone()
    for _ in range(tries):
        tv = torch.where(b.spread(done), nv, v0 + b.spread(lam) * dv)
        tm = torch.where(done, nm, lm + lam * dm)
        ts = torch.where(done, ns, ls + lam * ds)
        pt = b.ssum((b.forward(tv, tm.exp(), ts.exp()) - target) ** 2)
        ok = (pt < per) & ~done
        nv = torch.where(b.spread(ok), tv, nv)
        nm, ns = torch.where(ok, tm, nm), torch.where(ok, ts, ns)
        done |= ok
        if bool(done.all()):
            break
        lam = torch.where(done, lam, lam / 4)
    return nv, nm, ns, done


def run(n_scenes=32, steps=80, iters=10, inner=6, seed=4, n_min=16384, n_max=131072, every=10, obs_every=10,
        recover_v0=True, fd_log=0.05, fd_v=0.01, max_log_step=1.0, v0_grid=0.05, log=sys.stderr):
    sc, mu_true, sig_true = make_scenes(n_scenes, seed, n_min, n_max, steps)
    obs = list(range(obs_every, steps + 1, obs_every))
    b = _Batch(sc, steps, every, obs)
    grid = _Grid(sc, v0_grid if recover_v0 else 0.0)
    v0_true = torch.as_tensor(np.concatenate([s.vel for s in sc]), dtype=torch.float64)
    t0 = time.perf_counter()
    target = b.forward(v0_true, mu_true, sig_true)
    t_target = time.perf_counter() - t0
    v0 = torch.zeros_like(v0_true) if recover_v0 else v0_true.clone()
    lm = torch.full((n_scenes,), math.log(MU0), dtype=torch.float64)
    ls = torch.full((n_scenes,), math.log(SIGMA0), dtype=torch.float64)
    n_obs_values = 3.0 * len(obs) * b.sizes.to(torch.float64)
    zero = torch.zeros(n_scenes, dtype=torch.float64)

    def errors():
        mu, sig = lm.exp().numpy(), ls.exp().numpy()
        em, es = np.abs(mu - mu_true) / mu_true, np.abs(sig - sig_true) / sig_true
        ev = (b.ssum((v0 - v0_true) ** 2) / b.ssum(v0_true ** 2)).sqrt().numpy()
        return em, es, ev

    x = b.forward(v0, lm.exp(), ls.exp(), record=recover_v0)
    r = x - target
    per = b.ssum(r * r)
    history = []
    t_start = time.perf_counter()
    for it in range(iters):
        res = zero
        kept = 0
        if recover_v0:  # initial velocities, coefficients held (the recorded rollout is at this point)
            dv, res = _cgls_v(b, grid, x, r, v0, lm.exp(), ls.exp(), inner, fd_v)
            v0, lm, ls, ok = _backtrack(b, target, per, v0, lm, ls, dv, zero, zero)
            kept += int((~ok).sum())
            x = b.forward(v0, lm.exp(), ls.exp())
            r = x - target
            per = b.ssum(r * r)
        # coefficients, with the initial velocities free to move along v0 and along their last step
        cols = [(b.forward(v0, (lm + fd_log).exp(), ls.exp()) - x) / fd_log,
                (b.forward(v0, lm.exp(), (ls + fd_log).exp()) - x) / fd_log]
        dirs = []
        if recover_v0:
            dirs = [v0, dv]
            cols += [_jv(b, x, v0, lm.exp(), ls.exp(), d, fd_v) for d in dirs]
        c = _gn_small(b, r, cols)
        del cols
        dm, ds = c[:, 0].clamp(-max_log_step, max_log_step), c[:, 1].clamp(-max_log_step, max_log_step)
        dvc = torch.zeros_like(v0)
        for j, d in enumerate(dirs):
            dvc = dvc + b.spread(c[:, 2 + j]) * d
        v0, lm, ls, ok = _backtrack(b, target, per, v0, lm, ls, dvc, dm, ds)
        kept += int((~ok).sum())
        x = b.forward(v0, lm.exp(), ls.exp(), record=recover_v0)
        r = x - target
        per = b.ssum(r * r)
        em, es, ev = errors()
        loss = float((per / n_obs_values).sum())
        history.append(dict(iter=it, loss=loss, kept=kept, lsq_residual_median=float(res.median()),
                            mu_err_median=float(np.median(em)), mu_err_max=float(em.max()),
                            sigma_err_median=float(np.median(es)), sigma_err_max=float(es.max()),
                            v0_err_median=float(np.median(ev)), elapsed_s=time.perf_counter() - t_start))
        print(f"iter {it:2d} loss {loss:.4e}  viscosity median {np.median(em):.2e} max {em.max():.2e}  "
              f"surface tension median {np.median(es):.2e} max {es.max():.2e}  v0 median {np.median(ev):.3f}  "
              f"kept {kept}  ({time.perf_counter() - t_start:.0f} s)", file=log, flush=True)
    t_solve = time.perf_counter() - t_start
    em, es, ev = errors()
    # the initial velocity's response: the velocity after the first step, and the velocity at
    # the first observation (positions after steps k and k + 1), at the hidden coefficients
    k1 = obs[0]
    xr = [b.forward(v, mu_true, sig_true, obs=(1, k1, k1 + 1)) for v in (v0, v0_true, torch.zeros_like(v0))]
    resp = (b.ssum((xr[0][0] - xr[1][0]) ** 2) / b.ssum((xr[1][0] - xr[2][0]) ** 2)).sqrt().numpy()
    vk = [x[2] - x[1] for x in xr]
    resp_k = (b.ssum((vk[0] - vk[1]) ** 2) / b.ssum((vk[1] - vk[2]) ** 2)).sqrt().numpy()
    # the fit itself: rms distance to the observed positions against the rms displacement of the
    # observed positions from the start
    x0 = torch.as_tensor(np.concatenate([s_.pos for s_ in sc]), dtype=torch.float64)
    rms_fit = (per / (3.0 * len(obs) * b.sizes.to(torch.float64))).sqrt().numpy()
    rms_disp = (b.ssum((target - x0[None]) ** 2) / (3.0 * len(obs) * b.sizes.to(torch.float64))).sqrt().numpy()
    mu, sig = lm.exp().numpy(), ls.exp().numpy()
    return {
        "method": "Gauss-Newton by blocks with per-scene backtracking: " + (
            f"initial velocities by {inner} CGLS iterations, then a 4-column solve in (log viscosity, "
            "log surface tension, v0 scale, last v0 step)" if recover_v0
            else "hidden v0 given; a 2x2 solve in (log viscosity, log surface tension)"),
        "inner": inner if recover_v0 else 0,
        "scenes": n_scenes, "particles": b.n, "steps": steps, "observed_steps": obs, "iterations": iters,
        "fd_log": fd_log, "fd_v": fd_v, "max_log_step": max_log_step, "checkpoint_every": every,
        "v0_grid": v0_grid if recover_v0 else None, "v0_unknowns": (3 * grid.M if grid.h > 0 else 3 * b.n) if recover_v0 else 0,
        "sizes": f"log-uniform {n_min}..{n_max}, seed {seed}", "start": {"viscosity": MU0, "surface_tension": SIGMA0},
        "target_rollout_s": t_target, "solve_s": t_solve,
        "forward_rollouts": b.fwd[0], "forward_s_each": b.fwd[1] / max(b.fwd[0], 1),
        "adjoint_passes": b.bwd[0], "adjoint_s_each": b.bwd[1] / max(b.bwd[0], 1),
        "final_loss": history[-1]["loss"] if history else None,
        "viscosity": {"true": mu_true.tolist(), "recovered": mu.tolist(),
                      "rel_err_median": float(np.median(em)), "rel_err_max": float(em.max())},
        "surface_tension": {"true": sig_true.tolist(), "recovered": sig.tolist(),
                            "rel_err_median": float(np.median(es)), "rel_err_max": float(es.max())},
        "initial_velocity": {"recovered": bool(recover_v0), "rel_err_per_scene": ev.tolist(),
                             "rel_err_median": float(np.median(ev)),
                             "rel_err_all": float(torch.linalg.norm(v0 - v0_true) / torch.linalg.norm(v0_true)),
                             "first_step_response_err_per_scene": resp.tolist(),
                             "first_step_response_err_median": float(np.median(resp)),
                             "first_step_response_err_max": float(resp.max()),
                             "first_observation_velocity_err_per_scene": resp_k.tolist(),
                             "first_observation_velocity_err_median": float(np.median(resp_k)),
                             "first_observation_step": k1},
        "trajectory_fit": {"rms_distance_per_scene": rms_fit.tolist(), "rms_displacement_per_scene": rms_disp.tolist(),
                           "rms_distance_median": float(np.median(rms_fit)),
                           "rms_distance_over_displacement_median": float(np.median(rms_fit / rms_disp))},
        "history": history,
    }


def scan_sigma(n_scenes=32, steps=80, seed=4, n_min=16384, n_max=131072, obs_every=10, points=13, log=sys.stderr):
    """The trajectory loss of every scene against its surface-tension coefficient, at the hidden
    viscosity and initial velocity: `points` factors from 1/4 to 4 of the hidden value (log
    spaced). A scene whose loss hardly rises away from its hidden value cannot have its surface
    tension recovered from these observations. For scale, also the loss of the hidden
    coefficients with the initial velocity 0.1 % and 1 % too large."""
    sc, mu_true, sig_true = make_scenes(n_scenes, seed, n_min, n_max, steps)
    roll = DifferentiableRollout(sc, steps, checkpoint_every=steps)
    obs = list(range(obs_every, steps + 1, obs_every))
    v0 = torch.as_tensor(np.concatenate([s.vel for s in sc]), dtype=torch.float32)
    offs = np.cumsum([0] + [s.n for s in sc])
    mu = torch.as_tensor(mu_true)
    with torch.no_grad():
        target = roll.trajectory(v0, mu, torch.as_tensor(sig_true), obs)
        factors = np.exp(np.linspace(math.log(0.25), math.log(4.0), points))
        table = []
        for f in factors:
            x = roll.trajectory(v0, mu, torch.as_tensor(sig_true * f), obs)
            table.append(_per_scene_loss(x, target, offs).numpy())
            print(f"scan factor {f:.3f}: median loss {np.median(table[-1]):.3e}", file=log, flush=True)
        # the same loss for small errors in the initial velocity's amplitude, for comparison
        amp = [1e-3, 1e-2]
        amp_loss = [_per_scene_loss(roll.trajectory(v0 * (1 + e), mu, torch.as_tensor(sig_true), obs), target,
                                    offs).numpy().tolist() for e in amp]
    table = np.array(table)  # (points, scenes)
    # the loss at a factor-2 error relative to the positions' squared scale (mean |x|^2 per scene)
    scale = np.array([float((target[:, offs[q]:offs[q + 1]] ** 2).mean()) for q in range(n_scenes)])
    return {"factors": factors.tolist(), "loss": table.tolist(), "position_scale": scale.tolist(),
            "v0_amplitude_errors": amp, "v0_amplitude_loss": amp_loss,
            "sigma_true": sig_true.tolist(), "viscosity_true": mu_true.tolist(), "observed_steps": obs}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--scenes", type=int, default=32)
    ap.add_argument("--steps", type=int, default=80)
    ap.add_argument("--iters", type=int, default=10, help="Gauss-Newton iterations")
    ap.add_argument("--inner", type=int, default=6, help="CGLS iterations per Gauss-Newton iteration")
    ap.add_argument("--seed", type=int, default=4)
    ap.add_argument("--n-min", type=int, default=16384)
    ap.add_argument("--n-max", type=int, default=131072)
    ap.add_argument("--obs-every", type=int, default=10)
    ap.add_argument("--fd-log", type=float, default=0.05)
    ap.add_argument("--fd-v", type=float, default=0.01)
    ap.add_argument("--v0-grid", type=float, default=0.05,
                    help="spacing (m) of the trilinear lattice the initial velocities are sought on (0: per particle)")
    ap.add_argument("--known-v0", action="store_true", help="keep the hidden initial velocities (do not recover them)")
    ap.add_argument("--scan", type=int, default=0, help="instead: loss scan over this many surface-tension factors")
    ap.add_argument("--out", default=None)
    a = ap.parse_args()
    if a.scan:
        res = scan_sigma(a.scenes, a.steps, a.seed, a.n_min, a.n_max, a.obs_every, a.scan)
        print(json.dumps({"scan_points": a.scan, "scenes": a.scenes}), flush=True)
    else:
        res = run(a.scenes, a.steps, a.iters, a.inner, a.seed, a.n_min, a.n_max, obs_every=a.obs_every,
                  recover_v0=not a.known_v0, fd_log=a.fd_log, fd_v=a.fd_v, v0_grid=a.v0_grid)
        print(json.dumps({k: v for k, v in res.items() if k != "history" and not isinstance(v, dict)}), flush=True)
    if a.out:
        with open(a.out, "w") as f:
            json.dump(res, f, indent=1)
    return 0


if __name__ == "__main__":
    sys.exit(main())
