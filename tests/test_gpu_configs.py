"""The batched and differentiable workloads at their full scene sizes, on the GPU.

- Batched forward: the 64 scenes of 8k-256k particles (4.49M, seed 2) stepped together with
  solver chunks replayed from HIP graphs, against the smallest, the median and the largest
  scene stepped alone: bit-identical positions, velocities, densities and loop counts.
- Differentiable droplet: the 65,536-particle sphere (radius 0.25 m, h = 0.02 m, checkpoints
  every 10 steps). Its gradients of L = sum w . x_T (fixed random w) with respect to the
  viscosity, the surface-tension coefficient and the initial velocities (along random unit
  vectors per particle) against central differences (eps 1e-3) of the FP64 oracle's forward
  (its CG run to 1e-16 of the initial r.r), within 1e-3. The oracle takes about 0.5-1 s per step
  of this droplet at that tolerance, so the check runs 20 steps (two checkpoint segments), not
  the workload's 100: six oracle forwards of 100 steps would take several minutes.
- Batched parameter recovery: the 32 scenes of 16k-128k particles (2.12M, seed 4) with their
  hidden coefficients; the batched adjoint (a trajectory loss observed at two steps) gives, for
  the smallest, the median and the largest scene, the gradients of that scene run alone, bit for
  bit.
"""
import sys

import numpy as np
import pytest

from mps_cuda_amd import HipSim, scenes

pytestmark = pytest.mark.gpu


def say(m):
    print(m, flush=True)
    sys.stdout.flush()


def _sample(sizes):
    order = np.argsort(sizes, kind="stable")
    return sorted({int(order[0]), int(order[len(order) // 2]), int(order[-1])})


def test_batch_of_64_with_graph_replay_matches_scenes_alone():
    sc = scenes.batch_scenes(n_scenes=64, seed=2)
    say(f"[cfg-batch] 64 scenes, {sum(s.n for s in sc)} particles")
    walls = [scenes.sample_walls(HipSim, s) for s in sc]
    cfg = scenes.make_config(sc[0])
    cfg.use_graphs = 1
    bat = HipSim(cfg)
    scenes.load_batch(bat, sc, walls)
    picks = _sample([s.n for s in sc])
    alone = {}
    for q in picks:
        g = HipSim(scenes.make_config(sc[q]))
        scenes.load_plain(g, sc[q], walls[q])
        alone[q] = g
    steps = 8
    for t in range(steps):
        bat.step()
        for q, g in alone.items():
            sa, sb = g.step(), bat.scene_stats(q)
            assert (sb.l_df, sb.l_cg, sb.l_cd) == (sa.l_df, sa.l_cg, sa.l_cd), (t, q)
    say(f"[cfg-batch] {steps} steps; checking scenes {picks} ({[sc[q].n for q in picks]} particles)")
    for name in ("FLUID_POS", "FLUID_VEL", "FLUID_DENSITY"):
        a = bat.get(name)
        for q, g in alone.items():
            i = bat.scene_info(q)
            assert np.array_equal(a[i["fluid_offset"]:i["fluid_offset"] + i["n_fluid"]], g.get(name)), (name, q)


@pytest.mark.timeout(900)
def test_droplet_64k_adjoint_matches_fp64_finite_differences():
    from oracle.binding import OracleSim, set_threads
    steps, every = 20, 10
    sc = scenes.diff_droplet(n=65536, radius=0.25, seed=5, steps=steps)
    mu0, sig0 = 0.025, 0.1
    cfg = scenes.make_config(sc)
    cfg.cg_tol, cfg.cg_max_iter = -1.0e-12, 400   # the implicit derivative is as exact as this solve
    g = HipSim(cfg)
    scenes.load_into(g, sc)
    g.set_fluid_material(**dict(sc.material, viscosity=mu0, surface_tension=sig0))
    g.set("FLUID_VEL", sc.extra["v0"])
    w = np.random.default_rng(8).normal(size=(sc.n, 3))
    counts = g.adj_rollout(steps, every)
    gv0, gmu, gsig = g.adj_backward(w.astype(np.float32))
    say(f"[cfg-droplet] GPU adjoint done: {sc.n} particles x {steps} steps, l_cg mean {counts[:, 0, 1].mean():.1f}")

    def o_loss(mu, sig, v):
        c = scenes.make_config(sc)
        c.cg_tol, c.cg_max_iter = -1.0e-16, 300
        o = OracleSim(c, 64)
        scenes.load_into(o, sc)
        o.set_fluid_material(**dict(sc.material, viscosity=mu, surface_tension=sig))
        o.set("FLUID_VEL", v)
        for _ in range(steps):
            o.step()
        val = float(np.sum(w * o.get("FLUID_POS")))
        o.close()
        return val

    say(f"[cfg-droplet] oracle threads {set_threads(0)}")
    v0 = sc.extra["v0"]
    dv = np.random.default_rng(9).normal(size=v0.shape)
    dv /= np.linalg.norm(dv, axis=1, keepdims=True)
    e = 1.0e-3
    fd_mu = (o_loss(mu0 * (1 + e), sig0, v0) - o_loss(mu0 * (1 - e), sig0, v0)) / (2 * e * mu0)
    say(f"[cfg-droplet] mu: adjoint {gmu[0]:.6e} FD {fd_mu:.6e}")
    fd_sig = (o_loss(mu0, sig0 * (1 + e), v0) - o_loss(mu0, sig0 * (1 - e), v0)) / (2 * e * sig0)
    say(f"[cfg-droplet] sigma: adjoint {gsig[0]:.6e} FD {fd_sig:.6e}")
    fd_v = (o_loss(mu0, sig0, v0 + e * dv) - o_loss(mu0, sig0, v0 - e * dv)) / (2 * e)
    ad_v = float(np.sum(gv0.astype(np.float64) * dv))
    say(f"[cfg-droplet] v0 direction: adjoint {ad_v:.6e} FD {fd_v:.6e}")
    errs = (abs(gmu[0] - fd_mu) / abs(fd_mu), abs(gsig[0] - fd_sig) / abs(fd_sig), abs(ad_v - fd_v) / abs(fd_v))
    say(f"[cfg-droplet] relative errors mu {errs[0]:.2e} sigma {errs[1]:.2e} v0 {errs[2]:.2e}")
    assert max(errs) < 1e-3, errs


def test_recovery_batch_of_32_gradients_equal_scenes_alone():
    from mps_cuda_amd.recover import make_scenes
    steps, every, obs = 6, 3, [3, 6]
    sc, mu_true, sig_true = make_scenes(32, 4, 16384, 131072, steps)
    say(f"[cfg-recovery] 32 scenes, {sum(s.n for s in sc)} particles, {steps} steps, observed at {obs}")
    walls = [scenes.sample_walls(HipSim, s) for s in sc]
    bat = HipSim(scenes.make_config(sc[0]))
    scenes.load_batch(bat, sc, walls)
    for q, s in enumerate(sc):
        bat.scene_set_material(q, **dict(s.material, viscosity=float(mu_true[q]), surface_tension=float(sig_true[q])))
    xs, _, counts = bat.adj_rollout_traj(steps, obs, every)
    rng = np.random.default_rng(3)
    w = rng.normal(size=xs.shape).astype(np.float32)
    gv, gmu, gsig = bat.adj_backward_traj(obs, w)
    picks = _sample([s.n for s in sc])
    for q in picks:
        s = sc[q]
        i = bat.scene_info(q)
        lo, hi = i["fluid_offset"], i["fluid_offset"] + i["n_fluid"]
        g = HipSim(scenes.make_config(s))
        scenes.load_plain(g, s, walls[q])
        g.set_fluid_material(**dict(s.material, viscosity=float(mu_true[q]), surface_tension=float(sig_true[q])))
        xq, _, cq = g.adj_rollout_traj(steps, obs, every)
        assert np.array_equal(cq[:, 0, :], counts[:, q, :]), q
        assert np.array_equal(xq, xs[:, lo:hi]), q
        gv1, gmu1, gsig1 = g.adj_backward_traj(obs, w[:, lo:hi])
        assert np.array_equal(gv[lo:hi], gv1), q
        assert gmu[q] == gmu1[0] and gsig[q] == gsig1[0], q
        say(f"[cfg-recovery] scene {q} ({s.n} particles): gradients identical; mu {gmu1[0]:.4e} sigma {gsig1[0]:.4e}")
