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
from tests.adjoint_ref import VARIANTS, Replay, droplet, replay_for

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


@pytest.mark.parametrize("variant", VARIANTS)
def test_adjoint_matches_fp64_reference(variant):
    """Cubic spline, poly6/spiky, and two fluid radii (the reverse kernels' unequal-radius path)."""
    sc = droplet(variant, STEPS)
    g = _load(sc)
    counts = g.adj_rollout(STEPS, 4)
    xT = g.get("FLUID_POS")
    say(f"[adjoint] {variant}: loop counts per step (l_df, l_cg, l_cd): {counts[:, 0].tolist()}")
    rng = np.random.default_rng(0)
    w = rng.normal(size=(sc.n, 3))
    gv0, gmu, gsig = g.adj_backward(w.astype(np.float32))
    assert np.array_equal(g.get("FLUID_POS"), xT)  # left at the final state

    rep = replay_for(sc, g.get("FLUID_MASS"), g.get("BND_POS"), g.get("BND_RADIUS")[0])
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


def test_batched_adjoint_equals_scenes_alone():
    """Three droplets of different sizes and parameters in one context: per-scene gradients
    (initial velocities, viscosity, surface tension) identical to each droplet alone."""
    sc = [scenes.diff_droplet(n=n, radius=0.05 + 0.01 * k, seed=10 + k, steps=STEPS, h=0.04,
                              viscosity=0.01 + 0.01 * k, surface_tension=0.1 + 0.1 * k)
          for k, n in enumerate((180, 333, 260))]
    walls = [scenes.sample_walls(HipSim, s) for s in sc]
    bat = HipSim(scenes.make_config(sc[0]))
    scenes.load_batch(bat, sc, walls)
    bat.set("FLUID_VEL", np.concatenate([s.extra["v0"] for s in sc]))
    counts = bat.adj_rollout(STEPS, 4)
    rng = np.random.default_rng(2)
    w = [rng.normal(size=(s.n, 3)).astype(np.float32) for s in sc]
    gv, gmu, gsig = bat.adj_backward(np.concatenate(w))
    off = 0
    for q, s in enumerate(sc):
        g = HipSim(scenes.make_config(s))
        scenes.load_plain(g, s, walls[q])
        g.set("FLUID_VEL", s.extra["v0"])
        c1 = g.adj_rollout(STEPS, 4)
        assert np.array_equal(c1[:, 0, :], counts[:, q, :])
        gv1, gmu1, gsig1 = g.adj_backward(w[q])
        assert np.array_equal(gv[off:off + s.n], gv1), q
        assert gmu[q] == gmu1[0] and gsig[q] == gsig1[0], q
        off += s.n
    say(f"[adjoint] batch of {len(sc)}: gradients identical to each scene alone; mu grads {gmu.tolist()}")


def test_autograd_function_drives_the_adjoint():
    from mps_cuda_amd.autograd import DifferentiableRollout
    sc = scenes.diff_droplet(n=300, radius=0.08, seed=3, steps=STEPS, h=0.04)
    roll = DifferentiableRollout([sc], STEPS, checkpoint_every=2)
    v0 = torch.as_tensor(sc.extra["v0"], dtype=torch.float32).clone().requires_grad_(True)
    mu = torch.tensor([0.02], dtype=torch.float64, requires_grad=True)
    sig = torch.tensor([0.3], dtype=torch.float64, requires_grad=True)
    target = roll(v0.detach() * 0.5, mu.detach() * 2, sig.detach())
    x = roll(v0, mu, sig)
    loss = ((x - target) ** 2).mean()
    loss.backward()
    # same gradients through the C ABI directly
    g = _load(sc)
    g.adj_rollout(STEPS, 2)
    gx = (2.0 / x.numel()) * (g.get("FLUID_POS") - target.numpy())
    gv0, gmu, gsig = g.adj_backward(gx.astype(np.float32))
    assert _rel(v0.grad.numpy(), gv0) < 1e-6 and _rel(mu.grad.numpy(), gmu) < 1e-6 and _rel(sig.grad.numpy(), gsig) < 1e-6
    assert np.isfinite(loss.item()) and loss.item() > 0


def test_trajectory_cotangents_match_fp64_replay():
    """A loss over the whole trajectory, L = sum_k w_k . x(t_k) at steps 2, 4 (twice: a repeated
    step adds) and 6: the observed states and the gradients of the trajectory path
    (mps_adj_rollout_traj / mps_adj_backward_traj) against the FP64 replay with torch autograd;
    the single final observation equals mps_adj_backward bit for bit."""
    sc = scenes.diff_droplet(n=300, radius=0.08, seed=6, steps=STEPS, h=0.04)
    obs = [2, 4, 4, STEPS]
    g = _load(sc)
    say(f"[adjoint-traj] N={sc.n} steps={STEPS} obs={obs}")
    xs, _, counts = g.adj_rollout_traj(STEPS, obs, 4)
    rng = np.random.default_rng(7)
    w = rng.normal(size=(len(obs), sc.n, 3))
    gv0, gmu, gsig = g.adj_backward_traj(obs, w.astype(np.float32))
    rep = Replay(sc.pos, sc.radius[0], g.get("FLUID_MASS")[0], g.get("BND_POS"), g.get("BND_RADIUS")[0], sc.physics,
                 sc.material)
    v0 = torch.as_tensor(sc.extra["v0"]).clone().requires_grad_(True)
    mu = torch.tensor(sc.material["viscosity"], dtype=torch.float64, requires_grad=True)
    sig = torch.tensor(sc.material["surface_tension"], dtype=torch.float64, requires_grad=True)
    xr = rep.rollout(v0, mu, sig, counts[:, 0, :], observe=obs)
    assert _rel(xr.detach().numpy(), xs) < 1e-5
    L = (torch.as_tensor(w) * xr).sum()
    rv, rm, rs = torch.autograd.grad(L, (v0, mu, sig))
    e_v, e_mu, e_sig = _rel(gv0, rv.numpy()), _rel([gmu[0]], [rm.item()]), _rel([gsig[0]], [rs.item()])
    say(f"[adjoint-traj] rel err vs FP64: v0 {e_v:.2e} mu {e_mu:.2e} sigma {e_sig:.2e}")
    assert e_v < 1e-3 and e_mu < 1e-3 and e_sig < 1e-3
    # the final state alone through both entry points
    g2 = _load(sc)
    g2.adj_rollout(STEPS, 4)
    wf = w[-1].astype(np.float32)
    a = g2.adj_backward(wf)
    b = g2.adj_backward_traj([STEPS], wf[None])
    assert all(np.array_equal(p, q) for p, q in zip(a, b))


def test_adjoint_across_z_sorts():
    """The Z-sort every 2 frames (with faithful_quirks = 0, so wall volumes follow their walls)
    relabels the fluid during the rollout: observations after steps 2, 4 and 6 each come in the
    reference order of their step. The trajectory adjoint against the FP64 replay, whose
    particles keep their first labels, matched to the observed ones by position. The forward CG
    runs to 1e-12 of its initial r.r: the viscosity derivative is only as exact as that solve
    (at the default 1e-8 it is off by 1.2e-3 here)."""
    sc = droplet("cubic", STEPS)
    sc.config.update(zsort_interval=2, faithful_quirks=0, cg_tol=-1.0e-12, cg_max_iter=400)
    g = _load(sc)
    obs = [2, 4, 6]
    xs, _, counts = g.adj_rollout_traj(STEPS, obs, 3)
    w = np.random.default_rng(21).normal(size=xs.shape)
    gv0, gmu, gsig = g.adj_backward_traj(obs, w.astype(np.float32))
    rep = replay_for(sc, g.get("FLUID_MASS"), g.get("BND_POS"), g.get("BND_RADIUS")[0])
    v0 = torch.as_tensor(sc.extra["v0"]).clone().requires_grad_(True)
    mu = torch.tensor(sc.material["viscosity"], dtype=torch.float64, requires_grad=True)
    sig = torch.tensor(sc.material["surface_tension"], dtype=torch.float64, requires_grad=True)
    xr = rep.rollout(v0, mu, sig, counts[:, 0, :], observe=obs)
    L = torch.zeros((), dtype=torch.float64)
    relabelled = 0
    for k in range(len(obs)):
        nn = torch.cdist(xr[k].detach(), torch.as_tensor(xs[k], dtype=torch.float64)).argmin(dim=1).numpy()
        assert len(set(nn.tolist())) == sc.n
        assert _rel(xr[k].detach().numpy(), xs[k][nn]) < 1e-5
        relabelled += int((nn != np.arange(sc.n)).any())
        L = L + (torch.as_tensor(w[k][nn]) * xr[k]).sum()
    assert relabelled == len(obs)
    rv, rm, rs = torch.autograd.grad(L, (v0, mu, sig))
    e_v, e_mu, e_sig = _rel(gv0, rv.numpy()), _rel([gmu[0]], [rm.item()]), _rel([gsig[0]], [rs.item()])
    say(f"[adjoint] across Z-sorts (frames 1, 3, 5), observed at {obs}: rel err vs FP64 v0 {e_v:.2e} mu {e_mu:.2e} "
        f"sigma {e_sig:.2e}")
    assert e_v < 1e-3 and e_mu < 1e-3 and e_sig < 1e-3


def test_stale_rollout_is_refused():
    """A backward after the particles, walls or parameters were replaced raises before any
    copy (the rollout's checkpoints no longer describe the context)."""
    from mps_cuda_amd._capi import MPSError
    sc = scenes.diff_droplet(n=300, radius=0.08, seed=3, steps=STEPS, h=0.04)
    small = scenes.diff_droplet(n=120, radius=0.06, seed=4, steps=STEPS, h=0.04)
    w = np.ones((sc.n, 3), np.float32)
    for change in ("smaller fluid", "material", "walls"):
        g = _load(sc)
        g.adj_rollout(STEPS, 3)
        say(f"[adjoint-stale] {change}")
        if change == "smaller fluid":
            scenes.load_into(g, small)
        elif change == "material":
            g.scene_set_material(0, **dict(sc.material, viscosity=0.05))
        else:
            g.upload_boundary(g.get("BND_POS"), g.get("BND_RADIUS"))
        with pytest.raises(MPSError) as ei:
            g.adj_backward(w[:g.get("FLUID_RADIUS").shape[0]])
        assert ei.value.code == 4


def test_autograd_refuses_a_replaced_rollout():
    from mps_cuda_amd.autograd import DifferentiableRollout
    sc = scenes.diff_droplet(n=300, radius=0.08, seed=3, steps=STEPS, h=0.04)
    roll = DifferentiableRollout([sc], STEPS, checkpoint_every=3)
    v0 = torch.as_tensor(sc.extra["v0"], dtype=torch.float32).clone().requires_grad_(True)
    mu = torch.tensor([0.02], dtype=torch.float64, requires_grad=True)
    sig = torch.tensor([0.3], dtype=torch.float64, requires_grad=True)
    x1 = roll(v0, mu, sig)
    x2 = roll(v0, mu * 1.5, sig)
    with pytest.raises(RuntimeError, match="replaced"):
        x1.sum().backward()
    x2.sum().backward()      # the latest rollout still differentiates
    assert mu.grad is not None and np.isfinite(mu.grad.numpy()).all()
    assert v0.grad.dtype == v0.dtype and mu.grad.dtype == mu.dtype


def test_autograd_trajectory_matches_the_c_abi():
    from mps_cuda_amd.autograd import DifferentiableRollout
    sc = scenes.diff_droplet(n=300, radius=0.08, seed=8, steps=STEPS, h=0.04)
    obs = [3, STEPS]
    roll = DifferentiableRollout([sc], STEPS, checkpoint_every=2)
    v0 = torch.as_tensor(sc.extra["v0"], dtype=torch.float32).clone().requires_grad_(True)
    mu = torch.tensor([sc.material["viscosity"]], dtype=torch.float64, requires_grad=True)
    sig = torch.tensor([sc.material["surface_tension"]], dtype=torch.float64, requires_grad=True)
    xs = roll.trajectory(v0, mu, sig, obs)
    assert tuple(xs.shape) == (len(obs), sc.n, 3)
    w = torch.as_tensor(np.random.default_rng(9).normal(size=xs.shape), dtype=torch.float32)
    (w * xs).sum().backward()
    g = _load(sc)
    g.adj_rollout_traj(STEPS, obs, 2)
    gv0, gmu, gsig = g.adj_backward_traj(obs, w.numpy())
    assert _rel(v0.grad.numpy(), gv0) < 1e-6 and _rel(mu.grad.numpy(), gmu) < 1e-6 and _rel(sig.grad.numpy(), gsig) < 1e-6


@pytest.mark.parametrize("kernel", ["cubic", "poly6"])
def test_adjoint_matches_fp64_finite_differences(kernel):
    """A 4096-particle droplet over 30 steps at (viscosity 0.025, surface tension 0.1), with the
    cubic spline or the poly6/spiky kernels: the GPU adjoint's gradients of L = sum w . x_T (fixed
    random w, so both sides use the same cotangent) against central differences (eps 1e-3,
    relative for the coefficients, along random unit vectors per particle for the initial
    velocities) of the FP64 oracle's own forward (its CG run to 1e-16 of the initial r.r), within
    1e-3."""
    from mps_cuda_amd import _capi
    from oracle.binding import OracleSim, set_threads
    steps = 30
    sc = scenes.diff_droplet(n=4096, radius=0.1, seed=5, steps=steps, h=0.02)
    if kernel == "poly6":
        sc.config["kernel"] = _capi.KERNEL_POLY6_SPIKY
    g = _load(sc)
    mu0, sig0 = 0.025, 0.1
    w = np.random.default_rng(8).normal(size=(sc.n, 3))

    def o_loss(mu, sig, v):
        cfg = scenes.make_config(sc)
        cfg.cg_tol, cfg.cg_max_iter = -1.0e-16, 300
        o = OracleSim(cfg, 64)
        scenes.load_into(o, sc)
        o.set_fluid_material(**dict(sc.material, viscosity=mu, surface_tension=sig))
        o.set("FLUID_VEL", v)
        for _ in range(steps):
            o.step()
        return float(np.sum(w * o.get("FLUID_POS")))

    set_threads(0)
    g.set_fluid_material(**dict(sc.material, viscosity=mu0, surface_tension=sig0))
    g.adj_rollout(steps, 10)
    gv0, gmu, gsig = g.adj_backward(w.astype(np.float32))
    v0 = sc.extra["v0"]
    dv = np.random.default_rng(9).normal(size=v0.shape)
    dv /= np.linalg.norm(dv, axis=1, keepdims=True)
    e = 1.0e-3
    fd_mu = (o_loss(mu0 * (1 + e), sig0, v0) - o_loss(mu0 * (1 - e), sig0, v0)) / (2 * e * mu0)
    fd_sig = (o_loss(mu0, sig0 * (1 + e), v0) - o_loss(mu0, sig0 * (1 - e), v0)) / (2 * e * sig0)
    fd_v = (o_loss(mu0, sig0, v0 + e * dv) - o_loss(mu0, sig0, v0 - e * dv)) / (2 * e)
    ad_v = float(np.sum(gv0.astype(np.float64) * dv))
    errs = (abs(gmu[0] - fd_mu) / abs(fd_mu), abs(gsig[0] - fd_sig) / abs(fd_sig), abs(ad_v - fd_v) / abs(fd_v))
    say(f"[adjoint] {kernel}, 4096 x {steps} steps vs FP64 FD: mu {gmu[0]:.6e}/{fd_mu:.6e}  sigma {gsig[0]:.6e}/{fd_sig:.6e}  "
        f"v0 {ad_v:.6e}/{fd_v:.6e}  rel errs {errs}")
    assert max(errs) < 1e-3, errs


def test_adjoint_with_graph_replay():
    sc = scenes.diff_droplet(n=300, radius=0.08, seed=3, steps=STEPS, h=0.04)
    w = np.random.default_rng(4).normal(size=(sc.n, 3)).astype(np.float32)
    out = []
    for graphs in (0, 1):
        cfg = scenes.make_config(sc)
        cfg.use_graphs = graphs
        g = HipSim(cfg)
        scenes.load_into(g, sc)
        g.set("FLUID_VEL", sc.extra["v0"])
        g.adj_rollout(STEPS, 3)
        out.append(g.adj_backward(w))
    for a, b in zip(out[0], out[1]):
        assert np.array_equal(a, b)


# Phases isolated by configuration (the adjoint is exposed per rollout): each case switches on one
# more part of the step. "advection": gravity, integration and wall response only (no viscosity,
# surface tension or XSPH, pressure loops stopped after their first evaluation); "pressure": the
# DF/CD solves, on the droplet compressed to 0.78 of its size (1.27 rest density: the first step's
# constant-density solve pushes it apart); "viscosity" / "surface_tension": the implicit CG with
# only that term, on the same compressed droplet, run to 1e-12 of the initial r.r (the implicit
# derivative is only as exact as the forward solve).
PHASES = {
    "advection": (dict(viscosity=0.0, surface_tension=0.0, xsph=0.0),
                  dict(df_tol=1.0e30, cd_tol=1.0e30, df_min_iter=0, cd_min_iter=0), 1.0),
    "pressure": (dict(viscosity=0.0, surface_tension=0.0, xsph=0.0), {}, 0.78),
    "viscosity": (dict(viscosity=0.02, surface_tension=0.0), dict(cg_tol=-1.0e-12, cg_max_iter=400), 0.78),
    "surface_tension": (dict(viscosity=0.0, surface_tension=0.3, xsph=0.0), dict(cg_tol=-1.0e-12, cg_max_iter=400),
                        0.78),
}


@pytest.mark.parametrize("variant", VARIANTS)
@pytest.mark.parametrize("phase", list(PHASES))
def test_dot_product_per_phase(phase, variant):
    """Dot-product test <J v, w> = <v, J^T w> for the map (v0, viscosity, surface tension) ->
    (final positions, final velocities) of a 300-particle droplet over 6 steps, one phase of the
    step switched on at a time. J^T w comes from the GPU adjoint with cotangents w = (w_x, w_v);
    J v from the forward derivative (torch.autograd.functional.jvp) of the FP64 replay of the
    same frozen map (tests/adjoint_ref.py), along a dense random direction in v0 and along the
    coefficient the phase has. For each kernel set and for two fluid radii (tests/adjoint_ref.py
    VARIANTS). The v0 error is measured against the Cauchy-Schwarz scale |v| |J^T w| (a random
    direction can make <J v, w> itself nearly cancel), bar 5e-6 (about 1.5e-4 of a typical
    <J v, w> for these 900 components); the coefficients' against their derivative, bar 1e-3."""
    sc = droplet(variant, STEPS)
    mat, conf, squeeze = PHASES[phase]
    sc.material.update(mat)
    sc.config.update(conf)
    c = sc.pos.mean(axis=0)
    sc.pos = c + squeeze * (sc.pos - c)
    rng = np.random.default_rng(17)
    wx, wv, dv = rng.normal(size=(sc.n, 3)), rng.normal(size=(sc.n, 3)), rng.normal(size=(sc.n, 3))

    if phase == "pressure":  # the constant-density solve really applies forces in the first step
        p = _load(sc)
        st = p.step()
        assert st.err_cd[0] > 10 * sc.config.get("cd_tol", 1e-3) and st.l_cd >= 2, (list(st.err_cd)[:3], st.l_cd)
    g = _load(sc)
    counts = g.adj_rollout(STEPS, 3)
    gv0, gmu, gsig = g.adj_backward(wx.astype(np.float32), wv.astype(np.float32))

    rep = replay_for(sc, g.get("FLUID_MASS"), g.get("BND_POS"), g.get("BND_RADIUS")[0])
    tw = (torch.as_tensor(wx), torch.as_tensor(wv))

    def f(v, m, s):
        x, vT = rep.rollout(v, m, s, counts[:, 0, :])
        return (tw[0] * x).sum() + (tw[1] * vT).sum()

    args = (torch.as_tensor(sc.extra["v0"]), torch.tensor(sc.material["viscosity"], dtype=torch.float64),
            torch.tensor(sc.material["surface_tension"], dtype=torch.float64))
    zero = torch.zeros((), dtype=torch.float64)
    _, jv = torch.autograd.functional.jvp(f, args, (torch.as_tensor(dv), zero, zero))
    v_jtw = float(np.sum(gv0.astype(np.float64) * dv))
    scale = float(np.linalg.norm(dv) * np.linalg.norm(gv0.astype(np.float64)))
    errs = {"v0": abs(v_jtw - float(jv)) / scale, "v0_of_value": abs(v_jtw - float(jv)) / abs(float(jv))}
    msg = f"[dot-product] {phase}: <v, J^T w> {v_jtw:.9e}  <J v, w> {float(jv):.9e}"
    one = torch.ones((), dtype=torch.float64)
    for name, ad, tangent in (("viscosity", gmu[0], (zero, one, zero)), ("surface_tension", gsig[0], (zero, zero, one))):
        if sc.material[name] > 0:
            _, jc = torch.autograd.functional.jvp(f, args, (torch.zeros_like(args[0]),) + tangent[1:])
            errs[name] = abs(ad - float(jc)) / abs(float(jc))
            msg += f"  d/d{name} {ad:.6e}/{float(jc):.6e}"
    say(f"{msg} ({variant})  loops (l_df, l_cg, l_cd) {sorted(set(map(tuple, counts[:, 0, :].tolist())))}  rel errs {errs}")
    assert errs["v0"] < 5e-6, errs
    assert all(e < 1e-3 for k, e in errs.items() if not k.startswith("v0")), errs
