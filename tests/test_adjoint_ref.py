"""The FP64 reference of the adjoint (tests/adjoint_ref.py) pinned on the CPU: its replay of a
rollout reproduces the FP64 oracle's trajectory (given the oracle's loop counts), and its
autograd gradients match central finite differences of the same frozen map. Three variants of
the droplet: the cubic spline with equal radii, the poly6/spiky kernel set, and two fluid radii
(r and 0.8 r, so the pair kernels differ between the two ends)."""
import numpy as np
import pytest
import torch

from mps_cuda_amd import scenes
from oracle.binding import OracleSim
from tests.adjoint_ref import VARIANTS, droplet, replay_for

STEPS = 4


def _scene(variant="cubic"):
    return droplet(variant, STEPS)


def _oracle_run(sc, v0, mu, sigma):
    cfg = scenes.make_config(sc)
    cfg.cg_tol = -1.0e-26  # converge the FP64 solve; the replay solves it exactly
    o = OracleSim(cfg, 64)
    m = dict(sc.material, viscosity=mu, surface_tension=sigma)
    p = sc.physics
    o.set_physics(p["dt"], p["gravity"], p["box_min"], p["box_max"])
    o.set_fluid_material(**m)
    b = sc.boundary_mesh
    o.sample_boundary_mesh(b["verts"], b["tris"], b["radius"], b.get("density", 1.0), b.get("center"), b.get("size"))
    o.upload_fluid(sc.pos, sc.radius, v0)
    counts = []
    for _ in range(STEPS):
        st = o.step()
        counts.append((st.l_df, st.l_cg, st.l_cd))
    return o, np.array(counts)


def _replay(sc, o):
    return replay_for(sc, o.get("FLUID_MASS"), o.get("BND_POS"), o.get("BND_RADIUS")[0])


@pytest.mark.parametrize("variant", VARIANTS)
def test_replay_matches_oracle_trajectory(variant):
    sc = _scene(variant)
    mu, sig = 0.02, 0.3
    o0, _ = _oracle_run(sc, None, mu, sig)  # for the wall particles
    o, counts = _oracle_run(sc, sc.extra["v0"], mu, sig)
    assert counts[:, 0].max() >= 2 or counts[:, 2].max() >= 3  # some pressure iterations applied
    rep = _replay(sc, o0)
    x, v = rep.rollout(torch.as_tensor(sc.extra["v0"]), torch.tensor(mu, dtype=torch.float64),
                       torch.tensor(sig, dtype=torch.float64), counts)
    xo, vo = o.get("FLUID_POS"), o.get("FLUID_VEL")
    ex = np.linalg.norm(x.numpy() - xo) / np.linalg.norm(xo)
    ev = np.linalg.norm(v.numpy() - vo) / np.linalg.norm(vo)
    assert ex < 1e-9 and ev < 1e-6, (ex, ev)


@pytest.mark.parametrize("variant", VARIANTS)
def test_replay_gradients_match_finite_differences(variant):
    sc = _scene(variant)
    o0, counts = _oracle_run(sc, sc.extra["v0"], 0.02, 0.3)
    rep = _replay(sc, o0)
    rng = np.random.default_rng(0)
    w = torch.as_tensor(rng.normal(size=(sc.n, 3)))
    dv = torch.as_tensor(rng.normal(size=(sc.n, 3)))

    def loss(v0, mu, sig):
        x, _ = rep.rollout(v0, mu, sig, counts)
        return (w * x).sum()

    v0 = torch.as_tensor(sc.extra["v0"]).clone().requires_grad_(True)
    mu = torch.tensor(0.02, dtype=torch.float64, requires_grad=True)
    sig = torch.tensor(0.3, dtype=torch.float64, requires_grad=True)
    L = loss(v0, mu, sig)
    gv, gm, gs = torch.autograd.grad(L, (v0, mu, sig))
    with torch.no_grad():
        e = 1e-5
        fd_mu = (loss(v0, mu + e, sig) - loss(v0, mu - e, sig)) / (2 * e)
        fd_sig = (loss(v0, mu, sig + e) - loss(v0, mu, sig - e)) / (2 * e)
        fd_v = (loss(v0 + e * dv, mu, sig) - loss(v0 - e * dv, mu, sig)) / (2 * e)
    assert abs(fd_mu - gm) <= 1e-5 * abs(gm) + 1e-9, (float(fd_mu), float(gm))
    assert abs(fd_sig - gs) <= 1e-5 * abs(gs) + 1e-9, (float(fd_sig), float(gs))
    assert abs(fd_v - (gv * dv).sum()) <= 1e-5 * abs(fd_v), (float(fd_v), float((gv * dv).sum()))


def test_replay_without_volume_accumulation_matches_oracle():
    """faithful_quirks = 0 (wall volume 1/S every step, permuted with the walls by the Z-sort):
    the replay without accumulation follows the oracle with the Z-sort every 2 frames. The
    oracle relabels its fluid at each Z-sort, so positions are compared as point sets."""
    sc = _scene()
    sc.config.update(zsort_interval=2, faithful_quirks=0)
    mu, sig = 0.02, 0.3
    o, counts = _oracle_run(sc, sc.extra["v0"], mu, sig)
    rep = _replay(sc, o)
    assert not rep.accumulate
    x, _ = rep.rollout(torch.as_tensor(sc.extra["v0"]), torch.tensor(mu, dtype=torch.float64),
                       torch.tensor(sig, dtype=torch.float64), counts)
    xo = o.get("FLUID_POS")
    D = torch.cdist(x, torch.as_tensor(xo))
    nn = D.argmin(dim=1).numpy()
    assert len(set(nn.tolist())) == sc.n  # a one-to-one label map
    assert (nn != np.arange(sc.n)).any()  # the Z-sort did relabel
    ex = np.linalg.norm(x.numpy() - xo[nn]) / np.linalg.norm(xo)
    assert ex < 1e-9, ex
