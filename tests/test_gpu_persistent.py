"""Launch variants of the viscosity CG against the chunked launches with fold kernels:
bit-identical velocities, positions, loop counts and r.r logs, for one scene and for a batch
whose scenes leave the loop at different iterations.
- The persistent solve (all iterations in one cooperative launch with grid-wide barriers,
  every workgroup folding the sums itself in the fold kernels' order), and its give-up path,
  forced from the host with a barrier wait of one poll: the solve then reruns with chunked
  launches, with the same bits.
- The chunked launches whose sums are folded by the last workgroup of each launch.
"""
import sys

import numpy as np
import pytest

from mps_cuda_amd import HipSim, scenes

pytestmark = pytest.mark.gpu


def say(m):
    print(m, flush=True)
    sys.stdout.flush()


def _sim(sc, **cfgkw):
    cfg = scenes.make_config(sc)
    for k, v in cfgkw.items():
        setattr(cfg, k, v)
    g = HipSim(cfg)
    scenes.load_into(g, sc)
    if "v0" in sc.extra:
        g.set("FLUID_VEL", sc.extra["v0"])
    return g


def _same_state(a, b, what):
    for name in ("FLUID_POS", "FLUID_VEL"):
        assert np.array_equal(a.get(name), b.get(name)), (what, name)


@pytest.mark.parametrize("mode", [1, 2])
@pytest.mark.parametrize("spin,expect", [(0, 1), (1, 2)])
def test_persistent_cg_single_scene(spin, expect, mode):
    """The 65,536-particle droplet (256 workgroups) for 4 steps of ~100 CG iterations each, with
    either barrier (1: fences, 2: device-coherent accesses)."""
    sc = scenes.diff_droplet(n=65536, radius=0.25, seed=2, steps=4)
    a = _sim(sc, persistent_solver=0, cg_fold_in_kernel=0)
    b = _sim(sc, persistent_solver=mode, persistent_spin=spin)
    say(f"[persistent] droplet N={sc.n} mode={mode} spin={spin}")
    for t in range(4):
        sa, sb = a.step(), b.step()
        assert sa.cg_persistent == 0 and sb.cg_persistent == expect, (t, sb.cg_persistent)
        assert (sa.l_df, sa.l_cg, sa.l_cd) == (sb.l_df, sb.l_cg, sb.l_cd), t
        n = min(sa.l_cg + 1, 128)
        assert list(sa.rr_cg[:n]) == list(sb.rr_cg[:n]), t
        _same_state(a, b, t)
    say(f"[persistent] spin={spin}: l_cg {sa.l_cg}, identical to the chunked solve")


@pytest.mark.parametrize("mode", [1, 2])
def test_persistent_cg_batch_scenes_converging_apart(mode):
    sc = scenes.batch_scenes(n_scenes=5, seed=5, n_min=300, n_max=6000)
    walls = [scenes.sample_walls(HipSim, s) for s in sc]
    sims = []
    for persistent in (0, mode):
        cfg = scenes.make_config(sc[0])
        cfg.persistent_solver = persistent
        cfg.cg_fold_in_kernel = 0
        g = HipSim(cfg)
        scenes.load_batch(g, sc, walls)
        sims.append(g)
    a, b = sims
    apart = 0
    for t in range(12):
        sa, sb = a.step(), b.step()
        assert sb.cg_persistent == 1, t
        counts = []
        for q in range(len(sc)):
            qa, qb = a.scene_stats(q), b.scene_stats(q)
            assert (qa.l_df, qa.l_cg, qa.l_cd) == (qb.l_df, qb.l_cg, qb.l_cd), (t, q)
            n = min(qa.l_cg + 1, 128)
            assert list(qa.rr_cg[:n]) == list(qb.rr_cg[:n]), (t, q)
            counts.append(qa.l_cg)
        apart += len(set(counts)) > 1
        _same_state(a, b, t)
    assert apart > 0
    say(f"[persistent] batch of {len(sc)}: identical to chunked, scenes left the CG apart in {apart} of 12 steps")


@pytest.mark.parametrize("graphs", [0, 1])
def test_fold_in_last_workgroup_matches_fold_launches(graphs):
    """The chunked CG with its sums folded by the last workgroup of the A.p and update launches
    (cg_fold_in_kernel = 1, two launches per iteration) against the separate fold launches (0):
    the droplet for 4 steps, direct launches and graph replay."""
    sc = scenes.diff_droplet(n=65536, radius=0.25, seed=2, steps=4)
    a = _sim(sc, cg_fold_in_kernel=0, use_graphs=graphs)
    b = _sim(sc, cg_fold_in_kernel=1, use_graphs=graphs)
    say(f"[fold-in-kernel] droplet N={sc.n} graphs={graphs}")
    for t in range(4):
        sa, sb = a.step(), b.step()
        assert (sa.l_df, sa.l_cg, sa.l_cd) == (sb.l_df, sb.l_cg, sb.l_cd), t
        n = min(sa.l_cg + 1, 128)
        assert list(sa.rr_cg[:n]) == list(sb.rr_cg[:n]), t
        _same_state(a, b, t)
    say(f"[fold-in-kernel] l_cg {sa.l_cg}: identical to the fold launches")


def test_fold_in_last_workgroup_batch_scenes_converging_apart():
    sc = scenes.batch_scenes(n_scenes=5, seed=5, n_min=300, n_max=6000)
    walls = [scenes.sample_walls(HipSim, s) for s in sc]
    sims = []
    for fold_in in (0, 1):
        cfg = scenes.make_config(sc[0])
        cfg.cg_fold_in_kernel = fold_in
        g = HipSim(cfg)
        scenes.load_batch(g, sc, walls)
        sims.append(g)
    a, b = sims
    apart = 0
    for t in range(12):
        a.step(), b.step()
        counts = []
        for q in range(len(sc)):
            qa, qb = a.scene_stats(q), b.scene_stats(q)
            assert (qa.l_df, qa.l_cg, qa.l_cd) == (qb.l_df, qb.l_cg, qb.l_cd), (t, q)
            n = min(qa.l_cg + 1, 128)
            assert list(qa.rr_cg[:n]) == list(qb.rr_cg[:n]), (t, q)
            counts.append(qa.l_cg)
        apart += len(set(counts)) > 1
        _same_state(a, b, t)
    assert apart > 0
    say(f"[fold-in-kernel] batch of {len(sc)}: identical to the fold launches, scenes apart in {apart} of 12 steps")
