"""Batched scenes on the GPU: several independent scenes stepped in one context give, scene
by scene, exactly the results of the same scene stepped alone (bit-identical positions,
velocities, densities, neighbour lists and iteration counts), while each scene's solver loops
converge on their own."""
import sys

import numpy as np
import pytest

from mps_cuda_amd import HipSim, MPSError, scenes

pytestmark = pytest.mark.gpu


def say(m):
    print(m, flush=True)
    sys.stdout.flush()


def _make(n_scenes, seed, n_min, n_max):
    sc = scenes.batch_scenes(n_scenes=n_scenes, seed=seed, n_min=n_min, n_max=n_max)
    walls = [scenes.sample_walls(HipSim, s) for s in sc]
    return sc, walls


def _slice(a, off, n):
    return a[off:off + n]


def test_batch_matches_scenes_alone():
    sc, walls = _make(4, 5, 300, 3000)
    say(f"[batch] sizes {[s.n for s in sc]} kinds {[s.extra['kind'] for s in sc]}")
    cfg = scenes.make_config(sc[0])
    bat = HipSim(cfg)
    scenes.load_batch(bat, sc, walls)
    assert bat.scene_count == len(sc)
    alone = []
    for s, w in zip(sc, walls):
        g = HipSim(scenes.make_config(s))
        scenes.load_plain(g, s, w)
        alone.append(g)
    info = [bat.scene_info(q) for q in range(len(sc))]
    assert [i["n_fluid"] for i in info] == [s.n for s in sc]
    differing = 0
    for step in range(25):
        st = bat.step()
        ls = []
        for q, g in enumerate(alone):
            sa = g.step()
            sb = bat.scene_stats(q)
            assert (sb.l_df, sb.l_cd, sb.l_cg) == (sa.l_df, sa.l_cd, sa.l_cg), (step, q)
            assert list(sb.err_cd[:sa.l_cd]) == list(sa.err_cd[:sa.l_cd]), (step, q)
            ls.append((sa.l_df, sa.l_cd, sa.l_cg))
        assert st.l_cg == max(l[2] for l in ls) and st.l_cd == max(l[1] for l in ls)
        differing += len(set(ls)) > 1
    assert differing > 0  # scenes left their loops at different iterations
    for name in ("FLUID_POS", "FLUID_VEL", "FLUID_DENSITY", "FLUID_ALPHA"):
        a = bat.get(name)
        for q, g in enumerate(alone):
            assert np.array_equal(_slice(a, info[q]["fluid_offset"], info[q]["n_fluid"]), g.get(name)), (name, q)
    idx, nei = bat.get("NEI_FF_IDX"), bat.get("NEI_FF")
    fidx, fnei = bat.get("NEI_FB_IDX"), bat.get("NEI_FB")
    for q, g in enumerate(alone):
        o, n = info[q]["fluid_offset"], info[q]["n_fluid"]
        lo, hi = idx[o], idx[o + n]
        assert np.array_equal(idx[o:o + n + 1] - lo, g.get("NEI_FF_IDX"))
        assert np.array_equal(nei[lo:hi] - o, g.get("NEI_FF"))
        flo, fhi = fidx[o], fidx[o + n]
        assert np.array_equal(fnei[flo:fhi] - info[q]["boundary_offset"], g.get("NEI_FB"))
    say(f"[batch] 25 steps identical; loops differed across scenes in {differing} steps")
    with pytest.raises(MPSError):
        bat.get("FHASH_KEY")
    with pytest.raises(MPSError):
        bat.upload_fluid(np.zeros((1, 3)), 0.01)


def test_batch_material_change_and_recommit():
    sc, walls = _make(3, 11, 400, 1200)
    bat = HipSim(scenes.make_config(sc[0]))
    scenes.load_batch(bat, sc, walls)
    g = HipSim(scenes.make_config(sc[1]))
    scenes.load_plain(g, sc[1], walls[1])
    for _ in range(3):
        bat.step()
        g.step()
    m = dict(sc[1].material)
    m["viscosity"], m["surface_tension"] = 0.04, 0.2
    bat.scene_set_material(1, **m)
    g.set_fluid_material(**m)
    for _ in range(3):
        bat.step()
        g.step()
    o, n = bat.scene_info(1)["fluid_offset"], sc[1].n
    assert np.array_equal(bat.get("FLUID_VEL")[o:o + n], g.get("FLUID_VEL"))
    # a new batch replaces the old one
    scenes.load_batch(bat, sc[:1], walls[:1])
    assert bat.scene_count == 1 and bat.get("FLUID_POS").shape[0] == sc[0].n
    bat.step()


def test_batch_graph_replay_matches_direct_launches():
    """The batched step with its solver chunks replayed from cached HIP graphs (the graph-
    captured batched step) gives the same bits as direct launches, scene by scene, with the
    scenes' loops ending at different iterations."""
    sc, walls = _make(4, 7, 300, 3000)
    sims = []
    for graphs in (0, 1):
        cfg = scenes.make_config(sc[0])
        cfg.use_graphs = graphs
        g = HipSim(cfg)
        scenes.load_batch(g, sc, walls)
        sims.append(g)
    for step in range(20):
        sa, sb = sims[0].step(), sims[1].step()
        assert (sa.l_df, sa.l_cd, sa.l_cg) == (sb.l_df, sb.l_cd, sb.l_cg), step
        for q in range(len(sc)):
            qa, qb = sims[0].scene_stats(q), sims[1].scene_stats(q)
            assert (qa.l_df, qa.l_cd, qa.l_cg) == (qb.l_df, qb.l_cd, qb.l_cg), (step, q)
        for name in ("FLUID_POS", "FLUID_VEL"):
            assert np.array_equal(sims[0].get(name), sims[1].get(name)), (step, name)
    say("[batch] 20 graph-replayed batched steps identical to direct launches")
