"""Coupled fluid objects on the GPU (mps_config.couple_scenes): two fluids of different rest
density and viscosity filling one tank, as a batch of two scenes, against the oracle's
coupled world (the reference's "ref" overloads of density, factor, delta and pressure force,
joint pressure loops). Loop counts must agree at every step and positions within the step
tolerance; the same batch without coupling must differ (the coupling is really exercised)."""
import sys

import numpy as np
import pytest

from mps_cuda_amd import HipSim, scenes
from oracle.binding import OracleWorld
from tests.parity import rel_err

pytestmark = pytest.mark.gpu
TOL = 1e-5


def _say(m):
    print(m, flush=True)
    sys.stdout.flush()


def _phases(tight=False):
    sc = scenes.dam_break_cpu(n_side=12)
    if tight:  # the joint loops then run past the faster phase's own stopping point
        sc.config.update(df_tol=2e-5, cd_tol=1e-4)
    walls = scenes.sample_walls(HipSim, sc)
    cut = sc.pos[:, 0] < np.median(sc.pos[:, 0])
    mats = [dict(sc.material), dict(sc.material, density=600.0, viscosity=0.03, surface_tension=0.1)]
    vels = [None, np.tile([0.0, -1.5, 0.5], ((~cut).sum(), 1)) if tight else None]
    return sc, walls, [sc.pos[cut], sc.pos[~cut]], mats, vels


def _gpu(sc, walls, parts, mats, vels, coupled):
    cfg = scenes.make_config(sc)
    cfg.couple_scenes = int(coupled)
    g = HipSim(cfg)
    g.batch_clear()
    for p, m, v in zip(parts, mats, vels):
        g.batch_add_scene(sc.physics, m, p, sc.radius[0], vel=v, bpos=walls[0], bradius=walls[1], bmass=walls[2])
    g.batch_commit()
    return g


@pytest.mark.parametrize("tight", [False, True])
def test_coupled_phases_match_the_oracle_world(tight):
    sc, walls, parts, mats, vels = _phases(tight)
    g = _gpu(sc, walls, parts, mats, vels, True)
    w = OracleWorld(scenes.make_config(sc), 2, 32)
    for k in range(2):
        o = w.phase(k)
        p = sc.physics
        o.set_physics(p["dt"], p["gravity"], p["box_min"], p["box_max"])
        m = mats[k]
        o.set_fluid_material(m["radius"], m["density"], m["viscosity"], m["xsph"], m["surface_tension"])
        o.upload_boundary(walls[0], walls[1], walls[2])
        o.upload_fluid(parts[k], sc.radius[0], vels[k])
    n0 = len(parts[0])
    loops = set()
    for s in range(16):
        g.step()
        so = w.step()
        loops.add((so[0].l_df, so[0].l_cd))
        for q in range(2):
            sg = g.scene_stats(q)
            assert (sg.l_df, sg.l_cd, sg.l_cg, sg.zsorted) == (so[q].l_df, so[q].l_cd, so[q].l_cg, so[q].zsorted), (s, q)
        x = g.get("FLUID_POS")
        e = [rel_err(x[:n0], w.phase(0).get("FLUID_POS")), rel_err(x[n0:], w.phase(1).get("FLUID_POS"))]
        assert max(e) <= TOL, (s, e)
    _say(f"[phases] tight={tight}: 16 coupled steps, (l_df, l_cd) seen {sorted(loops)}, pos errs {e}")
    if tight:
        assert max(l for _, l in loops) > 3
    u = _gpu(sc, walls, parts, mats, vels, False)
    for _ in range(16):
        u.step()
    d = rel_err(u.get("FLUID_POS"), g.get("FLUID_POS"))
    # uncoupled, the particles at the interface miss the other fluid's share of their density
    rd = np.abs(u.get("FLUID_DENSITY") / g.get("FLUID_DENSITY") - 1.0).max()
    _say(f"[phases] uncoupled vs coupled: positions {d:.3e}, largest density ratio - 1 {rd:.3f}")
    assert d > 1e-6 and rd > 0.05


def test_coupled_scenes_need_one_physics():
    from mps_cuda_amd import MPSError
    sc, walls, parts, mats, _ = _phases()
    cfg = scenes.make_config(sc)
    cfg.couple_scenes = 1
    g = HipSim(cfg)
    g.batch_clear()
    g.batch_add_scene(sc.physics, mats[0], parts[0], sc.radius[0], bpos=walls[0], bradius=walls[1], bmass=walls[2])
    g.batch_add_scene(dict(sc.physics, dt=sc.physics["dt"] * 2), mats[1], parts[1], sc.radius[0], bpos=walls[0],
                      bradius=walls[1], bmass=walls[2])
    with pytest.raises(MPSError):
        g.batch_commit()
