"""Several wall objects on the GPU (mps_set_boundary_objects) against the oracle: the cube
walls of a small dam break split into two touching objects plus a far plate as a third.
Wall positions after the per-object Z-sort must be identical, volumes and fluid positions
within the step tolerance, loop counts equal, over 20 steps."""
import sys

import numpy as np
import pytest

from mps_cuda_amd import HipSim, scenes
from oracle.binding import OracleSim
from tests.parity import rel_err

pytestmark = pytest.mark.gpu
TOL = 1e-5


def _load(sim, sc, walls, counts):
    p, m = sc.physics, sc.material
    sim.set_physics(p["dt"], p["gravity"], p["box_min"], p["box_max"])
    sim.set_fluid_material(m["radius"], m["density"], m["viscosity"], m["xsph"], m["surface_tension"])
    sim.upload_boundary(walls[0], walls[1], walls[2])
    sim.set_boundary_objects(counts)
    sim.upload_fluid(sc.pos, sc.radius[0], None)


def test_wall_objects_match_the_oracle():
    sc = scenes.dam_break_cpu(n_side=12)
    cube = scenes.sample_walls(HipSim, sc)
    g0 = np.arange(12) * 0.01 + 0.3
    xx, zz = np.meshgrid(g0, g0, indexing="ij")
    plate = np.stack([xx.ravel(), np.full(xx.size, 2.0), zz.ravel()], axis=1).astype(np.float32)
    walls = (np.concatenate([cube[0], plate]), np.concatenate([cube[1], np.full(len(plate), 0.01, np.float32)]),
             np.concatenate([cube[2], np.zeros(len(plate), np.float32)]))
    n = len(cube[0])
    counts = [n // 2, n - n // 2, len(plate)]
    g = HipSim(scenes.make_config(sc))
    o = OracleSim(scenes.make_config(sc), 32)
    _load(g, sc, walls, counts)
    _load(o, sc, walls, counts)
    assert np.array_equal(g.get("BND_OBJECT"), o.get("BND_OBJECT"))
    for s in range(20):
        sg, so = g.step(), o.step()
        assert (sg.l_df, sg.l_cd, sg.l_cg, sg.zsorted) == (so.l_df, so.l_cd, so.l_cg, so.zsorted), s
        assert rel_err(g.get("FLUID_POS"), o.get("FLUID_POS")) <= TOL, s
        assert rel_err(g.get("BND_VOLUME"), o.get("BND_VOLUME")) <= TOL, s
    assert np.array_equal(g.get("BND_POS"), o.get("BND_POS"))
    print(f"[walls] 3 objects ({counts}), 20 steps: loops {sg.l_df}/{sg.l_cd}/{sg.l_cg}, "
          f"pos err {rel_err(g.get('FLUID_POS'), o.get('FLUID_POS')):.2e}", flush=True)
    sys.stdout.flush()
