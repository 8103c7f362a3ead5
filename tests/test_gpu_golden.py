"""The gfx950 library against the committed FP32 golden vectors (tests/golden/, written by the
oracle): six steps of the reference scene and of a small dam break. Neighbour lists must be
identical, loop counts equal and every step's positions and velocities within 1e-5 relative."""
import os
import sys

import numpy as np
import pytest

from mps_cuda_amd import HipSim, scenes
from tests.golden import make_golden as mg
from tests.parity import rel_err

pytestmark = pytest.mark.gpu
HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.mark.parametrize("name", sorted(mg.golden_scenes()))
def test_gpu_matches_golden_fp32(name):
    with np.load(os.path.join(HERE, f"steps_{name}_fp32.npz"), allow_pickle=False) as f:
        want = {k: f[k] for k in f.files}
    sc = mg.golden_scenes()[name]
    g = HipSim(scenes.make_config(sc))
    scenes.load_into(g, sc)
    for s in range(mg.STEPS):
        st = g.step()
        assert [st.l_df, st.l_cd, st.l_cg, st.zsorted] == want["loops"][s].tolist(), s
        assert rel_err(g.get("FLUID_POS"), want["pos"][s]) <= 1e-5, s
        assert rel_err(g.get("FLUID_VEL"), want["vel"][s]) <= 1e-5, s
    for k, arr in (("nei_ff_idx", "NEI_FF_IDX"), ("nei_ff", "NEI_FF"), ("nei_fb_idx", "NEI_FB_IDX"), ("nei_fb", "NEI_FB")):
        assert np.array_equal(g.get(arr), want[k]), k
    assert rel_err(g.get("FLUID_DENSITY"), want["density"]) <= 1e-5
    print(f"[golden] {name}: {mg.STEPS} steps match, pos err {rel_err(g.get('FLUID_POS'), want['pos'][-1]):.2e}", flush=True)
    sys.stdout.flush()
