"""Moving walls: the reference scene's box is stretched after a few steps and re-sampled
(ParticleSampling on the moved mesh, MPSParticleSamplingUtil.cu:14-133) on the GPU and in the
FP32 oracle. The re-sampled wall arrays must be identical (positions, radii, face ids,
barycentric weights, masses and velocities bit for bit; volumes, which come from the steps
before, to the step tolerance), and the steps after it must keep step parity. The Z-sort of
frame 1 has permuted the wall arrays before the re-sampling, so the copy of unchanged faces
by their old offsets (the reference's rule) is exercised on permuted data too."""
import sys

import numpy as np
import pytest

from mps_cuda_amd import scenes
from tests.parity import make_pair, rel_err

pytestmark = pytest.mark.gpu
TOL = 1e-5


def _say(m):
    print(m, flush=True)
    sys.stdout.flush()


def _fitted(sc):
    b = sc.boundary_mesh
    c, s = np.asarray(b["center"], np.float64), np.asarray(b["size"], np.float64)
    return (scenes.CUBE_VERTS - 0.5) * s + c     # Translate of the unit cube (its AABB is [0, 1]^3)


@pytest.mark.parametrize("faithful", [1, 0])
def test_resampled_walls_match_the_oracle(faithful):
    sc = scenes.surface_tension_test_scene(4)
    sc.config["faithful_quirks"] = faithful
    g, o = make_pair(sc)
    _say(f"[resample] faithful={faithful}: {g.get('BND_POS').shape[0]} wall particles")
    for _ in range(4):
        g.step()
        o.step()
    v = _fitted(sc)
    v[scenes.CUBE_VERTS[:, 0] == 1.0, 0] += 1.0          # push the +x wall out by 1 m
    kg = g.resample_boundary_mesh(v, sc.boundary_mesh["radius"])
    ko = o.resample_boundary_mesh(v, sc.boundary_mesh["radius"])
    _say(f"[resample] faces generated anew: gpu {kg}, oracle {ko}; walls now {g.get('BND_POS').shape[0]}")
    assert kg == ko > 0
    for k in ("BND_POS", "BND_RADIUS", "BND_FACE_ID", "BND_BCC", "BND_MASS", "BND_VEL"):
        a, b = g.get(k), o.get(k)
        assert a.shape == b.shape and np.array_equal(a, b.astype(a.dtype)), k
    assert rel_err(g.get("BND_VOLUME"), o.get("BND_VOLUME")) <= TOL
    for s in range(4):
        sg, so = g.step(), o.step()
        assert (sg.l_df, sg.l_cg, sg.l_cd) == (so.l_df, so.l_cg, so.l_cd), s
        assert rel_err(g.get("FLUID_POS"), o.get("FLUID_POS")) <= TOL, s
        assert rel_err(g.get("BND_VOLUME"), o.get("BND_VOLUME")) <= TOL, s
    # the same stretch again changes nothing
    assert g.resample_boundary_mesh(v, sc.boundary_mesh["radius"]) == 0


def test_resample_without_a_sampled_mesh_is_a_state_error():
    from mps_cuda_amd import HipSim, MPSError
    sc = scenes.surface_tension_test_scene(3)
    g = HipSim(scenes.make_config(sc))
    with pytest.raises(MPSError) as e:
        g.resample_boundary_mesh(scenes.CUBE_VERTS, 0.4)
    assert e.value.code == 4
