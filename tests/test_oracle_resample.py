"""Re-sampling a moved mesh in the oracle (ParticleSampling called again on the same MeshObject,
MPSParticleSamplingUtil.cu:14-133, .cuh:69-246), pinned without restating it:

- an unchanged mesh, or one moved rigidly by exact amounts, changes nothing (the reference
  keeps the old particles, so walls moved rigidly stay where they were sampled);
- after a stretch, every face whose sample count or shortest edge changed holds exactly the
  particles a fresh sampling of the stretched mesh gives that face, and every other face
  holds its old particles, at the offsets of the new counts;
- mass, velocity and volume are resized by index (a kept particle takes what was stored at
  its new index; new slots are zero), or travel with the particle without faithful quirks."""
import numpy as np
import pytest

from mps_cuda_amd import _capi as capi
from mps_cuda_amd.scenes import CUBE_TRIS, CUBE_VERTS
from oracle.binding import OracleSim, mesh_sample_counts

R_M = 0.04


def _sampled(faithful=1, prec=64):
    cfg = capi.default_config()
    cfg.faithful_quirks = faithful
    o = OracleSim(cfg, prec)
    o.sample_boundary_mesh(CUBE_VERTS, CUBE_TRIS, R_M, 2.0)
    return o


def _state(o):
    return {k: o.get(k).copy() for k in ("BND_POS", "BND_RADIUS", "BND_MASS", "BND_VEL", "BND_FACE_ID", "BND_BCC",
                                         "BND_VOLUME")}


def _short_edges(v, t):
    """Index of the shortest edge (e0 = n1-n0, e1 = n2-n1, e2 = n0-n2), ties to the lower index."""
    out = []
    for f in t:
        n = v[f]
        ls = [np.linalg.norm(n[1] - n[0]), np.linalg.norm(n[2] - n[1]), np.linalg.norm(n[0] - n[2])]
        i = 0
        if ls[i] > ls[1]:
            i = 1
        if ls[i] > ls[2]:
            i = 2
        out.append(i)
    return np.array(out)


def test_unchanged_mesh_changes_nothing():
    o = _sampled()
    before = _state(o)
    assert o.resample_boundary_mesh(CUBE_VERTS, R_M) == 0
    after = _state(o)
    assert all(np.array_equal(before[k], after[k]) for k in before)


def test_rigid_move_keeps_the_old_particles():
    o = _sampled()
    before = _state(o)
    # shifts by multiples of 0.25 keep every edge vector exact, so counts and short edges match
    assert o.resample_boundary_mesh(CUBE_VERTS + np.array([0.25, -0.5, 1.0]), R_M) == 0
    after = _state(o)
    assert all(np.array_equal(before[k], after[k]) for k in before)


@pytest.mark.parametrize("faithful", [1, 0])
def test_stretch_regenerates_changed_faces_like_a_fresh_sampling(faithful):
    o = _sampled(faithful)
    old = _state(o)
    B0 = len(old["BND_RADIUS"])
    vol = np.arange(1, B0 + 1, dtype=np.float64) * 1e-3      # distinct volumes to follow by index
    o.set("BND_VOLUME", vol)
    vel = np.zeros((B0, 3))
    vel[:, 0] = np.arange(B0)                                  # distinct velocities
    o.set("BND_VEL", vel)

    new_v = CUBE_VERTS.copy()
    new_v[new_v[:, 0] == 1.0, 0] = 1.5                         # stretch the x = 1 side outwards
    c_old = mesh_sample_counts(CUBE_VERTS, CUBE_TRIS, R_M).astype(np.int64)
    c_new = mesh_sample_counts(new_v, CUBE_TRIS, R_M).astype(np.int64)
    changed = (c_old != c_new) | (_short_edges(CUBE_VERTS, CUBE_TRIS) != _short_edges(new_v, CUBE_TRIS))
    assert 0 < changed.sum() < len(CUBE_TRIS)
    assert o.resample_boundary_mesh(new_v, R_M) == int(changed.sum())
    new = _state(o)

    fresh = OracleSim(capi.default_config(), 64)
    fresh.sample_boundary_mesh(new_v, CUBE_TRIS, R_M, 2.0)
    fr = _state(fresh)
    B1 = int(c_new.sum())
    assert len(new["BND_RADIUS"]) == B1 == len(fr["BND_RADIUS"])
    off_old = np.concatenate([[0], np.cumsum(c_old)])
    off_new = np.concatenate([[0], np.cumsum(c_new)])
    for t in range(len(CUBE_TRIS)):
        a, b = off_new[t], off_new[t + 1]
        if changed[t]:
            for k in ("BND_POS", "BND_RADIUS", "BND_FACE_ID", "BND_BCC"):
                assert np.array_equal(new[k][a:b], fr[k][a:b]), (t, k)
            # masses come from the vertex masses of the first sampling (area of the unit cube
            # 6 m^2, not the stretched one): same weights, scaled by the area ratio
            area_new = 2 * 1.0 + 4 * 1.5
            np.testing.assert_allclose(new["BND_MASS"][a:b], fr["BND_MASS"][a:b] * 6.0 / area_new, rtol=1e-12)
        else:
            p, q = off_old[t], off_old[t + 1]
            assert b - a == q - p
            for k in ("BND_POS", "BND_RADIUS", "BND_FACE_ID", "BND_BCC"):
                assert np.array_equal(new[k][a:b], old[k][p:q]), (t, k)
            if faithful:
                idx = np.arange(a, b)
                keep = np.where(idx < B0, old["BND_MASS"][np.minimum(idx, B0 - 1)], 0.0)
                assert np.array_equal(new["BND_MASS"][a:b], keep), t
                assert np.array_equal(new["BND_VEL"][a:b, 0], np.where(idx < B0, idx, 0).astype(float)), t
            else:
                assert np.array_equal(new["BND_MASS"][a:b], old["BND_MASS"][p:q]), t
                assert np.array_equal(new["BND_VEL"][a:b, 0], np.arange(p, q, dtype=float)), t
    # volumes never move with the particles
    n = min(B0, B1)
    assert np.array_equal(new["BND_VOLUME"][:n], vol[:n]) and not new["BND_VOLUME"][n:].any()


def test_resample_needs_a_sampled_mesh():
    o = OracleSim(capi.default_config(), 64)
    with pytest.raises(RuntimeError):
        o.resample_boundary_mesh(CUBE_VERTS, R_M)
    o = _sampled()
    with pytest.raises(RuntimeError):
        o.resample_boundary_mesh(CUBE_VERTS[:7], R_M)
    o.upload_boundary(np.zeros((3, 3)), np.full(3, 0.01))
    with pytest.raises(RuntimeError):
        o.resample_boundary_mesh(CUBE_VERTS, R_M)
