"""Boundary sampling of the oracle (MPSParticleSamplingUtil.cuh:10-294, MPSMeshObject.cuh:41-101)
pinned by hand-derived counts and ownership invariants."""
import itertools

import numpy as np

from mps_cuda_amd import _capi as capi
from mps_cuda_amd.scenes import CUBE_TRIS, CUBE_VERTS
from oracle.binding import OracleSim, mesh_rtri, mesh_sample_counts

SQUARE_V = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0]], np.float64)
SQUARE_T = np.array([[0, 1, 2], [0, 2, 3]], np.uint32)


def test_unit_square_counts():
    # d = R_m/2 = 0.1. Face 0 owns all three vertices and its three edges (the diagonal is
    # shared with face 1, which has the larger id): 3 + 9 + 9 + ceil((sqrt2-0.101)/0.1)=14,
    # plus 9 scan lines of lengths 0.9..0.1 -> 8+7+...+1+0 = 36 points: 71.
    # Face 1 owns vertex 3 and edges (2,3), (3,0): 1 + 9 + 9 + 36 = 55.
    counts = mesh_sample_counts(SQUARE_V, SQUARE_T, 0.2)
    assert counts.tolist() == [71, 55]


def _edges(tris):
    out = {}
    for f, t in enumerate(tris):
        for i in range(3):
            a, b = int(t[i]), int(t[(i + 1) % 3])
            out.setdefault(frozenset((a, b)), []).append((f, i))
    return out


def test_cube_vertices_and_edges_owned_once():
    rt = mesh_rtri(8, CUBE_TRIS)
    owners_v = {v: 0 for v in range(8)}
    for f, t in enumerate(CUBE_TRIS):
        for i in range(3):
            if (rt[f] >> i) & 1:
                owners_v[int(t[i])] += 1
    assert all(c == 1 for c in owners_v.values())
    for e, uses in _edges(CUBE_TRIS).items():
        owned = sum((rt[f] >> (i + 3)) & 1 for f, i in uses)
        assert owned == 1, e


def test_cube_samples_on_surface_without_duplicates():
    cfg = capi.default_config()
    o = OracleSim(cfg, 64)
    m = capi.MeshMaterial(0.04, 1.0)
    o.sample_boundary_mesh(CUBE_VERTS, CUBE_TRIS, 0.04, 1.0, [0.5, 0.5, 0.5], [1.0, 1.0, 1.0])
    p = o.get("BND_POS")
    assert len(p) == int(mesh_sample_counts(CUBE_VERTS, CUBE_TRIS, 0.04).sum())
    on_face = np.any(np.isclose(p, 0.0, atol=1e-12) | np.isclose(p, 1.0, atol=1e-12), axis=1)
    assert on_face.all()
    assert np.all((p >= -1e-12) & (p <= 1 + 1e-12))
    r = o.get("BND_RADIUS")
    assert np.all(r == 0.01)
    # no two samples closer than half the spacing (0.02): each corner/edge point appears once
    from scipy.spatial import cKDTree
    pairs = cKDTree(p).query_pairs(0.01)
    assert len(pairs) == 0
    # every cube corner is sampled exactly once
    for c in itertools.product([0.0, 1.0], repeat=3):
        assert np.sum(np.all(np.isclose(p, c, atol=1e-12), axis=1)) == 1
