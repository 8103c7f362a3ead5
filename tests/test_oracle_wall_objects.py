"""Several wall objects in the oracle: each wall's Akinci volume sums its own object and then the
other objects (the boundary "ref" overload, MPSSPHBaseUtil.cu:25-42, .cuh:43-76); the fluid's
wall sums go object by object; each object is Z-sorted within its own range. Pinned by:

- objects that never come near each other (or the same fluid) give the one-object result bit
  for bit, through a Z-sort that keeps them apart;
- a cube split into two touching objects changes volumes and densities only by rounding
  (FP64): the per-object sums add up to the one-object sum;
- two wall particles of different objects and radii follow the closed form, including the
  volume carried over from the previous step (never cleared in the reference)."""
import numpy as np
import pytest

from mps_cuda_amd import _capi as capi
from mps_cuda_amd import scenes
from oracle.binding import OracleSim, kernel_eval


def _plate(n=12, y=2.0, spacing=0.01):
    g = np.arange(n) * spacing
    xx, zz = np.meshgrid(g + 0.3, g + 0.3, indexing="ij")
    return np.stack([xx.ravel(), np.full(xx.size, y), zz.ravel()], axis=1)


def _ctx(sc, walls, counts=None, prec=32):
    o = OracleSim(scenes.make_config(sc), prec)
    p, m = sc.physics, sc.material
    o.set_physics(p["dt"], p["gravity"], p["box_min"], p["box_max"])
    o.set_fluid_material(m["radius"], m["density"], m["viscosity"], m["xsph"], m["surface_tension"])
    o.upload_boundary(walls[0], walls[1], walls[2])
    if counts is not None:
        o.set_boundary_objects(counts)
    o.upload_fluid(sc.pos, sc.radius[0], None)
    return o


def _cube_walls(sc):
    s = OracleSim(scenes.make_config(sc), 64)
    p = sc.physics
    s.set_physics(p["dt"], p["gravity"], p["box_min"], p["box_max"])
    b = sc.boundary_mesh
    s.sample_boundary_mesh(b["verts"], b["tris"], b["radius"], b.get("density", 1.0), b.get("center"), b.get("size"))
    return s.get("BND_POS"), s.get("BND_RADIUS"), s.get("BND_MASS")


def test_objects_apart_equal_one_object():
    sc = scenes.dam_break_cpu(n_side=8)
    cube = _cube_walls(sc)
    plate = _plate()
    walls = (np.concatenate([cube[0], plate]), np.concatenate([cube[1], np.full(len(plate), 0.01)]),
             np.concatenate([cube[2], np.zeros(len(plate))]))
    one = _ctx(sc, walls)
    two = _ctx(sc, walls, [len(cube[0]), len(plate)])
    assert two.get("BND_OBJECT").tolist() == [0] * len(cube[0]) + [1] * len(plate)
    for s in range(6):
        a, b = one.step(), two.step()
        assert (a.l_df, a.l_cd, a.l_cg, a.zsorted) == (b.l_df, b.l_cd, b.l_cg, b.zsorted), s
    assert np.array_equal(one.get("FLUID_POS"), two.get("FLUID_POS"))
    # the Z-sort kept the plate at the end in the two-object run
    assert np.array_equal(two.get("BND_OBJECT")[-len(plate):], np.ones(len(plate)))
    assert np.array_equal(np.sort(two.get("BND_POS")[-len(plate):], axis=0), np.sort(plate.astype(np.float32), axis=0))


def test_split_walls_sum_to_one_object():
    sc = scenes.dam_break_cpu(n_side=8)
    cube = _cube_walls(sc)
    n = len(cube[0])
    one = _ctx(sc, cube, prec=64)
    two = _ctx(sc, cube, [n // 2, n - n // 2], prec=64)
    for ph in ("HASH", "NEIGHBORS", "BOUNDARY_VOLUME", "DENSITY", "ALPHA"):
        one.run_phase(ph)
        two.run_phase(ph)
    for name in ("BND_VOLUME", "FLUID_DENSITY", "FLUID_ALPHA"):
        np.testing.assert_allclose(two.get(name), one.get(name), rtol=1e-13, atol=0, err_msg=name)
    # the split really separates neighbours: walls of object 0 have object-1 neighbours
    idx, nei = two.get("NEI_BB_IDX"), two.get("NEI_BB")
    obj = two.get("BND_OBJECT")
    cross = sum(int((obj[nei[idx[i]:idx[i + 1]]] != obj[i]).any()) for i in range(n))
    assert cross > 20


def test_two_wall_particles_closed_form():
    sc = scenes.dam_break_cpu(n_side=2)
    cfg = scenes.make_config(sc)
    cfg.kernel = capi.KERNEL_CUBIC
    o = OracleSim(cfg, 64)
    o.set_physics(1e-3, (0.0, 0.0, 0.0), (-1.0,) * 3, (2.0,) * 3)
    o.set_fluid_material(0.04, 1000.0, 0.0, 0.0, 0.0)
    ra, rb, d = 0.01, 0.015, 0.03                     # neighbours: d < (4 ra + 4 rb) / 2 = 0.05
    o.upload_boundary(np.array([[0.5, 0.5, 0.5], [0.5 + d, 0.5, 0.5]]), np.array([ra, rb]), None)
    o.set_boundary_objects([1, 1])
    o.upload_fluid(np.array([[0.1, 0.1, 0.1]]), 0.01, None)
    W = lambda q, h: float(kernel_eval(64, capi.KERNEL_CUBIC, np.array([q]), 1.0 / h)[0][0])
    ha, hb = 4 * ra, 4 * rb
    va = vb = 0.0
    for _ in range(2):
        for ph in ("HASH", "NEIGHBORS", "BOUNDARY_VOLUME"):
            o.run_phase(ph)
        va = 1.0 / ((va + W(0.0, ha)) + W(d, ha))     # own object (itself), then the other object
        vb = 1.0 / ((vb + W(0.0, hb)) + W(d, hb))
        got = o.get("BND_VOLUME")
        assert got[0] == pytest.approx(va, rel=1e-14) and got[1] == pytest.approx(vb, rel=1e-14)
