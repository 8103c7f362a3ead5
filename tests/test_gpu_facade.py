"""The C++ classes of include/mps_system.hpp driven like the reference's application
(System::Initalize, the simulation keys, OnUpdate per frame; MPSSystem.cpp:19-291) give the
same state, bit for bit, as the same scene loaded and stepped through the Python handle: the
header adds no arithmetic of its own. The C++ program (tests/cpp/facade_scene.cpp) runs as a
child process."""
import os
import subprocess
import sys

import numpy as np
import pytest

from mps_cuda_amd import HipSim, build, scenes

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
CUBE = os.path.join(HERE, "golden", "cube.obj")


def _facade(keys, steps, out):
    exe = build.FACADE_BIN
    assert os.path.exists(exe), "facade driver not built (python -m mps_cuda_amd.build)"
    print(f"[facade] keys={keys!r} steps={steps}", flush=True)
    r = subprocess.run([exe, "run", CUBE, keys, str(steps), str(out)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    b = out.read_bytes()
    n, frame, nb = np.frombuffer(b[:24], np.uint64)
    loops = np.frombuffer(b[24:40], np.uint32)
    pos = np.frombuffer(b[40:], np.float32).reshape(int(n), 3)
    return int(frame), int(nb), loops[:3].tolist(), pos


@pytest.mark.parametrize("keys,size,gravity", [("", 5, -9.8), ("+g", 10, 0.0), ("w", 5, -9.8)])
def test_facade_run_equals_python_path(tmp_path, keys, size, gravity):
    """'w': the +x wall moves out by 1 m after half of the steps and is re-sampled."""
    steps = 12
    frame, nb, loops, pos = _facade(keys, steps, tmp_path / "out.bin")
    sc = scenes.surface_tension_test_scene(size)
    sc.physics["gravity"] = (0.0, gravity, 0.0)
    g = HipSim(scenes.make_config(sc))
    scenes.load_into(g, sc)
    b = sc.boundary_mesh
    walls = (scenes.CUBE_VERTS - 0.5) * np.asarray(b["size"]) + np.asarray(b["center"])
    walls[scenes.CUBE_VERTS[:, 0] == 1.0, 0] += 1.0
    for s in range(steps):
        if "w" in keys and s == steps // 2:
            assert g.resample_boundary_mesh(walls, b["radius"]) > 0
        st = g.step()
    ref = g.get("FLUID_POS").astype(np.float32)
    print(f"[facade] size {size}: {pos.shape[0]} particles, {nb} walls, loops {loops}", flush=True)
    sys.stdout.flush()
    assert frame == steps and pos.shape == ref.shape
    assert nb == g.get("BND_POS").shape[0]
    assert loops == [st.l_df, st.l_cd, st.l_cg]
    assert np.array_equal(pos, ref)


def test_facade_spatial_hash_lists_equal_python_path(tmp_path):
    """SpatialHash::UpdateHash / BuildNeighorhood / GetNeighborhood through the C++ classes give
    the same ids and CSR lists (reference order) as the HASH and NEIGHBORS phases run from Python."""
    exe = build.FACADE_BIN
    out = tmp_path / "lists.bin"
    print("[facade] lists", flush=True)
    r = subprocess.run([exe, "lists", CUBE, str(out)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    b = out.read_bytes()
    counts = np.frombuffer(b[:56], np.uint64).astype(int)
    arrs, at = [], 56
    for n in counts:
        arrs.append(np.frombuffer(b[at:at + 4 * n], np.uint32))
        at += 4 * n
    sc = scenes.surface_tension_test_scene(5)
    g = HipSim(scenes.make_config(sc))
    scenes.load_into(g, sc)
    g.run_phase("HASH")
    g.run_phase("NEIGHBORS")
    names = ["FHASH_ID", "NEI_FF_IDX", "NEI_FF", "NEI_FB_IDX", "NEI_FB", "NEI_BB_IDX", "NEI_BB"]
    for name, a in zip(names, arrs):
        assert np.array_equal(a, g.get(name)), name
    assert counts[2] > 0 and counts[6] > 0    # (the fluid starts 3.5 m from the walls: no fluid->wall pairs yet)
