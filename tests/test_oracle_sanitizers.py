"""The oracle under AddressSanitizer + UndefinedBehaviorSanitizer (host code only): the sanitizer
build (make -C oracle asan) is loaded into a child Python with libasan preloaded and driven
through mesh sampling and re-sampling, wall objects, a coupled two-fluid world, both smoothing
kernels and both viscosity solvers. Any report fails the test (UB aborts; ASan exits non-zero)."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

SCRIPT = r"""
import numpy as np
from mps_cuda_amd import _capi as capi
from mps_cuda_amd import scenes
from oracle.binding import OracleSim, OracleWorld, set_threads
set_threads(2)
sc = scenes.dam_break_cpu(n_side=6)
for kernel, visc in ((capi.KERNEL_POLY6_SPIKY, capi.VISC_CG), (capi.KERNEL_CUBIC, capi.VISC_JACOBI)):
    cfg = scenes.make_config(sc)
    cfg.kernel, cfg.visc_solver = kernel, visc
    o = OracleSim(cfg, 32)
    scenes.load_into(o, sc)
    for _ in range(3):
        o.step()
v = (scenes.CUBE_VERTS - 0.5) * 1.0 + 0.5
v[scenes.CUBE_VERTS[:, 0] == 1.0, 0] += 0.3
assert o.resample_boundary_mesh(v, sc.boundary_mesh["radius"]) > 0
o.step()
B = o.get("BND_POS").shape[0]
o.set_boundary_objects([B // 2, B - B // 2])
o.step()
w = OracleWorld(scenes.make_config(sc), 2, 64)
cut = sc.pos[:, 0] < np.median(sc.pos[:, 0])
for k, part in enumerate((sc.pos[cut], sc.pos[~cut])):
    p, m = w.phase(k), sc.material
    p.set_physics(sc.physics["dt"], sc.physics["gravity"], sc.physics["box_min"], sc.physics["box_max"])
    p.set_fluid_material(m["radius"], m["density"] * (1 + k), m["viscosity"], m["xsph"], m["surface_tension"])
    p.upload_fluid(part, sc.radius[0], None)
for _ in range(3):
    w.step()
print("sanitized run ok", flush=True)
"""


def test_oracle_under_asan_and_ubsan():
    libasan = subprocess.run(["g++", "-print-file-name=libasan.so"], capture_output=True, text=True).stdout.strip()
    if not libasan or not os.path.isabs(libasan) or not os.path.exists(libasan):
        pytest.skip("libasan not available")
    r = subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "asan"], capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr
    env = dict(os.environ, LD_PRELOAD=libasan, ASAN_OPTIONS="detect_leaks=0",
               MPS_ORACLE_LIB=os.path.join(ROOT, "oracle", "_build", "libmps_oracle_asan.so"))
    r = subprocess.run([sys.executable, "-c", SCRIPT], cwd=ROOT, env=env, capture_output=True, text=True, timeout=600)
    out = r.stdout + r.stderr
    assert r.returncode == 0 and "sanitized run ok" in out, out[-4000:]
    assert "AddressSanitizer" not in out and "runtime error" not in out, out[-4000:]
