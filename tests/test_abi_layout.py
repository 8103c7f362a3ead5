"""The ctypes mirrors in mps_cuda_amd/_capi.py against the C declarations in include/mps_sph.h:
the size of every structure and the offset of every field, as the host C compiler lays them out
(a field added to one side only would shift every later field silently)."""
import ctypes
import os
import shutil
import subprocess

import pytest

from mps_cuda_amd import _capi

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PAIRS = {
    "mps_config": _capi.Config,
    "mps_physics": _capi.Physics,
    "mps_sph_material": _capi.SPHMaterial,
    "mps_mesh_material": _capi.MeshMaterial,
    "mps_step_stats": _capi.StepStats,
    "mps_kernel_stat": _capi.KernelStat,
    "mps_fence": _capi.Fence,
    "mps_tune_result": _capi.TuneResult,
    "mps_scene_range": _capi.SceneRange,
}


@pytest.mark.skipif(shutil.which("gcc") is None, reason="needs the host C compiler")
def test_ctypes_mirrors_match_the_header(tmp_path):
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "mps_sph.h"', "int main(void) {"]
    for cname, py in PAIRS.items():
        lines.append(f'    printf("{cname} size %zu\\n", sizeof({cname}));')
        for f in py._fields_:
            lines.append(f'    printf("{cname} {f[0]} %zu\\n", offsetof({cname}, {f[0]}));')
    lines += ["    return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", "-I", os.path.join(REPO, "include"), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    got = {}
    for line in out:
        if line:
            a, b, c = line.split()
            got[(a, b)] = int(c)
    for cname, py in PAIRS.items():
        assert got[(cname, "size")] == ctypes.sizeof(py), cname
        for f in py._fields_:
            assert got[(cname, f[0])] == getattr(py, f[0]).offset, (cname, f[0])
