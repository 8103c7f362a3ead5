"""The OBJ reader of MeshObject::LoadMesh (MPSMeshObject.cu:58-103) and the host side of the
C++ classes in include/mps_system.hpp, without a GPU.

The reader's rules are the reference's token loop: a token longer than one character is
skipped, `v` takes three reals, `f` three 1-based indices. Files below are written by hand so
that each rule (and each error the reference would loop or exit on) is exercised once."""
import os
import subprocess

import numpy as np
import pytest

from mps_cuda_amd import MPSError, build, read_obj, scenes

HERE = os.path.dirname(os.path.abspath(__file__))
CUBE = os.path.join(HERE, "golden", "cube.obj")


def _obj(tmp_path, text, name="m.obj"):
    p = tmp_path / name
    p.write_text(text)
    return str(p)


def test_golden_cube_reads_as_the_scene_cube():
    v, t = read_obj(CUBE)
    assert np.array_equal(v, scenes.CUBE_VERTS) and np.array_equal(t, scenes.CUBE_TRIS)


def test_long_tokens_are_skipped_and_one_character_words_are_not(tmp_path):
    text = ("# exported mesh, version 2\n"
            "o tri\nvn 0 0 1\nvt 0.5 0.5\nusemtl stone\n"
            "v 0 0 0\nv 1.5 0 0\nv 0 2.5e-1 0\n"
            "# a comment whose one-character word v 9 8 7 becomes a vertex, as in the reference\n"
            "s off\nf 1 2 3\n")
    v, t = read_obj(_obj(tmp_path, text))
    # `s` is one character but neither v nor f; `off` is skipped; the commented `v 9 8 7` counts
    assert v.tolist() == [[0, 0, 0], [1.5, 0, 0], [0, 0.25, 0], [9, 8, 7]]
    assert t.tolist() == [[0, 1, 2]]


def test_quad_keeps_its_first_triangle_and_slash_indices_use_the_vertex(tmp_path):
    text = "v 0 0 0\nv 1 0 0\nv 1 1 0\nv 0 1 0\nf 1 2 3 4\nf 1/1/1 3/3/3 4//4\n"
    _, t = read_obj(_obj(tmp_path, text))
    # the quad's 4th index is a one-character token that is neither v nor f: dropped
    assert t.tolist() == [[0, 1, 2], [0, 2, 3]]


def test_tokens_may_span_lines(tmp_path):
    v, t = read_obj(_obj(tmp_path, "v\n1\n2\n3 v 4 5 6 v 7 8 9 f\n1 2\n3\n"))
    assert v.tolist() == [[1, 2, 3], [4, 5, 6], [7, 8, 9]] and t.tolist() == [[0, 1, 2]]


@pytest.mark.parametrize("text", [
    "v 0 0 0\nv 1 0 0\nv 0 1 0\nf 0 1 2\n",        # index 0 (1-based file)
    "v 0 0 0\nv 1 0 0\nv 0 1 0\nf 1 2 4\n",        # past the vertex count
    "v 0 0 x\nv 1 0 0\nv 0 1 0\nf 1 2 3\n",        # malformed real
    "v 0 0 0\nv 1 0 0\nv 0 1 0\nf 1 2 -3\n",       # negative index
    "v 0 0 0\nv 1 0 0\nv 0 1 0\n",                 # no faces
    "f 1 2 3\n",                                   # no vertices
    "v 0 0 0\nv 1 0 0\nv 0 1 0\nf 1 2\n",          # face cut short at the end
])
def test_malformed_files_are_errors(tmp_path, text):
    with pytest.raises(MPSError):
        read_obj(_obj(tmp_path, text))


def test_missing_file_is_an_error(tmp_path):
    with pytest.raises(MPSError):
        read_obj(str(tmp_path / "none.obj"))


def _host_lines():
    exe = build.FACADE_BIN
    if not os.path.exists(exe):
        pytest.skip("facade driver not built (python -m mps_cuda_amd.build)")
    out = subprocess.run([exe, "host", CUBE], capture_output=True, text=True, timeout=60, check=True).stdout
    return {ln.split()[0]: ln.split()[1:] for ln in out.strip().splitlines()}


def test_cpp_material_classes_match_the_reference_constructors():
    d = _host_lines()
    r, rho, vol, mass, mu, xsph, sigma = map(float, d["sph_default"])
    # SPHMaterial::SPHMaterial (MPSSPHMaterial.cpp:4-11): SetParam(0.01, 1.0), 0.025, 0.1, 100
    assert (r, rho, mu, xsph, sigma) == (0.01, 1.0, 0.025, 0.1, 100.0)
    # V = R^3 / 48 * 3.141592 is (4/3) pi (R/4)^3 with pi cut to 3.141592: check against that
    assert vol == pytest.approx(4.0 / 3.0 * 3.141592 * (0.01 / 4) ** 3, rel=1e-15) and mass == vol
    r, vol, mass = map(float, d["sph_radius_0.4"])
    assert r == 0.4 and vol == pytest.approx(4.0 / 3.0 * 3.141592 * 0.1 ** 3, rel=1e-15) and mass == vol
    r, vol, mass = map(float, d["sph_param_0.04_1000"])
    assert vol == pytest.approx(4.0 / 3.0 * 3.141592 * 0.01 ** 3, rel=1e-15) and mass == pytest.approx(1000 * vol, rel=1e-15)
    r, rho, mass = map(float, d["mesh_default"])  # MeshMaterial::MeshMaterial: SetParam(0.1, 1.0)
    assert (r, rho) == (0.1, 1.0) and mass == pytest.approx(4.0 / 3.0 * 3.141592 * 0.025 ** 3, rel=1e-15)


def test_cpp_load_mesh_keeps_the_file_and_the_fit():
    d = _host_lines()
    assert d["mesh_counts"] == ["8", "12"]
    assert np.array_equal(np.array(d["mesh_verts"], float).reshape(-1, 3), scenes.CUBE_VERTS)
    assert np.array_equal(np.array(d["mesh_faces"], np.uint32).reshape(-1, 3), scenes.CUBE_TRIS)
    assert list(map(float, d["mesh_fit"])) == [1, 2, 3, 4, 5, 6, 2.5]
    assert d["missing_mesh"] == ["status", "1"]
