"""The C ABI library loads on a host without a GPU and exports every entry point that
include/mps_sph.h declares; the product package never imports the oracle."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    src = open(os.path.join(ROOT, "include", "mps_sph.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(mps_[a-z_0-9]+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    from mps_cuda_amd import _capi
    lib = ctypes.CDLL(_capi.LIB_PATH)
    names = _declared()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    # the Python binding covers the same set
    assert set(names) <= set(_capi.exported_symbols())


def test_config_defaults_are_the_reference_constants():
    from mps_cuda_amd import default_config
    c = default_config()
    assert (c.kernel, c.visc_solver, c.faithful_quirks) == (0, 0, 1)
    assert list(c.hash_size) == [64, 64, 64]
    assert (c.df_tol, c.df_min_iter, c.cd_tol, c.cd_min_iter, c.max_iter) == (1e-3, 1, 1e-4, 2, 100)
    assert (c.cg_tol, c.cg_max_iter, c.cg_factor, c.zsort_interval) == (1e-2, 100, 0.01, 50)


def test_product_does_not_use_the_oracle():
    pkg = os.path.join(ROOT, "mps_cuda_amd")
    pat = re.compile(r"^\s*(from\s+oracle|import\s+oracle)|sph_oracle|libmps_oracle|orc_[a-z]", re.M)
    for dirpath, _, files in os.walk(pkg):
        for f in files:
            if f.endswith((".py", ".cpp", ".hip", ".hpp", ".h")):
                text = open(os.path.join(dirpath, f)).read()
                assert not pat.search(text), f


def test_invalid_config_rejected():
    from mps_cuda_amd import MPSError, HipSim, default_config
    c = default_config()
    c.hash_size[0] = 48  # not a power of two
    with pytest.raises(MPSError):
        HipSim(c)


def test_missing_library_fails_loudly(tmp_path):
    """No fallback path: a missing libmps_sph.so is an error at load time."""
    from mps_cuda_amd import _capi
    with pytest.raises(RuntimeError, match="is missing"):
        _capi.load(str(tmp_path / "libmps_sph.so"))
