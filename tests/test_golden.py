"""The oracle reproduces the committed golden vectors of tests/golden/ exactly (written by
tests/golden/make_golden.py): smoothing kernels, cell keys and six steps of two scenes in FP32
and FP64, with lists, densities and loop counts. A change to the oracle that moves any of them
shows up here; the GPU is checked against the FP32 vectors in test_gpu_golden.py."""
import os

import numpy as np
import pytest

from oracle.binding import set_threads
from tests.golden import make_golden as mg

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _load(name):
    with np.load(os.path.join(HERE, name), allow_pickle=False) as f:
        return {k: f[k] for k in f.files}


def _same(got, want):
    assert set(got) == set(want)
    for k in want:
        assert np.array_equal(np.asarray(got[k]), want[k]), k


def test_kernel_values():
    _same(mg.kernel_table(), _load("kernels.npz"))


def test_cell_keys():
    _same(mg.key_table(), _load("cell_keys.npz"))


@pytest.mark.parametrize("prec", [32, 64])
@pytest.mark.parametrize("name", sorted(mg.golden_scenes()))
def test_scene_steps(name, prec):
    set_threads(4)    # the oracle's results do not depend on its thread count
    _same(mg.run_steps(mg.golden_scenes()[name], prec), _load(f"steps_{name}_fp{prec}.npz"))
