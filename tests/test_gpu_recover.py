"""Batched parameter recovery (mps_cuda_amd.recover) on a small batch: four scenes of 2k-4k
particles, 40 steps observed every 10. With the hidden initial velocities given, the Gauss-Newton
solve in (log viscosity, log surface tension) must bring both coefficients of every scene close
to their hidden values; with the initial velocities unknown as well, the trajectory loss and the
initial velocities' error must fall."""
import sys

import numpy as np
import pytest

from mps_cuda_amd import recover

pytestmark = pytest.mark.gpu
SMALL = dict(n_scenes=4, steps=40, seed=7, n_min=2048, n_max=4096, every=10, obs_every=10)


def say(m):
    print(m, flush=True)
    sys.stdout.flush()


def test_coefficients_recovered_with_known_initial_velocity():
    say("[recover] 4 scenes, 40 steps, hidden v0 given, 6 Gauss-Newton iterations")
    r = recover.run(iters=6, recover_v0=False, **SMALL)
    mu, sig = r["viscosity"], r["surface_tension"]
    say(f"[recover] viscosity median {mu['rel_err_median']:.2e} max {mu['rel_err_max']:.2e}; "
        f"surface tension median {sig['rel_err_median']:.2e} max {sig['rel_err_max']:.2e}")
    assert r["history"][-1]["loss"] < 1e-3 * r["history"][0]["loss"]
    assert mu["rel_err_max"] < 1e-2
    assert sig["rel_err_median"] < 0.05 and sig["rel_err_max"] < 0.25


def test_joint_recovery_reduces_loss_and_velocity_error():
    say("[recover] 4 scenes, 40 steps, v0 unknown, 3 Gauss-Newton iterations of 3 CGLS steps")
    r = recover.run(iters=3, inner=3, recover_v0=True, **SMALL)
    h = r["history"]
    iv = r["initial_velocity"]
    say(f"[recover] loss {h[0]['loss']:.3e} -> {h[-1]['loss']:.3e}; v0 error {iv['rel_err_median']:.3f}, "
        f"response error {iv['first_step_response_err_median']:.3f}")
    assert all(b["loss"] <= a["loss"] for a, b in zip(h, h[1:]))
    assert h[-1]["loss"] < 0.5 * h[0]["loss"]
    assert iv["rel_err_median"] < 0.95 and iv["first_step_response_err_median"] < 0.9
    assert np.isfinite(r["viscosity"]["recovered"]).all() and np.isfinite(r["surface_tension"]["recovered"]).all()
