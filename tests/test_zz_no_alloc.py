"""mps_step_noalloc (the step that allocates no device memory) against mps_step.

Two contexts take the same scenes: one steps with mps_step, the other with mps_step_noalloc
after one ordinary sizing run. The results must be bit-identical: positions, velocities,
neighbour lists and loop counts, over steps that cross the frame-1 Z-sort, on two calls in a
row where the second scene is smaller. A context that no step has sized refuses the call
before any physics and then steps on normally from the untouched state.
"""
import sys

import numpy as np
import pytest

from mps_cuda_amd import HipSim, scenes
from mps_cuda_amd._capi import MPSError

pytestmark = pytest.mark.gpu


def say(m):
    print(m, flush=True)
    sys.stdout.flush()


def _same(a, b, what):
    for name in ("FLUID_POS", "FLUID_VEL", "FLUID_DENSITY", "NEI_FF_IDX", "NEI_FF", "BND_VOLUME"):
        assert np.array_equal(a.get(name), b.get(name)), (what, name)


def _counts(st):
    return (st.frame, st.zsorted, st.l_df, st.l_cg, st.l_cd, st.nnz_ff, st.nnz_fb)


def test_noalloc_step_equals_step_on_two_calls_second_smaller():
    big = scenes.dam_break_cpu(n_side=12)
    small = scenes.dam_break_cpu(n_side=8)
    cfg = scenes.make_config(big)
    ref, na = HipSim(cfg), HipSim(cfg)
    scenes.load_into(na, big)
    for _ in range(2):          # sizes every buffer, the frame-1 Z-sort's included
        na.step()
    for k, sc in enumerate((big, small)):
        scenes.load_into(ref, sc)
        scenes.load_into(na, sc)
        say(f"[noalloc] call {k}: N={sc.n}")
        for t in range(4):
            sa, sb = ref.step(), na.step_noalloc()
            assert _counts(sa) == _counts(sb), (k, t, _counts(sa), _counts(sb))
            _same(ref, na, (k, t))
        assert sb.frame == 3


def test_unsized_context_refuses_then_steps_normally():
    sc = scenes.dam_break_cpu(n_side=8)
    cfg = scenes.make_config(sc)
    ref, na = HipSim(cfg), HipSim(cfg)
    scenes.load_into(ref, sc)
    scenes.load_into(na, sc)
    say("[noalloc] unsized context")
    with pytest.raises(MPSError) as ei:
        na.step_noalloc()
    assert ei.value.code == 4 and "mps_step_noalloc" in str(ei.value)
    for t in range(3):
        sa, sb = ref.step(), na.step()
        assert _counts(sa) == _counts(sb), t
    _same(ref, na, "after refusal")
