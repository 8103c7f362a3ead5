"""Writes the golden vectors of tests/golden/ from the oracle (run from the repository root:
python tests/golden/make_golden.py). Test infrastructure: the fixtures pin the oracle against
later changes and give the GPU a fixed target. Every file is a plain .npz (no pickled objects).

- kernels.npz: W, G, STW, STG of both smoothing-kernel families at 41 distances, FP32 and FP64;
- cell_keys.npz: linear and Morton cell keys of seeded positions (negative coordinates,
  hash wrap-around) in FP32 and FP64;
- steps_<scene>_fp<32|64>.npz: 6 steps of a scene: per-step positions, velocities and loop
  counts, plus the fluid-fluid and fluid-wall lists and densities of the last step.
"""
from __future__ import annotations

import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from mps_cuda_amd import _capi as capi  # noqa: E402
from mps_cuda_amd import scenes  # noqa: E402
from oracle.binding import OracleSim, cell_keys, kernel_eval, set_threads  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))
STEPS = 6


def golden_scenes():
    return {"surface_tension_3": scenes.surface_tension_test_scene(3), "dam_break_6": scenes.dam_break_cpu(n_side=6)}


def kernel_table():
    q = np.linspace(0.0, 1.0, 41) * 0.04
    out = {"dist": q, "invh": np.array(1.0 / 0.04)}
    for prec in (32, 64):
        for fam, k in (("spiky", capi.KERNEL_POLY6_SPIKY), ("cubic", capi.KERNEL_CUBIC)):
            w, g, stw, stg = kernel_eval(prec, k, q, 1.0 / 0.04)
            for name, a in (("W", w), ("G", g), ("STW", stw), ("STG", stg)):
                out[f"{name}_{fam}_fp{prec}"] = a
    return out


def key_table():
    rng = np.random.default_rng(7)
    pos = rng.uniform(-1.3, 1.3, size=(200, 3))
    out = {"pos": pos, "cell": np.array(0.04), "hash": np.array([16, 32, 64], np.uint32)}
    for prec in (32, 64):
        out[f"linear_fp{prec}"] = cell_keys(prec, pos, 0.04, [16, 32, 64], False)
        out[f"morton_fp{prec}"] = cell_keys(prec, pos, 0.04, [16, 32, 64], True)
    return out


def run_steps(sc, prec: int):
    o = OracleSim(scenes.make_config(sc), prec)
    scenes.load_into(o, sc)
    pos, vel, loops = [], [], []
    for _ in range(STEPS):
        st = o.step()
        pos.append(o.get("FLUID_POS"))
        vel.append(o.get("FLUID_VEL"))
        loops.append([st.l_df, st.l_cd, st.l_cg, st.zsorted])
    return dict(pos=np.stack(pos), vel=np.stack(vel), loops=np.array(loops, np.uint32),
                nei_ff_idx=o.get("NEI_FF_IDX"), nei_ff=o.get("NEI_FF"), nei_fb_idx=o.get("NEI_FB_IDX"),
                nei_fb=o.get("NEI_FB"), density=o.get("FLUID_DENSITY"))


def main():
    set_threads(1)
    np.savez(os.path.join(HERE, "kernels.npz"), **kernel_table())
    np.savez(os.path.join(HERE, "cell_keys.npz"), **key_table())
    for name, sc in golden_scenes().items():
        for prec in (32, 64):
            np.savez_compressed(os.path.join(HERE, f"steps_{name}_fp{prec}.npz"), **run_steps(sc, prec))
    print("golden vectors written to", HERE)


if __name__ == "__main__":
    main()
