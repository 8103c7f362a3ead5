"""Checkpoint / resume through the C ABI, shared by the CPU and GPU tests (test infrastructure).

A faithful resume carries the fluid state (positions, velocities, radii, masses in the current
particle order), the walls in their current (Z-sorted) order with their masses, velocities and
the never-cleared Akinci volumes, and the frame counter that sets the Z-sort cadence."""
from __future__ import annotations

import numpy as np


def snapshot(sim) -> dict:
    keys = ("FLUID_POS", "FLUID_VEL", "FLUID_RADIUS", "FLUID_MASS", "BND_POS", "BND_RADIUS", "BND_MASS", "BND_VEL",
            "BND_VOLUME")
    snap = {k: sim.get(k).copy() for k in keys}
    snap["frame"] = sim.frame
    return snap


def restore(sim, sc, snap):
    p, m = sc.physics, sc.material
    sim.set_physics(p["dt"], p["gravity"], p["box_min"], p["box_max"])
    sim.set_fluid_material(m["radius"], m["density"], m["viscosity"], m["xsph"], m["surface_tension"])
    sim.upload_boundary(snap["BND_POS"], snap["BND_RADIUS"], snap["BND_MASS"])
    sim.set("BND_VEL", snap["BND_VEL"])
    sim.set("BND_VOLUME", snap["BND_VOLUME"])
    sim.upload_fluid(snap["FLUID_POS"], snap["FLUID_RADIUS"], snap["FLUID_VEL"], snap["FLUID_MASS"])
    sim.frame = snap["frame"]


def run_resume(make, sc, split: int, total: int):
    """Steps `total` frames in one context and, from a snapshot at `split`, in a fresh one;
    returns (positions uninterrupted, positions resumed, loop counts of both runs)."""
    a = make()
    loops_a, loops_b = [], []
    snap = None
    for s in range(total):
        if s == split:
            snap = snapshot(a)
        st = a.step()
        if s >= split:
            loops_a.append((st.l_df, st.l_cd, st.l_cg, st.zsorted))
    b = make(load=False)
    restore(b, sc, snap)
    for _ in range(total - split):
        st = b.step()
        loops_b.append((st.l_df, st.l_cd, st.l_cg, st.zsorted))
    return np.asarray(a.get("FLUID_POS")), np.asarray(b.get("FLUID_POS")), loops_a, loops_b
