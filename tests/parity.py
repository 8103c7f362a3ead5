"""Shared helpers for GPU-vs-oracle parity tests (test infrastructure)."""
from __future__ import annotations

import numpy as np

from mps_cuda_amd import scenes
from oracle.binding import OracleSim

PHASE_ORDER = ["HASH", "NEIGHBORS", "BOUNDARY_VOLUME", "DENSITY", "ALPHA", "GRAVITY",
               "DIVERGENCE", "VISCOSITY", "CONSTANT_DENSITY", "ADVECT"]

INT_ARRAYS = {
    "HASH": ["FHASH_KEY", "FHASH_ID", "FHASH_START", "FHASH_END", "BHASH_KEY", "BHASH_ID", "BHASH_START", "BHASH_END"],
    "NEIGHBORS": ["NEI_FF_IDX", "NEI_FF", "NEI_FB_IDX", "NEI_FB", "NEI_BB_IDX", "NEI_BB"],
}
REAL_ARRAYS = {
    "BOUNDARY_VOLUME": ["BND_VOLUME"],
    "DENSITY": ["FLUID_DENSITY", "FLUID_COLOR"],
    "ALPHA": ["FLUID_ALPHA"],
    "GRAVITY": ["FLUID_VEL"],
    "DIVERGENCE": ["FLUID_VEL", "FLUID_KAPPA", "FLUID_DELTA"],
    "VISCOSITY": ["FLUID_VEL", "FLUID_FACTOR_ST", "FLUID_CG_B"],
    "CONSTANT_DENSITY": ["FLUID_VEL", "FLUID_KAPPA", "FLUID_DELTA"],
    "ADVECT": ["FLUID_POS", "FLUID_VEL"],
}


def rel_err(a, b) -> float:
    """Norm-wise relative error ||a-b|| / max(||b||, tiny)."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    den = max(float(np.linalg.norm(b)), 1e-30)
    return float(np.linalg.norm(a - b)) / den


def max_rel_elem(a, b, floor: float) -> float:
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.max(np.abs(a - b) / np.maximum(np.abs(b), floor))) if a.size else 0.0


def make_pair(scene, precision: int = 32):
    from mps_cuda_amd import HipSim
    cfg = scenes.make_config(scene)
    g = HipSim(cfg)
    o = OracleSim(scenes.make_config(scene), precision)
    scenes.load_into(g, scene)
    scenes.load_into(o, scene)
    return g, o
