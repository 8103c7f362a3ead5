"""ctypes handle over the CPU oracle (oracle/_build/libmps_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg. Mirrors mps_cuda_amd.HipSim method for method so that parity tests
drive both through the same calls. Inputs are passed in double; arrays come back in the
oracle's working precision (float32 or float64).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

from mps_cuda_amd import _capi as capi

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("MPS_ORACLE_LIB") or os.path.join(HERE, "_build", "libmps_oracle.so")

_P = C.c_void_p
_SIGS = {
    "orc_create": (C.c_int, [C.POINTER(capi.Config), C.c_int, C.POINTER(_P)]),
    "orc_destroy": (None, [_P]),
    "orc_last_error": (C.c_char_p, [_P]),
    "orc_precision": (C.c_int, [_P]),
    "orc_set_threads": (C.c_int, [C.c_int]),
    "orc_set_physics": (C.c_int, [_P, C.POINTER(capi.Physics)]),
    "orc_set_fluid_material": (C.c_int, [_P, C.POINTER(capi.SPHMaterial)]),
    "orc_upload_fluid": (C.c_int, [_P, C.c_size_t, _P, _P, _P, _P]),
    "orc_upload_boundary": (C.c_int, [_P, C.c_size_t, _P, _P, _P]),
    "orc_sample_boundary_mesh": (C.c_int, [_P, _P, C.c_size_t, _P, C.c_size_t, _P, _P, C.POINTER(capi.MeshMaterial)]),
    "orc_set_boundary_objects": (C.c_int, [_P, C.c_uint32, _P]),
    "orc_world_create": (C.c_int, [C.POINTER(capi.Config), C.c_int, C.c_uint32, C.POINTER(_P)]),
    "orc_world_destroy": (None, [_P]),
    "orc_world_phase": (_P, [_P, C.c_uint32]),
    "orc_world_step": (C.c_int, [_P, _P]),
    "orc_world_run_phase": (C.c_int, [_P, C.c_int32, _P]),
    "orc_resample_boundary_mesh": (C.c_int, [_P, _P, C.c_size_t, C.POINTER(capi.MeshMaterial), C.POINTER(C.c_size_t)]),
    "orc_step": (C.c_int, [_P, C.POINTER(capi.StepStats)]),
    "orc_run_phase": (C.c_int, [_P, C.c_int32, C.POINTER(capi.StepStats)]),
    "orc_array_info": (C.c_int, [_P, C.c_int32, C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)]),
    "orc_get_array": (C.c_int, [_P, C.c_int32, _P, C.c_size_t]),
    "orc_set_array": (C.c_int, [_P, C.c_int32, _P, C.c_size_t]),
    "orc_get_frame": (C.c_int, [_P, C.POINTER(C.c_uint32)]),
    "orc_set_frame": (C.c_int, [_P, C.c_uint32]),
    "orc_cg_operator": (C.c_int, [_P, _P, _P, C.c_double]),
    "orc_kernel_eval": (C.c_int, [C.c_int, C.c_int, C.c_size_t, _P, C.c_double, _P, _P, _P, _P]),
    "orc_cell_keys": (C.c_int, [C.c_int, C.c_size_t, _P, C.c_double, _P, C.c_int, _P]),
    "orc_mesh_rtri": (C.c_int, [C.c_size_t, _P, C.c_size_t, _P]),
    "orc_mesh_sample_counts": (C.c_int, [_P, C.c_size_t, _P, C.c_size_t, C.c_double, _P]),
}

_lib = None


def build() -> str:
    r = subprocess.run(["make", "-C", HERE], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"oracle build failed:\n{r.stdout}\n{r.stderr}")
    return LIB_PATH


def load() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        lib = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def set_threads(n: int) -> int:
    return load().orc_set_threads(int(n))


class OracleSim:
    """CPU oracle context with the same methods as mps_cuda_amd.HipSim."""

    def __init__(self, cfg: capi.Config, precision: int = 32):
        self.lib = load()
        self.cfg = cfg
        self.precision = precision
        self.real = np.float32 if precision == 32 else np.float64
        h = C.c_void_p()
        rc = self.lib.orc_create(C.byref(cfg), precision, C.byref(h))
        if rc:
            raise capi.MPSError(rc, "orc_create failed")
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            self.lib.orc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, rc):
        if rc:
            raise capi.MPSError(rc, self.lib.orc_last_error(self.h).decode())

    def set_physics(self, dt, gravity, box_min, box_max):
        p = capi.Physics(dt, (C.c_double * 3)(*gravity), (C.c_double * 3)(*box_min), (C.c_double * 3)(*box_max))
        self._chk(self.lib.orc_set_physics(self.h, C.byref(p)))

    def set_fluid_material(self, radius, density, viscosity, xsph, surface_tension):
        m = capi.SPHMaterial(radius, density, viscosity, xsph, surface_tension)
        self._chk(self.lib.orc_set_fluid_material(self.h, C.byref(m)))

    def upload_fluid(self, pos, radius, vel=None, mass=None):
        pos = np.ascontiguousarray(pos, np.float64).reshape(-1, 3)
        n = pos.shape[0]
        radius = np.ascontiguousarray(np.broadcast_to(np.asarray(radius, np.float64), (n,)))
        vel = None if vel is None else np.ascontiguousarray(vel, np.float64).reshape(n, 3)
        mass = None if mass is None else np.ascontiguousarray(np.broadcast_to(np.asarray(mass, np.float64), (n,)))
        self._chk(self.lib.orc_upload_fluid(self.h, n, _ptr(pos), _ptr(vel), _ptr(radius), _ptr(mass)))

    def upload_boundary(self, pos, radius, mass=None):
        pos = np.ascontiguousarray(pos, np.float64).reshape(-1, 3)
        n = pos.shape[0]
        radius = np.ascontiguousarray(np.broadcast_to(np.asarray(radius, np.float64), (n,)))
        mass = None if mass is None else np.ascontiguousarray(mass, np.float64)
        self._chk(self.lib.orc_upload_boundary(self.h, n, _ptr(pos), _ptr(radius), _ptr(mass)))

    def sample_boundary_mesh(self, verts, tris, radius, density=1.0, center=None, size=None):
        verts = np.ascontiguousarray(verts, np.float64).reshape(-1, 3)
        tris = np.ascontiguousarray(tris, np.uint32).reshape(-1, 3)
        c = None if center is None else np.ascontiguousarray(center, np.float64)
        s = None if size is None else np.ascontiguousarray(size, np.float64)
        m = capi.MeshMaterial(radius, density)
        self._chk(self.lib.orc_sample_boundary_mesh(self.h, _ptr(verts), verts.shape[0], _ptr(tris), tris.shape[0],
                                                    _ptr(c), _ptr(s), C.byref(m)))

    def set_boundary_objects(self, counts):
        c = np.ascontiguousarray(counts, np.uint32)
        self._chk(self.lib.orc_set_boundary_objects(self.h, len(c), _ptr(c)))

    def resample_boundary_mesh(self, verts, radius) -> int:
        verts = np.ascontiguousarray(verts, np.float64).reshape(-1, 3)
        m = capi.MeshMaterial(radius, 1.0)
        k = C.c_size_t()
        self._chk(self.lib.orc_resample_boundary_mesh(self.h, _ptr(verts), verts.shape[0], C.byref(m), C.byref(k)))
        return int(k.value)

    def step(self) -> capi.StepStats:
        st = capi.StepStats()
        self._chk(self.lib.orc_step(self.h, C.byref(st)))
        return st

    def run_phase(self, phase) -> capi.StepStats:
        st = capi.StepStats()
        ph = capi.PHASE[phase] if isinstance(phase, str) else int(phase)
        self._chk(self.lib.orc_run_phase(self.h, ph, C.byref(st)))
        return st

    def synchronize(self):
        pass

    def _dtype(self, kind):
        return {"real": self.real, "f32": np.float32, "u32": np.uint32}[kind]

    def get(self, name: str) -> np.ndarray:
        aid, kind, w = capi.ARRAYS[name]
        n, eb = C.c_size_t(), C.c_size_t()
        self._chk(self.lib.orc_array_info(self.h, aid, C.byref(n), C.byref(eb)))
        out = np.empty((n.value, w) if w > 1 else (n.value,), self._dtype(kind))
        assert out.nbytes == n.value * eb.value, (name, out.nbytes, n.value, eb.value)
        self._chk(self.lib.orc_get_array(self.h, aid, _ptr(out), out.nbytes))
        return out

    def set(self, name: str, arr):
        aid, kind, w = capi.ARRAYS[name]
        a = np.ascontiguousarray(arr, self._dtype(kind))
        self._chk(self.lib.orc_set_array(self.h, aid, _ptr(a), a.nbytes))

    def cg_operator(self, p, f: float):
        p = np.ascontiguousarray(p, self.real).reshape(-1, 3)
        out = np.empty_like(p)
        self._chk(self.lib.orc_cg_operator(self.h, _ptr(p), _ptr(out), float(f)))
        return out

    @property
    def frame(self) -> int:
        f = C.c_uint32()
        self._chk(self.lib.orc_get_frame(self.h, C.byref(f)))
        return f.value

    @frame.setter
    def frame(self, v: int):
        self._chk(self.lib.orc_set_frame(self.h, int(v)))


class _Phase(OracleSim):
    """One fluid object of an OracleWorld (the world owns the context)."""

    def __init__(self, world, handle, cfg, precision):  # noqa: super().__init__ not called on purpose
        self.lib = world.lib
        self.cfg = cfg
        self.precision = precision
        self.real = np.float32 if precision == 32 else np.float64
        self.h = handle
        self._world = world

    def close(self):
        self.h = None


class OracleWorld:
    """Coupled fluid objects in the oracle: phase k is loaded like an OracleSim
    (world.phase(k).set_physics(...), ...); step() advances all of them together."""

    def __init__(self, cfg: capi.Config, nphase: int, precision: int = 32):
        self.lib = load()
        self.n = nphase
        h = C.c_void_p()
        rc = self.lib.orc_world_create(C.byref(cfg), precision, nphase, C.byref(h))
        if rc:
            raise capi.MPSError(rc, "orc_world_create failed")
        self.h = h
        self.phases = [_Phase(self, self.lib.orc_world_phase(h, k), cfg, precision) for k in range(nphase)]

    def phase(self, k: int) -> OracleSim:
        return self.phases[k]

    def close(self):
        if getattr(self, "h", None):
            for p in self.phases:
                p.close()
            self.lib.orc_world_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _stats(self, fn, *args):
        st = (capi.StepStats * self.n)()
        rc = fn(self.h, *args, C.cast(st, C.c_void_p))
        if rc:
            raise capi.MPSError(rc, "oracle world call failed")
        return list(st)

    def step(self) -> list:
        return self._stats(self.lib.orc_world_step)

    def run_phase(self, phase) -> list:
        ph = capi.PHASE[phase] if isinstance(phase, str) else int(phase)
        return self._stats(self.lib.orc_world_run_phase, ph)


def kernel_eval(precision: int, kernel: int, dist, invh: float):
    d = np.ascontiguousarray(dist, np.float64)
    w, g, stw, stg = (np.empty_like(d) for _ in range(4))
    load().orc_kernel_eval(precision, kernel, d.size, _ptr(d), float(invh), _ptr(w), _ptr(g), _ptr(stw), _ptr(stg))
    return w, g, stw, stg


def cell_keys(precision: int, pos, cell: float, hash_size, morton: bool = False):
    p = np.ascontiguousarray(pos, np.float64).reshape(-1, 3)
    H = np.ascontiguousarray(hash_size, np.uint32)
    out = np.empty(p.shape[0], np.uint32)
    load().orc_cell_keys(precision, p.shape[0], _ptr(p), float(cell), _ptr(H), int(morton), _ptr(out))
    return out


def mesh_rtri(nv: int, tris):
    t = np.ascontiguousarray(tris, np.uint32).reshape(-1, 3)
    out = np.empty(t.shape[0], np.uint32)
    load().orc_mesh_rtri(nv, _ptr(t), t.shape[0], _ptr(out))
    return out


def mesh_sample_counts(verts, tris, radius: float):
    v = np.ascontiguousarray(verts, np.float64).reshape(-1, 3)
    t = np.ascontiguousarray(tris, np.uint32).reshape(-1, 3)
    out = np.empty(t.shape[0], np.uint32)
    load().orc_mesh_sample_counts(_ptr(v), v.shape[0], _ptr(t), t.shape[0], float(radius), _ptr(out))
    return out
