"""ctypes mirror of include/mps_sph.h and a thin Python handle over the gfx950 library.

The library is loaded from the package directory (libmps_sph.so, built by
mps_cuda_amd.build). There is no fallback: if the library is missing, loading raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# MPS_SPH_LIB names another build of the same library (e.g. one compiled with other tile
# parameters); by default the in-tree build is loaded.
LIB_PATH = os.environ.get("MPS_SPH_LIB") or os.path.join(HERE, "libmps_sph.so")

MPS_OK, MPS_ERR_INVALID, MPS_ERR_HIP, MPS_ERR_NOMEM, MPS_ERR_STATE, MPS_ERR_NONFINITE = range(6)
MAX_ITER_LOG = 128

KERNEL_POLY6_SPIKY, KERNEL_CUBIC = 0, 1
VISC_CG, VISC_JACOBI, VISC_NONE = 0, 1, 2

PHASE = dict(
    ZSORT=1, HASH=2, NEIGHBORS=3, BOUNDARY_VOLUME=4, DENSITY=5, ALPHA=6, GRAVITY=7,
    DIVERGENCE=8, VISCOSITY=9, CONSTANT_DENSITY=10, ADVECT=11,
)

# array id -> (kind, width); kind "real" = float32 in the HIP library, oracle precision there
ARRAYS = {
    "FLUID_POS": (1, "real", 3), "FLUID_VEL": (2, "real", 3), "FLUID_RADIUS": (3, "real", 1),
    "FLUID_MASS": (4, "real", 1), "FLUID_COLOR": (5, "f32", 4), "FLUID_DENSITY": (6, "real", 1),
    "FLUID_ALPHA": (7, "real", 1), "FLUID_KAPPA": (8, "real", 1), "FLUID_DELTA": (9, "real", 1),
    "FLUID_FACTOR_ST": (10, "real", 1), "FLUID_CG_B": (11, "real", 3),
    "BND_POS": (20, "real", 3), "BND_RADIUS": (21, "real", 1), "BND_VOLUME": (22, "real", 1),
    "BND_MASS": (23, "real", 1), "BND_VEL": (24, "real", 3), "BND_FACE_ID": (25, "u32", 1),
    "BND_BCC": (26, "real", 2), "BND_OBJECT": (27, "u32", 1),
    "FHASH_KEY": (30, "u32", 1), "FHASH_ID": (31, "u32", 1), "FHASH_START": (32, "u32", 1),
    "FHASH_END": (33, "u32", 1), "BHASH_KEY": (34, "u32", 1), "BHASH_ID": (35, "u32", 1),
    "BHASH_START": (36, "u32", 1), "BHASH_END": (37, "u32", 1),
    "NEI_FF_IDX": (40, "u32", 1), "NEI_FF": (41, "u32", 1), "NEI_FB_IDX": (42, "u32", 1),
    "NEI_FB": (43, "u32", 1), "NEI_BB_IDX": (44, "u32", 1), "NEI_BB": (45, "u32", 1),
}


class Config(C.Structure):
    _fields_ = [
        ("kernel", C.c_int32), ("visc_solver", C.c_int32), ("faithful_quirks", C.c_int32),
        ("device", C.c_int32), ("hash_size", C.c_uint32 * 3), ("cell_size", C.c_double),
        ("df_tol", C.c_double), ("df_min_iter", C.c_uint32), ("cd_tol", C.c_double),
        ("cd_min_iter", C.c_uint32), ("max_iter", C.c_uint32), ("cg_tol", C.c_double),
        ("cg_max_iter", C.c_uint32), ("cg_factor", C.c_double), ("jacobi_max_iter", C.c_uint32),
        ("jacobi_tol", C.c_double), ("zsort_interval", C.c_uint32), ("nei_capacity", C.c_uint32),
        ("host_sync_every", C.c_uint32), ("adj_cg_max_iter", C.c_uint32), ("use_graphs", C.c_uint32),
        ("halo_limit", C.c_uint32), ("couple_scenes", C.c_uint32), ("cg_stream_coeffs", C.c_uint32),
        ("persistent_solver", C.c_uint32), ("persistent_spin", C.c_uint32),
        ("adj_cg_tol", C.c_double), ("cg_fold_in_kernel", C.c_uint32), ("cg_halo_depth", C.c_uint32),
    ]


class Physics(C.Structure):
    _fields_ = [("dt", C.c_double), ("gravity", C.c_double * 3), ("box_min", C.c_double * 3),
                ("box_max", C.c_double * 3)]


class SPHMaterial(C.Structure):
    _fields_ = [("radius", C.c_double), ("density", C.c_double), ("viscosity", C.c_double),
                ("xsph", C.c_double), ("surface_tension", C.c_double)]


class MeshMaterial(C.Structure):
    _fields_ = [("radius", C.c_double), ("density", C.c_double)]


class StepStats(C.Structure):
    _fields_ = [
        ("frame", C.c_uint32), ("zsorted", C.c_uint32), ("l_df", C.c_uint32), ("l_cd", C.c_uint32),
        ("l_cg", C.c_uint32), ("hit_cap_df", C.c_uint32), ("hit_cap_cd", C.c_uint32),
        ("hit_cap_cg", C.c_uint32), ("nonfinite", C.c_uint32), ("n_fluid", C.c_uint32),
        ("n_boundary", C.c_uint32), ("nei_regrow", C.c_uint32), ("nnz_ff", C.c_uint64),
        ("nnz_fb", C.c_uint64), ("nnz_bb", C.c_uint64), ("err_df", C.c_double * MAX_ITER_LOG),
        ("err_cd", C.c_double * MAX_ITER_LOG), ("rr_cg", C.c_double * MAX_ITER_LOG),
        ("halo_max", C.c_uint32), ("halo_spill", C.c_uint32), ("halo_blocks", C.c_uint32),
        ("cg_persistent", C.c_uint32),
    ]

    def as_dict(self) -> dict:
        d = {k: getattr(self, k) for k, _ in self._fields_ if not k.startswith(("err_", "rr_"))}
        d["err_df"] = list(self.err_df[: max(0, min(self.l_df, MAX_ITER_LOG))])
        d["err_cd"] = list(self.err_cd[: max(0, min(self.l_cd, MAX_ITER_LOG))])
        d["rr_cg"] = list(self.rr_cg[: max(0, min(self.l_cg + 1, MAX_ITER_LOG))])
        return d


class KernelStat(C.Structure):
    _fields_ = [("name", C.c_char * 48), ("launches", C.c_uint64), ("useful", C.c_uint64), ("total_ms", C.c_double)]


class Fence(C.Structure):
    _fields_ = [("event", C.c_void_p), ("frame", C.c_uint64)]


class TuneResult(C.Structure):
    _fields_ = [("use_graphs", C.c_uint32), ("host_sync_every", C.c_uint32), ("n_tried", C.c_uint32),
                ("pad", C.c_uint32), ("ms_per_step", C.c_double * 8), ("cg_stream_coeffs", C.c_uint32),
                ("persistent_solver", C.c_uint32), ("ms_coeffs", C.c_double * 2), ("ms_persistent", C.c_double * 2),
                ("cg_fold_in_kernel", C.c_uint32), ("pad2", C.c_uint32), ("ms_fold", C.c_double * 2)]


class SceneRange(C.Structure):
    _fields_ = [("fluid_offset", C.c_size_t), ("n_fluid", C.c_size_t), ("boundary_offset", C.c_size_t),
                ("n_boundary", C.c_size_t)]


_P = C.c_void_p
_SIGS = {
    "mps_config_init": (C.c_int, [C.POINTER(Config)]),
    "mps_create": (C.c_int, [C.POINTER(Config), C.POINTER(_P)]),
    "mps_destroy": (None, [_P]),
    "mps_last_error": (C.c_char_p, [_P]),
    "mps_version": (C.c_char_p, []),
    "mps_set_physics": (C.c_int, [_P, C.POINTER(Physics)]),
    "mps_set_fluid_material": (C.c_int, [_P, C.POINTER(SPHMaterial)]),
    "mps_upload_fluid": (C.c_int, [_P, C.c_size_t, _P, _P, _P, _P]),
    "mps_upload_boundary": (C.c_int, [_P, C.c_size_t, _P, _P, _P]),
    "mps_sample_boundary_mesh": (C.c_int, [_P, _P, C.c_size_t, _P, C.c_size_t, _P, _P, C.POINTER(MeshMaterial)]),
    "mps_set_boundary_objects": (C.c_int, [_P, C.c_uint32, _P]),
    "mps_resample_boundary_mesh": (C.c_int, [_P, _P, C.c_size_t, C.POINTER(MeshMaterial), C.POINTER(C.c_size_t)]),
    "mps_read_obj": (C.c_int, [_P, C.c_char_p, _P, C.POINTER(C.c_size_t), _P, C.POINTER(C.c_size_t)]),
    "mps_sample_boundary_obj": (C.c_int, [_P, C.c_char_p, _P, _P, C.POINTER(MeshMaterial)]),
    "mps_step": (C.c_int, [_P, C.POINTER(StepStats)]),
    "mps_step_noalloc": (C.c_int, [_P, C.POINTER(StepStats)]),
    "mps_run_phase": (C.c_int, [_P, C.c_int32, C.POINTER(StepStats)]),
    "mps_synchronize": (C.c_int, [_P]),
    "mps_array_info": (C.c_int, [_P, C.c_int32, C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)]),
    "mps_get_array": (C.c_int, [_P, C.c_int32, _P, C.c_size_t]),
    "mps_set_array": (C.c_int, [_P, C.c_int32, _P, C.c_size_t]),
    "mps_get_frame": (C.c_int, [_P, C.POINTER(C.c_uint32)]),
    "mps_set_frame": (C.c_int, [_P, C.c_uint32]),
    "mps_readback_async": (C.c_int, [_P, _P, _P, _P, C.POINTER(Fence)]),
    "mps_readback_async_f64": (C.c_int, [_P, _P, _P, _P, C.POINTER(Fence)]),
    "mps_fence_wait": (C.c_int, [_P, C.POINTER(Fence)]),
    "mps_host_alloc": (C.c_int, [_P, C.c_size_t, C.POINTER(_P)]),
    "mps_host_free": (C.c_int, [_P, _P]),
    "mps_profile_enable": (C.c_int, [_P, C.c_int32]),
    "mps_calibration_copy": (C.c_int, [_P, C.c_size_t, C.c_int32]),
    "mps_profile_read": (C.c_int, [_P, C.POINTER(KernelStat), C.c_size_t, C.POINTER(C.c_size_t)]),
    "mps_batch_clear": (C.c_int, [_P]),
    "mps_batch_add_scene": (C.c_int, [_P, C.POINTER(Physics), C.POINTER(SPHMaterial), C.c_size_t, _P, _P, _P, _P,
                                      C.c_size_t, _P, _P, _P]),
    "mps_batch_commit": (C.c_int, [_P]),
    "mps_scene_count": (C.c_int, [_P, C.POINTER(C.c_uint32)]),
    "mps_scene_info": (C.c_int, [_P, C.c_uint32, C.POINTER(SceneRange)]),
    "mps_scene_stats": (C.c_int, [_P, C.c_uint32, C.POINTER(StepStats)]),
    "mps_scene_set_material": (C.c_int, [_P, C.c_uint32, C.POINTER(SPHMaterial)]),
    "mps_adj_rollout": (C.c_int, [_P, C.c_uint32, C.c_uint32, _P]),
    "mps_autotune": (C.c_int, [_P, C.c_uint32, C.POINTER(TuneResult)]),
    "mps_adj_backward": (C.c_int, [_P, _P, _P, _P, _P, _P]),
    "mps_adj_backward_traj": (C.c_int, [_P, C.c_uint32, _P, _P, _P, _P, _P, _P]),
    "mps_adj_rollout_traj": (C.c_int, [_P, C.c_uint32, C.c_uint32, C.c_uint32, _P, _P, _P, _P]),
}

_libs: dict[str, C.CDLL] = {}


def load(path: str = LIB_PATH) -> C.CDLL:
    """Load libmps_sph.so (once per path); raises if it has not been built (no fallback path
    exists)."""
    key = os.path.abspath(path)
    if key in _libs:
        return _libs[key]
    if not os.path.exists(key):
        raise RuntimeError(f"{path} is missing: build it with `python -m mps_cuda_amd.build`")
    lib = C.CDLL(key)
    for name, (res, args) in _SIGS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _libs[key] = lib
    return lib


def exported_symbols() -> list[str]:
    return list(_SIGS)


def default_config() -> Config:
    cfg = Config()
    rc = load().mps_config_init(C.byref(cfg))
    if rc:
        raise RuntimeError("mps_config_init failed")
    return cfg


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def read_obj(path: str) -> tuple[np.ndarray, np.ndarray]:
    """Vertices (nv, 3) float64 and 0-based triangles (nt, 3) uint32 of a Wavefront OBJ file,
    read as MeshObject::LoadMesh reads it (MPSMeshObject.cu:58-94). Host-only: no GPU needed."""
    lib = load()
    nv, nt = C.c_size_t(), C.c_size_t()
    p = os.fsencode(path)
    rc = lib.mps_read_obj(None, p, None, C.byref(nv), None, C.byref(nt))
    if rc:
        raise MPSError(rc, f"cannot read OBJ file {path}")
    v = np.zeros((nv.value, 3), np.float64)
    t = np.zeros((nt.value, 3), np.uint32)
    rc = lib.mps_read_obj(None, p, _ptr(v), C.byref(nv), _ptr(t), C.byref(nt))
    if rc:
        raise MPSError(rc, f"cannot read OBJ file {path}")
    return v, t


class MPSError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"mps error {code}: {msg}")
        self.code = code


class HipSim:
    """One simulation context of the gfx950 library (positions etc. in float32)."""

    real = np.float32

    def __init__(self, cfg: Config | None = None):
        self.lib = load()
        self.cfg = cfg if cfg is not None else default_config()
        h = C.c_void_p()
        rc = self.lib.mps_create(C.byref(self.cfg), C.byref(h))
        if rc:
            raise MPSError(rc, "mps_create failed")
        self.h = h
        self.rollout_generation = 0  # bumped by every adj_rollout
        self._pinned = []            # pinned host buffers of host_array, freed in close()

    def close(self):
        if getattr(self, "h", None):
            self.lib.mps_synchronize(self.h)
            for p in getattr(self, "_pinned", []):
                self.lib.mps_host_free(self.h, p)
            self._pinned = []
            self.lib.mps_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, rc: int):
        if rc:
            raise MPSError(rc, self.lib.mps_last_error(self.h).decode())

    # ---- setup
    def set_physics(self, dt, gravity, box_min, box_max):
        p = Physics(dt, (C.c_double * 3)(*gravity), (C.c_double * 3)(*box_min), (C.c_double * 3)(*box_max))
        self._chk(self.lib.mps_set_physics(self.h, C.byref(p)))

    def set_fluid_material(self, radius, density, viscosity, xsph, surface_tension):
        m = SPHMaterial(radius, density, viscosity, xsph, surface_tension)
        self._chk(self.lib.mps_set_fluid_material(self.h, C.byref(m)))

    def upload_fluid(self, pos, radius, vel=None, mass=None):
        pos = np.ascontiguousarray(pos, np.float32).reshape(-1, 3)
        n = pos.shape[0]
        radius = np.ascontiguousarray(np.broadcast_to(np.asarray(radius, np.float32), (n,)))
        vel = None if vel is None else np.ascontiguousarray(vel, np.float32).reshape(n, 3)
        mass = None if mass is None else np.ascontiguousarray(np.broadcast_to(np.asarray(mass, np.float32), (n,)))
        self._chk(self.lib.mps_upload_fluid(self.h, n, _ptr(pos), _ptr(vel), _ptr(radius), _ptr(mass)))

    def upload_boundary(self, pos, radius, mass=None):
        pos = np.ascontiguousarray(pos, np.float32).reshape(-1, 3)
        n = pos.shape[0]
        radius = np.ascontiguousarray(np.broadcast_to(np.asarray(radius, np.float32), (n,)))
        mass = None if mass is None else np.ascontiguousarray(mass, np.float32)
        self._chk(self.lib.mps_upload_boundary(self.h, n, _ptr(pos), _ptr(radius), _ptr(mass)))

    def sample_boundary_mesh(self, verts, tris, radius, density=1.0, center=None, size=None):
        verts = np.ascontiguousarray(verts, np.float64).reshape(-1, 3)
        tris = np.ascontiguousarray(tris, np.uint32).reshape(-1, 3)
        c = None if center is None else np.ascontiguousarray(center, np.float64)
        s = None if size is None else np.ascontiguousarray(size, np.float64)
        m = MeshMaterial(radius, density)
        self._chk(self.lib.mps_sample_boundary_mesh(self.h, _ptr(verts), verts.shape[0], _ptr(tris), tris.shape[0],
                                                    _ptr(c), _ptr(s), C.byref(m)))

    def set_boundary_objects(self, counts):
        """The walls as len(counts) objects of these sizes (contiguous, current order)."""
        c = np.ascontiguousarray(counts, np.uint32)
        self._chk(self.lib.mps_set_boundary_objects(self.h, len(c), _ptr(c)))

    def resample_boundary_mesh(self, verts, radius) -> int:
        """ParticleSampling again after the sampled mesh's vertices moved; returns the number
        of faces generated anew."""
        verts = np.ascontiguousarray(verts, np.float64).reshape(-1, 3)
        m = MeshMaterial(radius, 1.0)
        k = C.c_size_t()
        self._chk(self.lib.mps_resample_boundary_mesh(self.h, _ptr(verts), verts.shape[0], C.byref(m), C.byref(k)))
        return int(k.value)

    def sample_boundary_obj(self, path, radius, density=1.0, center=None, size=None):
        """MeshObject::LoadMesh(path, center, size, density) + ParticleSampling."""
        c = None if center is None else np.ascontiguousarray(center, np.float64)
        s = None if size is None else np.ascontiguousarray(size, np.float64)
        m = MeshMaterial(radius, density)
        self._chk(self.lib.mps_sample_boundary_obj(self.h, os.fsencode(path), _ptr(c), _ptr(s), C.byref(m)))

    # ---- batched scenes
    def batch_clear(self):
        self._chk(self.lib.mps_batch_clear(self.h))

    def batch_add_scene(self, physics: dict, material: dict, pos, radius, vel=None, mass=None, bpos=None,
                        bradius=None, bmass=None):
        """Stage one scene. physics: dt, gravity, box_min, box_max; material: radius, density,
        viscosity, xsph, surface_tension."""
        p = Physics(physics["dt"], (C.c_double * 3)(*physics["gravity"]), (C.c_double * 3)(*physics["box_min"]),
                    (C.c_double * 3)(*physics["box_max"]))
        m = SPHMaterial(material["radius"], material["density"], material["viscosity"], material["xsph"],
                        material["surface_tension"])
        pos = np.ascontiguousarray(pos, np.float32).reshape(-1, 3)
        n = pos.shape[0]
        radius = np.ascontiguousarray(np.broadcast_to(np.asarray(radius, np.float32), (n,)))
        vel = None if vel is None else np.ascontiguousarray(vel, np.float32).reshape(n, 3)
        mass = None if mass is None else np.ascontiguousarray(np.broadcast_to(np.asarray(mass, np.float32), (n,)))
        nb = 0
        if bpos is not None:
            bpos = np.ascontiguousarray(bpos, np.float32).reshape(-1, 3)
            nb = bpos.shape[0]
            bradius = np.ascontiguousarray(np.broadcast_to(np.asarray(bradius, np.float32), (nb,)))
            bmass = None if bmass is None else np.ascontiguousarray(bmass, np.float32).reshape(nb)
        self._chk(self.lib.mps_batch_add_scene(self.h, C.byref(p), C.byref(m), n, _ptr(pos), _ptr(vel), _ptr(radius),
                                               _ptr(mass), nb, _ptr(bpos), _ptr(bradius), _ptr(bmass)))

    def batch_commit(self):
        self._chk(self.lib.mps_batch_commit(self.h))

    @property
    def scene_count(self) -> int:
        n = C.c_uint32()
        self._chk(self.lib.mps_scene_count(self.h, C.byref(n)))
        return n.value

    def scene_info(self, scene: int) -> dict:
        r = SceneRange()
        self._chk(self.lib.mps_scene_info(self.h, int(scene), C.byref(r)))
        return dict(fluid_offset=r.fluid_offset, n_fluid=r.n_fluid, boundary_offset=r.boundary_offset,
                    n_boundary=r.n_boundary)

    def scene_stats(self, scene: int) -> StepStats:
        st = StepStats()
        self._chk(self.lib.mps_scene_stats(self.h, int(scene), C.byref(st)))
        return st

    def scene_set_material(self, scene: int, radius, density, viscosity, xsph, surface_tension):
        m = SPHMaterial(radius, density, viscosity, xsph, surface_tension)
        self._chk(self.lib.mps_scene_set_material(self.h, int(scene), C.byref(m)))

    # ---- autotuning
    def autotune(self, steps: int = 5) -> dict:
        r = TuneResult()
        self._chk(self.lib.mps_autotune(self.h, int(steps), C.byref(r)))
        return dict(use_graphs=r.use_graphs, host_sync_every=r.host_sync_every,
                    ms_per_step=list(r.ms_per_step[:8]), cg_stream_coeffs=r.cg_stream_coeffs,
                    ms_coeffs=list(r.ms_coeffs), persistent_solver=r.persistent_solver,
                    ms_persistent=list(r.ms_persistent), cg_fold_in_kernel=r.cg_fold_in_kernel, ms_fold=list(r.ms_fold))

    # ---- adjoint
    def adj_rollout(self, n_steps: int, checkpoint_every: int = 10) -> np.ndarray:
        """Step n_steps keeping checkpoints; returns the loop counts (n_steps, scenes, 3) =
        (l_df, l_cg, l_cd) of every step."""
        counts = np.zeros((n_steps, self.scene_count, 3), np.uint32)
        self._chk(self.lib.mps_adj_rollout(self.h, int(n_steps), int(checkpoint_every), _ptr(counts)))
        self.rollout_generation += 1
        return counts

    def adj_rollout_traj(self, n_steps: int, obs_steps, checkpoint_every: int = 10, velocities: bool = False):
        """adj_rollout that also returns the positions (and velocities) after each step in
        obs_steps (1..n_steps): (pos (k, n, 3), vel (k, n, 3) or None, loop counts)."""
        n = self.get("FLUID_RADIUS").shape[0]
        obs = np.ascontiguousarray(obs_steps, np.uint32).reshape(-1)
        pos = np.zeros((obs.shape[0], n, 3), np.float32)
        vel = np.zeros((obs.shape[0], n, 3), np.float32) if velocities else None
        counts = np.zeros((n_steps, self.scene_count, 3), np.uint32)
        self._chk(self.lib.mps_adj_rollout_traj(self.h, int(n_steps), int(checkpoint_every), obs.shape[0], _ptr(obs),
                                                _ptr(pos), _ptr(vel), _ptr(counts)))
        self.rollout_generation += 1
        return pos, vel, counts

    def adj_backward(self, dloss_dpos, dloss_dvel=None):
        """(dL/dv0 (n, 3) float32, dL/dviscosity (scenes,), dL/dsurface_tension (scenes,))."""
        n = self.get("FLUID_RADIUS").shape[0]
        gx = np.ascontiguousarray(dloss_dpos, np.float32).reshape(n, 3)
        gv = None if dloss_dvel is None else np.ascontiguousarray(dloss_dvel, np.float32).reshape(n, 3)
        out = np.zeros((n, 3), np.float32)
        gmu = np.zeros(self.scene_count, np.float64)
        gsig = np.zeros(self.scene_count, np.float64)
        self._chk(self.lib.mps_adj_backward(self.h, _ptr(gx), _ptr(gv), _ptr(out), _ptr(gmu), _ptr(gsig)))
        return out, gmu, gsig

    def adj_backward_traj(self, obs_steps, dloss_dpos, dloss_dvel=None):
        """Trajectory loss: obs_steps (k,) in 1..n_steps, dloss_dpos / dloss_dvel (k, n, 3) at
        those states. Returns (dL/dv0 (n, 3), dL/dviscosity (scenes,), dL/dsurface_tension)."""
        n = self.get("FLUID_RADIUS").shape[0]
        obs = np.ascontiguousarray(obs_steps, np.uint32).reshape(-1)
        k = obs.shape[0]
        gx = np.ascontiguousarray(dloss_dpos, np.float32).reshape(k, n, 3)
        gv = None if dloss_dvel is None else np.ascontiguousarray(dloss_dvel, np.float32).reshape(k, n, 3)
        out = np.zeros((n, 3), np.float32)
        gmu = np.zeros(self.scene_count, np.float64)
        gsig = np.zeros(self.scene_count, np.float64)
        self._chk(self.lib.mps_adj_backward_traj(self.h, k, _ptr(obs), _ptr(gx), _ptr(gv), _ptr(out), _ptr(gmu),
                                                 _ptr(gsig)))
        return out, gmu, gsig

    # ---- stepping
    def step(self) -> StepStats:
        st = StepStats()
        self._chk(self.lib.mps_step(self.h, C.byref(st)))
        return st

    def step_noalloc(self) -> StepStats:
        """One step that allocates no device memory (mps_step_noalloc); the context must have
        been sized by an earlier mps_step on a scene at least this large."""
        st = StepStats()
        self._chk(self.lib.mps_step_noalloc(self.h, C.byref(st)))
        return st

    def run_phase(self, phase) -> StepStats:
        st = StepStats()
        ph = PHASE[phase] if isinstance(phase, str) else int(phase)
        self._chk(self.lib.mps_run_phase(self.h, ph, C.byref(st)))
        return st

    def synchronize(self):
        self._chk(self.lib.mps_synchronize(self.h))

    # ---- arrays
    def _dtype(self, kind):
        return {"real": self.real, "f32": np.float32, "u32": np.uint32}[kind]

    def get(self, name: str) -> np.ndarray:
        aid, kind, w = ARRAYS[name]
        n, eb = C.c_size_t(), C.c_size_t()
        self._chk(self.lib.mps_array_info(self.h, aid, C.byref(n), C.byref(eb)))
        dt = self._dtype(kind)
        out = np.empty((n.value, w) if w > 1 else (n.value,), dt)
        assert out.nbytes == n.value * eb.value, (name, out.nbytes, n.value, eb.value)
        self._chk(self.lib.mps_get_array(self.h, aid, _ptr(out), out.nbytes))
        return out

    def set(self, name: str, arr):
        aid, kind, w = ARRAYS[name]
        a = np.ascontiguousarray(arr, self._dtype(kind))
        self._chk(self.lib.mps_set_array(self.h, aid, _ptr(a), a.nbytes))

    @property
    def frame(self) -> int:
        f = C.c_uint32()
        self._chk(self.lib.mps_get_frame(self.h, C.byref(f)))
        return f.value

    @frame.setter
    def frame(self, v: int):
        self._chk(self.lib.mps_set_frame(self.h, int(v)))

    # ---- render feed
    def readback_async(self, host_pos=None, host_radius=None, host_rgba=None) -> Fence:
        f = Fence()
        self._chk(self.lib.mps_readback_async(self.h, _ptr(host_pos), _ptr(host_radius), _ptr(host_rgba), C.byref(f)))
        return f

    def readback_async_f64(self, host_pos=None, host_radius=None, host_rgba=None) -> Fence:
        """The renderer's layout: positions and radii as float64 arrays, colours as float32."""
        f = Fence()
        self._chk(self.lib.mps_readback_async_f64(self.h, _ptr(host_pos), _ptr(host_radius), _ptr(host_rgba),
                                                  C.byref(f)))
        return f

    def host_array(self, shape, dtype):
        """A numpy view of pinned host memory (mps_host_alloc), freed with the context."""
        n = int(np.prod(shape)) * np.dtype(dtype).itemsize
        p = C.c_void_p()
        self._chk(self.lib.mps_host_alloc(self.h, max(n, 1), C.byref(p)))
        self._pinned.append(p)
        buf = (C.c_char * max(n, 1)).from_address(p.value)
        return np.frombuffer(buf, dtype=dtype, count=int(np.prod(shape))).reshape(shape)

    def fence_wait(self, f: Fence):
        self._chk(self.lib.mps_fence_wait(self.h, C.byref(f)))

    # ---- per-kernel timing
    def profile(self, on: bool = True):
        self._chk(self.lib.mps_profile_enable(self.h, int(bool(on))))

    def calibration_copy(self, nbytes: int, reps: int = 3):
        self._chk(self.lib.mps_calibration_copy(self.h, int(nbytes), int(reps)))

    def profile_read(self) -> dict:
        cap = 64
        arr = (KernelStat * cap)()
        n = C.c_size_t()
        self._chk(self.lib.mps_profile_read(self.h, arr, cap, C.byref(n)))
        out = {}
        for k in range(min(n.value, cap)):
            st = arr[k]
            out[st.name.decode()] = dict(launches=int(st.launches), useful=int(st.useful), total_ms=float(st.total_ms))
        return out
