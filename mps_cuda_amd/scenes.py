"""Scene definitions: the reference's own test scenes and the dam-break / droplet workloads.

A Scene is plain data (numpy arrays + parameter dicts) that both the gfx950 library and
the CPU oracle can be loaded from, so parity runs start from identical inputs.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from . import _capi as capi

# 12-triangle cube standing in for the missing obj/cube.obj (MPSSystem.cpp:63). Vertex 0 is
# the minimum corner so AABB::AddPoint's else-if (MPSAABB.h:20-28) sees the full extent.
CUBE_VERTS = np.array(
    [[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0], [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]], np.float64)
CUBE_TRIS = np.array(
    [[0, 2, 1], [0, 3, 2],   # z = 0
     [4, 5, 6], [4, 6, 7],   # z = 1
     [0, 1, 5], [0, 5, 4],   # y = 0
     [3, 7, 6], [3, 6, 2],   # y = 1
     [0, 4, 7], [0, 7, 3],   # x = 0
     [1, 2, 6], [1, 6, 5]],  # x = 1
    np.uint32)


@dataclass
class Scene:
    name: str
    pos: np.ndarray                 # (N, 3) float64 (exact float32 values)
    radius: np.ndarray              # (N,)
    vel: np.ndarray | None
    physics: dict
    material: dict
    config: dict
    boundary_mesh: dict | None = None   # verts, tris, center, size, radius, density
    boundary_pos: np.ndarray | None = None
    boundary_radius: np.ndarray | None = None
    steps: int = 1
    extra: dict = field(default_factory=dict)

    @property
    def n(self) -> int:
        return int(self.pos.shape[0])


def make_config(scene: Scene) -> capi.Config:
    cfg = capi.default_config()
    for k, v in scene.config.items():
        if k == "hash_size":
            for i in range(3):
                cfg.hash_size[i] = int(v[i])
        else:
            setattr(cfg, k, v)
    return cfg


def load_into(sim, scene: Scene):
    """Push a scene into a HipSim (or an oracle handle with the same methods)."""
    p = scene.physics
    sim.set_physics(p["dt"], p["gravity"], p["box_min"], p["box_max"])
    m = scene.material
    sim.set_fluid_material(m["radius"], m["density"], m["viscosity"], m["xsph"], m["surface_tension"])
    if scene.boundary_mesh is not None:
        b = scene.boundary_mesh
        sim.sample_boundary_mesh(b["verts"], b["tris"], b["radius"], b.get("density", 1.0), b.get("center"), b.get("size"))
    elif scene.boundary_pos is not None:
        sim.upload_boundary(scene.boundary_pos, scene.boundary_radius)
    sim.upload_fluid(scene.pos, scene.radius, scene.vel)


def _f32(a) -> np.ndarray:
    return np.asarray(a, np.float32).astype(np.float64)


# ---------------------------------------------------------------------------- reference scenes
def surface_tension_test_scene(size: int = 5) -> Scene:
    """System::SurfaceTensionTestScene (MPSSystem.cpp:552-651): two size^3 blocks, r = 0.1."""
    radius = 0.1
    gap = radius * float(np.float32(0.9))
    offset = float(np.float32(0.001))
    stride = gap + gap + offset
    L = stride * (size - 1) + gap + gap
    start = -L
    pos = []
    for z in range(size):
        for y in range(size):
            for x in range(size):
                pos.append([start - L * 0.025 + x * stride + gap, start + y * stride + gap, start + z * stride + gap])
                pos.append([L * 0.025 + x * stride + gap, start + y * stride + gap, start + z * stride + gap])
    pos = _f32(pos)
    bmin = [start - L * 0.025 + gap - 3.5, start + gap - 3.5, start + gap - 3.5]
    bmax = [L * 0.025 + (size - 1) * stride + gap + 3.5, start + (size - 1) * stride + gap + 3.5,
            start + (size - 1) * stride + gap + 3.5]
    center = [(a + b) * 0.5 for a, b in zip(bmin, bmax)]
    ext = [b - a for a, b in zip(bmin, bmax)]
    return Scene(
        name=f"surface_tension_test_{size}",
        pos=pos, radius=np.full(pos.shape[0], radius), vel=None,
        physics=dict(dt=0.005, gravity=(0.0, -9.8, 0.0), box_min=bmin, box_max=bmax),
        material=dict(radius=radius * 4, density=1.0, viscosity=0.025, xsph=0.1, surface_tension=100.0),
        config=dict(kernel=capi.KERNEL_POLY6_SPIKY, hash_size=(64, 64, 64), cell_size=radius * 4),
        boundary_mesh=dict(verts=CUBE_VERTS, tris=CUBE_TRIS, center=center, size=ext, radius=radius * 4, density=1.0),
    )


def viscosity_test_scene(size: int = 5, height: int = 20) -> Scene:
    """System::ViscosityTestScene (MPSSystem.cpp:455-550): a size x height x size column."""
    radius = 0.1
    gap = radius * float(np.float32(0.9))
    offset = float(np.float32(0.001))
    stride = gap + gap + offset
    lw = stride * (size - 1) + gap + gap
    start = -lw * float(np.float32(0.5))
    start_h = 0.0
    pos = []
    for y in range(height):
        for z in range(size):
            for x in range(size):
                pos.append([start + x * stride + gap, start + y * stride + gap, start + z * stride + gap])
    pos = _f32(pos)
    bmin = [start + gap - 5.5, start_h - 1.0, start + gap - 5.5]
    bmax = [start + (size - 1) * stride + gap + 5.5, start_h + (height - 1) * stride + gap + 5.5,
            start + (size - 1) * stride + gap + 5.5]
    center = [(a + b) * 0.5 for a, b in zip(bmin, bmax)]
    ext = [b - a for a, b in zip(bmin, bmax)]
    return Scene(
        name=f"viscosity_test_{size}x{height}",
        pos=pos, radius=np.full(pos.shape[0], radius), vel=None,
        physics=dict(dt=0.005, gravity=(0.0, -9.8, 0.0), box_min=bmin, box_max=bmax),
        material=dict(radius=radius * 4, density=1.0, viscosity=0.025, xsph=0.1, surface_tension=100.0),
        config=dict(kernel=capi.KERNEL_POLY6_SPIKY, hash_size=(64, 64, 64), cell_size=radius * 4),
        boundary_mesh=dict(verts=CUBE_VERTS, tris=CUBE_TRIS, center=center, size=ext, radius=radius * 4, density=1.0),
    )


# ---------------------------------------------------------------------------- dam breaks
def dam_break(nx: int, ny: int, nz: int, spacing: float, jitter: float, seed: int, box_min, box_max,
              dt: float, steps: int, hash_size, viscosity=0.01, surface_tension=0.05, name="dam_break") -> Scene:
    """Fluid block x = spacing*(i+1) + U(-jitter, jitter) in the lower corner of a walled tank.

    Particle radius = spacing/2 (support h = 2*spacing); walls are the tank's 12-triangle
    mesh sampled at spacing (R_m = 2*spacing, boundary radius spacing/2).
    """
    rng = np.random.default_rng(seed)
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    base = np.stack([i.ravel(), j.ravel(), k.ravel()], axis=1).astype(np.float64)
    base = np.asarray(box_min, np.float64)[None, :] + spacing * (base + 1.0)
    pos = _f32(base + rng.uniform(-jitter, jitter, size=base.shape))
    h = 2.0 * spacing
    center = [(a + b) * 0.5 for a, b in zip(box_min, box_max)]
    ext = [b - a for a, b in zip(box_min, box_max)]
    return Scene(
        name=name, pos=pos, radius=np.full(pos.shape[0], spacing * 0.5), vel=None,
        physics=dict(dt=dt, gravity=(0.0, -9.81, 0.0), box_min=list(box_min), box_max=list(box_max)),
        material=dict(radius=h, density=1000.0, viscosity=viscosity, xsph=0.1, surface_tension=surface_tension),
        config=dict(kernel=capi.KERNEL_CUBIC, hash_size=tuple(hash_size), cell_size=h, df_tol=1.0e-2, cd_tol=1.0e-3),
        boundary_mesh=dict(verts=CUBE_VERTS, tris=CUBE_TRIS, center=center, size=ext, radius=h, density=1.0),
        steps=steps,
    )


def dam_break_cpu(n_side: int = 32, steps: int = 500) -> Scene:
    """The CPU reference workload: 32^3 fluid on a 0.02 m lattice (+-0.002 m jitter, seed 0)
    in a 1 m box, h = 0.04 m, dt = 1 ms, 500 steps."""
    return dam_break(n_side, n_side, n_side, 0.02, 0.002, 0, (0.0, 0.0, 0.0), (1.0, 1.0, 1.0), 1.0e-3, steps,
                     (64, 64, 64), name=f"dam_break_{n_side}^3_h0.04")


def dam_break_1m(n_side: int = 100, steps: int = 1000) -> Scene:
    """The single-GPU workload: 100^3 = 1,000,000 fluid on a 0.01 m lattice (+-0.001 m jitter,
    seed 0) in a 3.0 x 1.5 x 1.5 m tank, h = 0.02 m, dt = 0.5 ms, hash 256 x 128 x 128."""
    return dam_break(n_side, n_side, n_side, 0.01, 0.001, 0, (0.0, 0.0, 0.0), (3.0, 1.5, 1.5), 5.0e-4, steps,
                     (256, 128, 128), name=f"dam_break_{n_side}^3_h0.02")


def droplet(n: int = 65536, radius: float = 0.25, seed: int = 0, steps: int = 100) -> Scene:
    """A jittered sphere of n particles (spacing ~0.01 m, h = 0.02 m) above the floor of a 1 m box."""
    rng = np.random.default_rng(seed)
    sp = (4.0 / 3.0 * np.pi * radius ** 3 / n) ** (1.0 / 3.0)
    m = int(np.ceil(2 * radius / sp)) + 2
    g = (np.arange(m) - (m - 1) / 2.0) * sp
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    pts = np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1)
    d = np.linalg.norm(pts, axis=1)
    pts = pts[np.argsort(d, kind="stable")[:n]]
    center = np.array([0.5, 0.45, 0.5])
    pos = _f32(pts + center + rng.uniform(-0.1 * sp, 0.1 * sp, size=pts.shape))
    h = 0.02
    return Scene(
        name=f"droplet_{n}", pos=pos, radius=np.full(n, h / 4), vel=None,
        physics=dict(dt=1.0e-3, gravity=(0.0, -9.81, 0.0), box_min=[0.0, 0.0, 0.0], box_max=[1.0, 1.0, 1.0]),
        material=dict(radius=h, density=1000.0, viscosity=0.02, xsph=0.1, surface_tension=0.3),
        config=dict(kernel=capi.KERNEL_CUBIC, hash_size=(64, 64, 64), cell_size=h, df_tol=1.0e-2, cd_tol=1.0e-3),
        boundary_mesh=dict(verts=CUBE_VERTS, tris=CUBE_TRIS, center=[0.5, 0.5, 0.5], size=[1.0, 1.0, 1.0], radius=h, density=1.0),
        steps=steps,
    )


# ---------------------------------------------------------------------------- batched scenes
def batch_scenes(n_scenes: int = 64, seed: int = 2, n_min: int = 8192, n_max: int = 262144, spacing: float = 0.02,
                 dt: float = 1.0e-3, steps: int = 300) -> list[Scene]:
    """Independent scenes for the batched step: sizes log-uniform in [n_min, n_max] (seeded);
    each a jittered fluid block in the corner or a sphere falling from near the top of its
    own box (coin flip); viscosity ~ U(0.005, 0.05), surface tension ~ U(0.01, 0.5). The box
    edge is 1 m, grown (and recorded in `extra`) when the fluid would not fit. All scenes
    share one configuration (cubic kernel, h = 2*spacing, hash 64^3, DF tol 1e-2, CD tol 1e-3),
    as the scenes of one context must."""
    rng = np.random.default_rng(seed)
    sizes = np.rint(np.exp(rng.uniform(np.log(n_min), np.log(n_max), n_scenes))).astype(np.int64)
    kinds = rng.integers(0, 2, n_scenes)
    mus = rng.uniform(0.005, 0.05, n_scenes)
    sigmas = rng.uniform(0.01, 0.5, n_scenes)
    h = 2.0 * spacing
    out = []
    for s in range(n_scenes):
        n = int(sizes[s])
        if kinds[s] == 0:  # block in the lower corner, filled layer by layer
            a = int(np.ceil(n ** (1.0 / 3.0)))
            k, j, i = np.meshgrid(np.arange(a), np.arange(a), np.arange(a), indexing="ij")
            lat = np.stack([i.ravel(), k.ravel(), j.ravel()], axis=1)[:n].astype(np.float64)  # y (height) last
            base = spacing * (lat + 1.0)
            L = max(1.0, spacing * (a + 2))
            kind = "block"
        else:  # sphere, nearest lattice points to its centre
            R = spacing * (3.0 * n / (4.0 * np.pi)) ** (1.0 / 3.0)
            m = int(np.ceil(2 * R / spacing)) + 4
            g = (np.arange(m) - (m - 1) / 2.0) * spacing
            X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
            pts = np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1)
            pts = pts[np.argsort(np.linalg.norm(pts, axis=1), kind="stable")[:n]]
            L = max(1.0, 2 * R + 8 * spacing)
            base = pts + np.array([L / 2, L - R - 3 * spacing, L / 2])
            kind = "sphere"
        pos = _f32(base + rng.uniform(-0.1 * spacing, 0.1 * spacing, size=base.shape))
        out.append(Scene(
            name=f"batch{s}_{kind}_{n}", pos=pos, radius=np.full(n, spacing * 0.5), vel=None,
            physics=dict(dt=dt, gravity=(0.0, -9.81, 0.0), box_min=[0.0, 0.0, 0.0], box_max=[L, L, L]),
            material=dict(radius=h, density=1000.0, viscosity=float(mus[s]), xsph=0.1, surface_tension=float(sigmas[s])),
            config=dict(kernel=capi.KERNEL_CUBIC, hash_size=(64, 64, 64), cell_size=h, df_tol=1.0e-2, cd_tol=1.0e-3),
            boundary_mesh=dict(verts=CUBE_VERTS, tris=CUBE_TRIS, center=[L / 2] * 3, size=[L] * 3, radius=h, density=1.0),
            steps=steps, extra=dict(kind=kind, box_edge=L),
        ))
    return out


def sample_walls(sim_factory, scene: Scene):
    """Boundary particles of a scene's wall mesh, sampled by a fresh single-scene context:
    (positions (B, 3), radii (B,), masses (B,)) as float32."""
    sim = sim_factory(make_config(scene))
    p = scene.physics
    sim.set_physics(p["dt"], p["gravity"], p["box_min"], p["box_max"])
    b = scene.boundary_mesh
    sim.sample_boundary_mesh(b["verts"], b["tris"], b["radius"], b.get("density", 1.0), b.get("center"), b.get("size"))
    out = (sim.get("BND_POS").astype(np.float32), sim.get("BND_RADIUS").astype(np.float32),
           sim.get("BND_MASS").astype(np.float32))
    if hasattr(sim, "close"):
        sim.close()
    return out


def load_plain(sim, scene: Scene, walls):
    """One scene into a single-scene context from pre-sampled walls (same inputs as a batch)."""
    p, m = scene.physics, scene.material
    sim.set_physics(p["dt"], p["gravity"], p["box_min"], p["box_max"])
    sim.set_fluid_material(m["radius"], m["density"], m["viscosity"], m["xsph"], m["surface_tension"])
    sim.upload_boundary(walls[0], walls[1], walls[2])
    sim.upload_fluid(scene.pos, scene.radius, scene.vel)


def load_batch(sim, scenes: list[Scene], walls: list):
    """All scenes into one context (the batch entry points), in order."""
    sim.batch_clear()
    for sc, w in zip(scenes, walls):
        sim.batch_add_scene(sc.physics, sc.material, sc.pos, sc.radius, vel=sc.vel, bpos=w[0], bradius=w[1], bmass=w[2])
    sim.batch_commit()


# ---------------------------------------------------------------------------- differentiable droplet
def fourier_velocity(pos: np.ndarray, seed: int, modes: int = 3, amplitude: float = 0.5) -> np.ndarray:
    """Smooth velocity field sum_k a_k sin(2 pi f_k . x + phi_k) (seeded directions, wave
    vectors of 1-3 cycles per metre, phases), scaled so that max |v| = amplitude."""
    rng = np.random.default_rng(seed)
    v = np.zeros_like(pos, dtype=np.float64)
    for _ in range(modes):
        a = rng.normal(size=3)
        f = rng.uniform(1.0, 3.0, size=3) * rng.choice([-1.0, 1.0], size=3)
        phi = rng.uniform(0.0, 2.0 * np.pi)
        v += np.sin(2.0 * np.pi * pos @ f + phi)[:, None] * a[None, :]
    v *= amplitude / max(np.linalg.norm(v, axis=1).max(), 1e-12)
    return _f32(v)


def diff_droplet(n: int = 65536, radius: float = 0.25, seed: int = 0, steps: int = 100, viscosity: float = 0.02,
                 surface_tension: float = 0.3, h: float = 0.02) -> Scene:
    """The differentiable droplet: a jittered sphere of n particles (spacing from n and the
    radius, support h) falling in a 1 m box at dt = 1 ms, with a seeded smooth initial velocity
    (3 Fourier modes, 0.5 m/s) in `extra["v0"]`. Configured for the adjoint: cubic kernel,
    no Z-sort (with the reference's volume quirk it reorders wall volumes), and a relative CG
    tolerance (r.r below 1e-8 of its initial value, reachable in FP32) so the viscosity solve
    matches its implicit derivative."""
    sc = droplet(n=n, radius=radius, seed=seed, steps=steps)
    if h != 0.02:
        sc.radius = np.full(n, h / 4)
        sc.material["radius"] = h
        sc.config["cell_size"] = h
        sc.boundary_mesh["radius"] = h
    sc.name = f"diff_droplet_{n}"
    sc.material["viscosity"] = viscosity
    sc.material["surface_tension"] = surface_tension
    sc.config.update(zsort_interval=0, cg_tol=-1.0e-8, cg_max_iter=200)
    sc.extra["v0"] = fourier_velocity(sc.pos, seed + 1)
    return sc
