#!/usr/bin/env python3
"""Throughput of the SPH time step on one MI355X.

Workload (default): plain DFSPH dam break, 100^3 = 1,000,000 fluid particles on a 0.01 m
lattice (+-0.001 m jitter, seed 0) in a 3.0 x 1.5 x 1.5 m tank with wall particles at
0.01 m, cubic kernel, h = 0.02 m, dt = 0.5 ms, density tol 0.1 %, divergence tol 1 %,
viscosity 0.01, surface tension 0.05, FP32.

A "step" is one System::OnUpdate: hash, neighbour lists, boundary volume, density/alpha,
divergence-free solve, viscosity + surface-tension CG, constant-density solve, advection.
`value` = fluid particles x timed steps / wall time of the timed steps (inputs resident in
HBM). The W warm-up steps run first; the scene is then reloaded and advanced --settle steps
untimed (default 500: past the start-up transient, where the lattice collapses from 0.52 rho0
towards rest density, the neighbour counts rise from ~28 to ~40 per particle and the CG iteration
count from ~13 to ~50). The library's autotuner then picks its launch settings by timing trial
steps from that state, which it restores, and the state is kept. The timed steps are steps
settle..settle+K-1 from it (no
device allocation inside the region, no per-kernel events); the same steps then run again from
the kept state with HIP events around every kernel on the context's stream, which gives the
per-kernel times. The CPU baseline steps from the same kept state.

Prints ONE JSON line. Extra objects: `roofline` (dominant kernel: algorithmic bytes from the
measured neighbour counts and iteration counts / its HIP-event time; `traffic` = its HBM bytes
per launch from the committed rocprofv3 counter summary, null unless that summary covers the
same steps), `cpu_baseline` (the FP32 CPU
oracle on a bounded sample of the same workload), per-kernel table.
"""
from __future__ import annotations

import argparse
import json
import os
import re
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

METRIC = "particle-steps/s at 0.1% density err: batched vs plain; adjoint/forward cost"
HBM_PEAK_GBS = 8000.0
COPY_BYTES = 512 * 1024 * 1024  # per buffer of the in-run copy measurement (read + written)

# Algorithmic bytes per launch, per fluid particle N and per neighbour entry E (FP32 SoA,
# 4-byte indices, compulsory traffic only): the per-pass figures of the measurement plan.
#   A.p (ST + XSPH + viscosity + final):   68 N + 4 E_ff
#   DF/CD evaluate (delta, stiffness, sum): 44 N + 4 E_f
#   DF/CD apply (force + velocity):         48 N + 4 E_f
#   density + alpha (two passes fused):     2 x (28 N + 4 E_f)
#   ST factor / CG right-hand side:         36 N + 4 E_ff / 56 N + 4 E_ff
#   CG vector update + direction:           132 N per iteration (x, r, p, Ap)
#   neighbour search ff + fb:               96 N + 4 E_f
def kernel_bytes(name: str, N: int, Eff: float, Efb: float) -> float | None:
    Ef = Eff + Efb
    table = {
        "k_cg_apply<1>": 68 * N + 4 * Eff,
        "k_cg_apply<0>": 68 * N + 4 * Eff,
        "k_pressure_eval<DF>": 44 * N + 4 * Ef,
        "k_pressure_eval<CD>": 44 * N + 4 * Ef,
        "k_pressure_apply": 48 * N + 4 * Ef,
        "k_density_alpha": 2 * (28 * N + 4 * Ef) + 24 * N,
        "k_st_factor": 36 * N + 4 * Eff,
        "k_cg_setup": 56 * N + 4 * Eff,
        "k_neighbors_fluid": 96 * N + 4 * Ef,
        "k_advect": 48 * N,
    }
    return table.get(name)


TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "dam_break_1m_hbm_bytes.md")


def say(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def traffic_window() -> list | None:
    """The steps the counter passes of the traffic summary covered (its title: "steps A..B")."""
    try:
        with open(TRAFFIC_PROFILE) as f:
            m = re.search(r"steps (\d+)\.\.(\d+)", f.readline())
    except OSError:
        return None
    return [int(m.group(1)), int(m.group(2))] if m else None


def measured_traffic(kernel: str, window: list, variant: str = "") -> float | None:
    """HBM bytes per launch of `kernel` on this workload: FETCH_SIZE + WRITE_SIZE from two
    rocprofv3 counter passes over the same timed window, each scaled by what the counter reads
    for a plain copy of known size (the summary's last column: MB per working dispatch, mean).
    None when the summary covers other steps than `window` (the lists and halos grow as the
    fluid settles, so bytes per launch depend on the window). The profiler spells out the
    template arguments that the library's timing table leaves out: `variant` selects among them
    (`k_cg_apply<1>` with variant "false" is the summary's `k_cg_apply<1, false, ...>` row)."""
    if traffic_window() != list(window):
        return None
    want = kernel[:-1] + ", " + variant if variant else kernel[:-1] + ","
    try:
        with open(TRAFFIC_PROFILE) as f:
            for line in f:
                cells = [c.strip() for c in line.strip().strip("|").split("|")]
                if len(cells) < 5:
                    continue
                name = cells[0]
                if name == kernel or (kernel.endswith(">") and name.startswith(want)):
                    return float(cells[-1]) * 1.0e6
    except (OSError, ValueError):
        return None
    return None


def step_bytes(N, B, Eff, Efb, Ebb, l_df, l_cd, l_cg, cap, C):
    """Whole-step algorithmic bytes (measurement plan table), from measured counts."""
    Ef = Eff + Efb
    J_df, J_cd = min(l_df, cap), min(l_cd, cap)
    A_df, A_cd = l_df - 1, l_cd - 1
    I_cg = max(min(l_cg - 1, cap), 0)
    b = 48 * N + 48 * B + 16 * C
    b += 48 * N + 4 * Eff + 48 * N + 4 * Efb + 48 * B + 4 * Ebb
    b += 36 * B + 4 * Ebb
    b += 2 * (28 * N + 4 * Ef) + 24 * N
    b += (J_df + J_cd) * (44 * N + 4 * Ef) + (A_df + A_cd) * (48 * N + 4 * Ef)
    b += (36 * N + 4 * Eff) + (56 * N + 4 * Eff)
    b += (1 + I_cg) * (68 * N + 4 * Eff) + I_cg * 132 * N + 96 * N
    b += 48 * N
    b += (28 * N + 96 * N + 28 * B + 128 * B) / 50.0
    return b


def _host_cpu() -> dict:
    """What the CPU baseline ran on: logical CPUs of the host, those this process may use, and
    the model name from /proc/cpuinfo."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            model = next((l.split(":", 1)[1].strip() for l in f if l.startswith("model name")), None)
    except OSError:
        pass
    usable = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else None
    return {"logical_cpus": os.cpu_count(), "usable_cpus": usable, "model": model}


SNAP_KEYS = ("FLUID_POS", "FLUID_VEL", "FLUID_RADIUS", "FLUID_MASS", "BND_POS", "BND_RADIUS", "BND_MASS", "BND_VEL",
             "BND_VOLUME")


def snapshot(sim) -> dict:
    """Everything a step reads that earlier steps wrote, in reference order: the fluid, the
    walls in their current order with their never-cleared volumes, and the frame counter."""
    snap = {k: sim.get(k).copy() for k in SNAP_KEYS}
    snap["frame"] = sim.frame
    return snap


def restore(sim, scene, snap) -> None:
    """Put a snapshot back (same sizes: no device allocation); stepping on from it gives the
    same results bit for bit as stepping on from the state it was taken of."""
    p, m = scene.physics, scene.material
    sim.set_physics(p["dt"], p["gravity"], p["box_min"], p["box_max"])
    sim.set_fluid_material(m["radius"], m["density"], m["viscosity"], m["xsph"], m["surface_tension"])
    sim.upload_boundary(snap["BND_POS"], snap["BND_RADIUS"], snap["BND_MASS"])
    sim.set("BND_VEL", snap["BND_VEL"])
    sim.set("BND_VOLUME", snap["BND_VOLUME"])
    sim.upload_fluid(snap["FLUID_POS"], snap["FLUID_RADIUS"], snap["FLUID_VEL"], snap["FLUID_MASS"])
    sim.frame = snap["frame"]


def cpu_baseline(scene, snap, budget_s: float, max_steps: int) -> dict:
    """FP32 CPU oracle on the same workload and window: from the kept state `snap`, one untimed
    step, then timed steps with every OpenMP thread until the budget is used (at least one),
    then one more step on one thread."""
    from mps_cuda_amd import scenes
    from oracle import binding

    threads = binding.set_threads(0)
    o = binding.OracleSim(scenes.make_config(scene), 32)
    restore(o, scene, snap)
    f0 = snap["frame"]
    o.step()
    n_done, t_used = 0, 0.0
    while n_done < max_steps and (n_done == 0 or t_used < budget_s):
        t0 = time.perf_counter()
        o.step()
        t_used += time.perf_counter() - t0
        n_done += 1
    rate = scene.n * n_done / t_used
    binding.set_threads(1)
    t0 = time.perf_counter()
    o.step()
    t_one = time.perf_counter() - t0
    binding.set_threads(threads)
    del o
    # the reference's CPU-runnable workload (32^3 dam break): steady state over steps 50..499,
    # or as many of them as fit the budget
    small = scenes.dam_break_cpu()
    q = binding.OracleSim(scenes.make_config(small), 32)
    scenes.load_into(q, small)
    for _ in range(50):
        q.step()
    n_small, t_small = 0, 0.0
    while 50 + n_small < small.steps and t_small < budget_s:
        t0 = time.perf_counter()
        q.step()
        t_small += time.perf_counter() - t0
        n_small += 1
    return {
        "value": rate, "unit": "particle-steps/s", "cores": threads, "kind": "port",
        "sample": f"FP32 oracle (oracle/sph_oracle.cpp, OpenMP {threads} threads) on the same "
                  f"{scene.n}-particle workload from the same kept state: steps {f0 + 1}..{f0 + n_done} after "
                  f"one untimed step, {t_used:.1f} s; then step {f0 + n_done + 1} on one thread, {t_one:.1f} s",
        "single_thread_value": scene.n / t_one,
        "small_workload": {
            "value": small.n * n_small / t_small, "unit": "particle-steps/s", "cores": threads,
            "sample": f"{small.name}, {small.n} particles: steps 50..{49 + n_small} after 50 untimed steps, {t_small:.1f} s",
        },
        "host": _host_cpu(),
    }


def batched_leg(n_scenes: int, steps: int, warmup: int, seed: int = 2, graphs: str = "auto") -> dict:
    """The batched form against the plain path on the same GPU: `n_scenes` independent scenes
    (sizes log-uniform 8k-256k, blocks or falling spheres, per-scene viscosity and surface
    tension) stepped together in one context, then each scene stepped alone in a context of
    its own. Both start from the same state and run the same steps (after `warmup` untimed
    steps); particle-steps/s = sum of scene sizes x steps / wall time."""
    from mps_cuda_amd import HipSim, scenes

    sc = scenes.batch_scenes(n_scenes=n_scenes, seed=seed)
    walls = [scenes.sample_walls(HipSim, s) for s in sc]
    total = sum(s.n for s in sc)
    cfg = scenes.make_config(sc[0])
    if graphs != "auto":
        cfg.use_graphs = int(graphs)
    print(f"batched leg: {n_scenes} scenes, {total} particles", file=sys.stderr, flush=True)

    bat = HipSim(cfg)
    scenes.load_batch(bat, sc, walls)
    for _ in range(warmup):
        bat.step()
    # launch mode (graph replay of solver chunks or direct launches, host-poll interval) picked
    # by the library's autotuner on this batch, unless given; the timed steps use it
    tuning = bat.autotune(2) if graphs == "auto" else None
    bat.synchronize()
    l_cg = []
    t0 = time.perf_counter()
    for _ in range(steps):
        st = bat.step()
        l_cg.append(st.l_cg)
    bat.synchronize()
    t_bat = time.perf_counter() - t0
    bat.close()
    print(f"batched: {1000.0 * t_bat / steps:.2f} ms/step", file=sys.stderr, flush=True)

    t_plain = 0.0
    for s, w in zip(sc, walls):
        g = HipSim(cfg)
        scenes.load_plain(g, s, w)
        for _ in range(warmup):
            g.step()
        g.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            g.step()
        g.synchronize()
        t_plain += time.perf_counter() - t0
        g.close()
    print(f"plain, scene by scene: {1000.0 * t_plain / steps:.2f} ms/step", file=sys.stderr, flush=True)
    v_bat, v_plain = total * steps / t_bat, total * steps / t_plain
    return {
        "scenes": n_scenes, "particles": total, "steps": steps, "warmup": warmup,
        "sizes": "log-uniform 8192..262144, seed %d" % seed,
        "value": v_bat, "ms_per_step": 1000.0 * t_bat / steps,
        "plain_value": v_plain, "plain_ms_per_step": 1000.0 * t_plain / steps,
        "speedup": v_bat / v_plain, "l_cg_max_mean": sum(l_cg) / len(l_cg),
        "timed_launch_mode": ({"use_graphs": tuning["use_graphs"], "host_sync_every": tuning["host_sync_every"]} if tuning
                              else {"use_graphs": int(cfg.use_graphs), "host_sync_every": int(cfg.host_sync_every)}),
        "solver_tuning": tuning,
    }


def adjoint_leg(n: int, steps: int, every: int, seed: int = 0, graphs: str = "auto") -> dict:
    """The differentiable droplet: a sphere of n particles falling for `steps` steps; target
    from viscosity 0.02, surface tension 0.3 and a smooth seeded initial velocity; loss = MSE of
    the final positions. At the guess (0.025, 0.1, same initial velocity) time one plain forward,
    the checkpointed rollout and the backward pass. The gradients are checked against central
    differences of the FP64 oracle's forward by the GPU tests (tests/test_gpu_adjoint.py), not
    here: finite differences of this FP32 forward are limited by its rounding."""
    from mps_cuda_amd import HipSim, scenes

    sc = scenes.diff_droplet(n=n, steps=steps, seed=seed)
    cfg = scenes.make_config(sc)
    if graphs != "auto":
        cfg.use_graphs = int(graphs)
    g = HipSim(cfg)
    v0 = sc.extra["v0"]

    def reset(mu, sig, v):
        scenes.load_into(g, sc)
        g.set_fluid_material(**dict(sc.material, viscosity=mu, surface_tension=sig))
        g.set("FLUID_VEL", v)

    def forward(mu, sig, v):
        reset(mu, sig, v)
        g.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            g.step()
        g.synchronize()
        return g.get("FLUID_POS").astype(np.float64), time.perf_counter() - t0

    target, _ = forward(0.02, 0.3, v0)
    # graph replay and the poll interval for this small, many-iteration scene (state kept),
    # unless given
    tuning = g.autotune(2) if graphs == "auto" else None
    mu0, sig0 = 0.025, 0.1
    x, t_fwd = forward(mu0, sig0, v0)

    def loss(xx):
        return float(np.mean((xx - target) ** 2))

    reset(mu0, sig0, v0)
    g.synchronize()
    t0 = time.perf_counter()
    counts = g.adj_rollout(steps, every)
    t_roll = time.perf_counter() - t0
    xr = g.get("FLUID_POS").astype(np.float64)
    gx = (2.0 / xr.size) * (xr - target)
    t0 = time.perf_counter()
    gv0, gmu, gsig = g.adj_backward(gx.astype(np.float32))
    t_back = time.perf_counter() - t0
    print(f"adjoint leg: forward {1000 * t_fwd / steps:.2f} ms/step, rollout {1000 * t_roll / steps:.2f}, "
          f"backward {1000 * t_back / steps:.2f}", file=sys.stderr, flush=True)
    return {
        "particles": n, "steps": steps, "checkpoint_every": every, "loss": loss(xr),
        "forward_ms_per_step": 1000 * t_fwd / steps, "rollout_ms_per_step": 1000 * t_roll / steps,
        "backward_ms_per_step": 1000 * t_back / steps, "adjoint_over_forward": t_back / t_fwd,
        "rollout_plus_backward_over_forward": (t_roll + t_back) / t_fwd,
        "l_cg_mean": float(counts[:, 0, 1].mean()),
        "gradients": {"viscosity": float(gmu[0]), "surface_tension": float(gsig[0]),
                      "v0_norm": float(np.linalg.norm(gv0))},
        "gradient_check": "tests/test_gpu_adjoint.py: central differences (eps 1e-3) of the FP64 oracle forward",
        "solver_tuning": tuning,
    }


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--settle", type=int, default=500,
                    help="untimed steps from the initial state before the timed window (past the start-up transient)")
    ap.add_argument("--n-side", type=int, default=100, help="fluid block edge in particles (default 100 -> 1M)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=12.0, help="seconds of CPU oracle work (all threads)")
    ap.add_argument("--cpu-max-steps", type=int, default=60)
    ap.add_argument("--no-profile", action="store_true", help="do not time kernels with HIP events")
    ap.add_argument("--profile-steps", type=int, default=0,
                    help="steps of the profiled replay that times every kernel (0 = as many as --steps)")
    ap.add_argument("--kernels", action="store_true", help="print the per-kernel table to stderr")
    ap.add_argument("--batch-scenes", type=int, default=64, help="scenes of the batched leg (0 = skip it)")
    ap.add_argument("--batch-steps", type=int, default=300, help="timed steps of the batched leg")
    ap.add_argument("--adjoint-particles", type=int, default=65536, help="droplet of the adjoint leg (0 = skip it)")
    ap.add_argument("--adjoint-steps", type=int, default=100)
    ap.add_argument("--graphs", default="auto", choices=("auto", "0", "1"),
                    help="replay solver iteration chunks from HIP graphs: auto = the library's autotuner "
                         "picks its launch settings (graphs, host-poll interval, operator form, folds, "
                         "persistent solver) in every leg, for the dam break from the settled state; "
                         "0 or 1 fixes graph replay and runs no autotuner")
    ap.add_argument("--cg-stream-coeffs", type=int, default=None, choices=(0, 1),
                    help="viscosity operator: 1 streams both pair coefficients, 0 streams a_ij and recomputes "
                         "b_ij (default: the library's configuration default)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="launch the library's plain copy kernel (k_calib_copy) 3 times on 512 MiB right before "
                         "and right after the timed window: hardware byte counters are scaled against a copy of "
                         "known size, and the copies mark the window in a kernel trace")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py measures one GPU only (this build has no multi-GPU path); got --gpus {args.gpus}",
              file=sys.stderr)
        return 2

    from mps_cuda_amd import HipSim, scenes

    scene = scenes.dam_break_1m(n_side=args.n_side, steps=args.steps)
    cfg = scenes.make_config(scene)
    if args.graphs != "auto":
        cfg.use_graphs = int(args.graphs)
    if args.cg_stream_coeffs is not None:
        cfg.cg_stream_coeffs = args.cg_stream_coeffs
    sim = HipSim(cfg)
    scenes.load_into(sim, scene)
    for _ in range(args.warmup):
        sim.step()
    scenes.load_into(sim, scene)
    for _ in range(args.settle):
        sim.step()
    # the launch settings timed on the state the window starts from (restored after every trial)
    tuned = sim.autotune(2) if args.graphs == "auto" else None
    kept = snapshot(sim)
    say(f"settled: {args.settle} untimed steps, frame {kept['frame']}")

    def run(steps: int, profile: bool, mark: bool = False):
        """Steps settle..settle+steps-1 of the workload from the kept state: (seconds, per-step
        stats, halo stats, per-kernel HIP-event times when profiling)."""
        restore(sim, scene, kept)
        if mark:
            sim.calibration_copy(COPY_BYTES, 3)
        sim.synchronize()
        sim.profile(profile)
        stats, halo = [], []
        sim.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            st = sim.step()
            stats.append((st.l_df, st.l_cd, st.l_cg, st.nnz_ff, st.nnz_fb, st.nnz_bb, st.hit_cap_cd,
                          st.err_cd[max(0, min(st.l_cd, 128) - 1)], st.nonfinite, st.n_boundary))
            halo.append((st.halo_max, st.halo_spill, st.halo_blocks))
        sim.synchronize()
        elapsed = time.perf_counter() - t0
        if mark:
            sim.calibration_copy(COPY_BYTES, 3)
            sim.synchronize()
        prof = sim.profile_read() if profile else {}
        sim.profile(False)
        return elapsed, stats, halo, prof

    N = scene.n
    # the throughput: no per-kernel events (they add ~1.5 ms per step on this workload)
    elapsed, stats, halo, _ = run(args.steps, False, args.copy_calibration)
    value = N * args.steps / elapsed
    say(f"timed region: {args.steps} steps, {1000.0 * elapsed / args.steps:.3f} ms per step")
    # per-kernel times: the same steps again from the initial state, with HIP events around
    # every kernel on the context's stream
    prof_steps = 0 if args.no_profile else (args.profile_steps or args.steps)
    prof, pstats, p_elapsed = {}, [], 0.0
    copy_gbps = None
    if prof_steps:
        p_elapsed, pstats, _, prof = run(prof_steps, True)
        say(f"profiled replay: {prof_steps} steps, {1000.0 * p_elapsed / prof_steps:.3f} ms per step")
        # a STREAM-style copy in the same run (the library's k_stream_copy: one 16-byte vector
        # per thread, non-temporal): what this device's memory system delivers to a copy
        sim.profile(True)
        sim.calibration_copy(COPY_BYTES, 5)
        cp = sim.profile_read().get("k_stream_copy")
        sim.profile(False)
        if cp and cp["total_ms"] > 0:
            copy_gbps = 2.0 * COPY_BYTES / (cp["total_ms"] * 1e-3 / cp["launches"]) / 1e9
            say(f"copy kernel: {copy_gbps:.0f} GB/s")

    cap = cfg.max_iter
    C = cfg.hash_size[0] * cfg.hash_size[1] * cfg.hash_size[2]
    B = stats[0][9] if stats else 0
    tot_step_bytes = 0.0
    kbytes = {}
    for (l_df, l_cd, l_cg, eff, efb, ebb, _, _, _, _) in stats:
        tot_step_bytes += step_bytes(N, B, eff, efb, ebb, l_df, l_cd, l_cg, cap, C)
    for (l_df, l_cd, l_cg, eff, efb, ebb, _, _, _, _) in pstats:
        uses = {
            "k_cg_apply<1>": max(min(l_cg - 1, cfg.cg_max_iter), 0), "k_cg_apply<0>": 1,
            "k_pressure_eval<DF>": min(l_df, cap), "k_pressure_eval<CD>": min(l_cd, cap),
            "k_pressure_apply": (l_df - 1) + (l_cd - 1), "k_density_alpha": 1, "k_st_factor": 1,
            "k_cg_setup": 1, "k_neighbors_fluid": 1, "k_advect": 1,
        }
        for name, u in uses.items():
            kb = kernel_bytes(name, N, eff, efb)
            kbytes[name] = kbytes.get(name, 0.0) + u * kb

    # the viscosity operator that ran: both pair coefficients streamed, or b_ij recomputed
    stream_coeffs = int(tuned["cg_stream_coeffs"]) if tuned else int(cfg.cg_stream_coeffs)
    recomputed = stream_coeffs == 0
    roofline = None
    kernels = {}
    for name, p in prof.items():
        k = dict(p)
        k["avg_us"] = 1000.0 * p["total_ms"] / max(p["launches"], 1)
        if name in kbytes and p["total_ms"] > 0:
            k["algorithmic_GBps"] = kbytes[name] / (p["total_ms"] * 1e-3) / 1e9
        kernels[name] = k
    if prof:
        dom = max(prof.items(), key=lambda kv: kv[1]["total_ms"])[0]
        if dom not in kbytes:
            dom = max((kv for kv in prof.items() if kv[0] in kbytes), key=lambda kv: kv[1]["total_ms"])[0]
        p = prof[dom]
        achieved = kbytes[dom] / (p["total_ms"] * 1e-3) / 1e9
        roofline = {
            "kernel": dom, "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": measured_traffic(dom, [args.settle, args.settle + args.steps - 1],
                                        ("true" if recomputed else "false") if dom.startswith("k_cg_apply<") else ""),
            "traffic_source": os.path.relpath(TRAFFIC_PROFILE, ROOT), "traffic_window": traffic_window(),
            "algorithmic_bytes_per_launch": kbytes[dom] / max(p["useful"], 1),
            "avg_launch_us": 1000.0 * p["total_ms"] / max(p["launches"], 1),
            "launches": p["launches"], "useful_launches": p["useful"],
            "share_of_gpu_time": p["total_ms"] / max(sum(q["total_ms"] for q in prof.values()), 1e-9),
            "profiled_steps": prof_steps, "profiled_ms_per_step": 1000.0 * p_elapsed / prof_steps,
            "step_algorithmic_GBps": tot_step_bytes / elapsed / 1e9,
            "step_frac": tot_step_bytes / elapsed / 1e9 / HBM_PEAK_GBS,
            "measured_copy_GBps": copy_gbps,
            "frac_of_measured_copy": achieved / copy_gbps if copy_gbps else None,
        }

    sim.profile(False)
    sim.close()
    batched = batched_leg(args.batch_scenes, args.batch_steps, args.warmup, graphs=args.graphs) if args.batch_scenes > 0 else None
    if batched:
        say(f"batched leg: {batched.get('speedup', batched)}")
    adjoint = adjoint_leg(args.adjoint_particles, args.adjoint_steps, 10, graphs=args.graphs) if args.adjoint_particles > 0 else None
    if adjoint:
        say("adjoint leg done")

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(scene, kept, args.cpu_budget, args.cpu_max_steps)
        say(f"cpu baseline: {cpu['value']:.3e} particle-steps/s")

    n = len(stats)
    iters = {
        "l_df_mean": sum(s[0] for s in stats) / n, "l_cd_mean": sum(s[1] for s in stats) / n,
        "l_cg_mean": sum(s[2] for s in stats) / n, "cd_capped_steps": sum(s[6] for s in stats),
        "cd_final_err_mean": sum(s[7] for s in stats) / n, "nonfinite_steps": sum(s[8] for s in stats),
        "nnz_ff_last": stats[-1][3], "nnz_fb_last": stats[-1][4], "nnz_bb": stats[-1][5],
        # LDS halos of the viscosity operator: largest staged halo, and the share of 256-slot
        # blocks over the LDS cap (those gather from global memory)
        "halo_max": max(h[0] for h in halo),
        "halo_spill_frac": sum(h[1] for h in halo) / max(sum(h[2] for h in halo), 1),
    }
    out = {
        "metric": METRIC, "value": value, "unit": "particle-steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": 1000.0 * elapsed / args.steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {
            "workload": f"plain DFSPH dam break, {N} fluid + {B} wall particles, 0.01 m lattice, "
                        f"3.0x1.5x1.5 m tank, cubic kernel h=0.02 m, dt=0.5 ms, CD tol 1e-3, DF tol 1e-2, "
                        f"viscosity 0.01 + surface tension 0.05 (CG), steps {args.settle}..{args.settle + args.steps - 1} "
                        f"(timed after {args.settle} untimed steps from the initial lattice)",
            "window": [args.settle, args.settle + args.steps - 1],
            "n_fluid": N, "n_boundary": B, "hash_size": list(cfg.hash_size),
            "cg_stream_coeffs": stream_coeffs,
            "solver_tuning": tuned,
        },
        "roofline": roofline, "cpu_baseline": cpu, "iterations": iters, "batched": batched,
        "adjoint": adjoint,
    }
    if args.kernels:
        for name, k in sorted(kernels.items(), key=lambda kv: -kv[1]["total_ms"]):
            print(f"{name:24s} launches={k['launches']:6d} useful={k['useful']:6d} total={k['total_ms']:9.2f} ms "
                  f"avg={k['avg_us']:8.1f} us GB/s={k.get('algorithmic_GBps', float('nan')):8.1f}", file=sys.stderr)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
