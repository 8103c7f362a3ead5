"""GPU (gfx950) feature coverage against the FP32 oracle: boundary sampling, kernel and
solver variants, hash wrap-around, neighbour-capacity regrow, failure paths, readback,
and the 1M-particle workload's hash and neighbour lists (bit-exact at full size)."""
import sys

import numpy as np
import pytest

from mps_cuda_amd import HipSim, MPSError, scenes
from mps_cuda_amd import _capi as capi
from oracle.binding import OracleSim
from tests.parity import INT_ARRAYS, PHASE_ORDER, REAL_ARRAYS, make_pair, rel_err

pytestmark = pytest.mark.gpu


def say(m):
    print(m, flush=True)
    sys.stdout.flush()


def _phases_match(g, o, label, steps=2, tol=1e-5):
    for s in range(steps):
        for ph in PHASE_ORDER:
            sg, so = g.run_phase(ph), o.run_phase(ph)
            for name in INT_ARRAYS.get(ph, []):
                a, b = g.get(name), o.get(name)
                assert a.shape == b.shape and np.array_equal(a, b), (label, s, ph, name)
            for name in REAL_ARRAYS.get(ph, []):
                e = rel_err(g.get(name), o.get(name))
                assert e <= tol, (label, s, ph, name, e)
            if ph == "DIVERGENCE":
                assert sg.l_df == so.l_df
            if ph == "VISCOSITY":
                assert sg.l_cg == so.l_cg
            if ph == "CONSTANT_DENSITY":
                assert sg.l_cd == so.l_cd
        g.set("FLUID_POS", o.get("FLUID_POS"))
        g.set("FLUID_VEL", o.get("FLUID_VEL"))


def test_boundary_sampling_bit_exact():
    for sc in (scenes.dam_break_cpu(n_side=8), scenes.surface_tension_test_scene()):
        g, o = make_pair(sc)
        for name in ("BND_POS", "BND_RADIUS", "BND_FACE_ID"):
            assert np.array_equal(g.get(name), o.get(name)), (sc.name, name)
        np.testing.assert_allclose(g.get("BND_MASS"), o.get("BND_MASS"), rtol=1e-6)
        say(f"[sampling] {sc.name}: {g.get('BND_POS').shape[0]} boundary particles identical")


@pytest.mark.parametrize("kernel", [capi.KERNEL_POLY6_SPIKY, capi.KERNEL_CUBIC])
def test_kernel_families(kernel):
    sc = scenes.dam_break_cpu(n_side=10)
    sc.config["kernel"] = kernel
    g, o = make_pair(sc)
    _phases_match(g, o, f"kernel{kernel}")


def test_jacobi_solver():
    sc = scenes.dam_break_cpu(n_side=10)
    sc.config["visc_solver"] = capi.VISC_JACOBI
    g, o = make_pair(sc)
    _phases_match(g, o, "jacobi")


def test_non_faithful_boundary_volume():
    sc = scenes.surface_tension_test_scene()
    sc.config["faithful_quirks"] = 0
    g, o = make_pair(sc)
    for s in range(55):  # crosses the Z-sort at frames 1 and 51
        sg, so = g.step(), o.step()
        assert (sg.l_df, sg.l_cg, sg.l_cd) == (so.l_df, so.l_cg, so.l_cd), s
    assert rel_err(g.get("BND_VOLUME"), o.get("BND_VOLUME")) < 1e-6
    assert rel_err(g.get("FLUID_POS"), o.get("FLUID_POS")) < 1e-5


def test_hash_smaller_than_domain():
    sc = scenes.dam_break_cpu(n_side=10)
    sc.config["hash_size"] = (8, 4, 8)  # cells alias across the 1 m box
    g, o = make_pair(sc)
    _phases_match(g, o, "wrap", steps=1)


def test_search_global_fallback_coarse_cells():
    """Cells twice the support radius put ~8x more candidates in every 27-cell neighbourhood,
    more than the search stages in LDS, so blocks search the global cell tables: same lists."""
    sc = scenes.dam_break_cpu(n_side=12)
    sc.config["cell_size"] = 2.0 * sc.config["cell_size"]
    g, o = make_pair(sc)
    _phases_match(g, o, "coarse", steps=1)


def test_neighbour_capacity_regrow():
    sc = scenes.dam_break_cpu(n_side=10)
    sc.config["nei_capacity"] = 8
    g, o = make_pair(sc)
    st = g.step()
    o.step()
    assert st.nei_regrow >= 1
    for name in ("NEI_FF", "NEI_FB", "NEI_BB", "NEI_FF_IDX"):
        assert np.array_equal(g.get(name), o.get(name)), name


def test_iteration_cap_and_nonfinite():
    sc = scenes.dam_break_cpu(n_side=8)
    sc.config["cd_tol"] = 0.0
    sc.config["max_iter"] = 6
    g, o = make_pair(sc)
    sg, so = g.step(), o.step()
    assert (sg.l_cd, sg.hit_cap_cd) == (7, 1) == (so.l_cd, so.hit_cap_cd)
    v = g.get("FLUID_VEL")
    v[3, 0] = np.nan
    g.set("FLUID_VEL", v)
    with pytest.raises(MPSError) as ei:
        g.step()
    assert ei.value.code == capi.MPS_ERR_NONFINITE


def test_empty_fluid():
    sc = scenes.dam_break_cpu(n_side=8)
    g = HipSim(scenes.make_config(sc))
    scenes.load_into(g, sc)
    g.upload_fluid(np.zeros((0, 3)), np.zeros(0))
    st = g.step()
    assert st.n_fluid == 0 and st.n_boundary > 0


def test_readback_async():
    sc = scenes.dam_break_cpu(n_side=8)
    g, o = make_pair(sc)
    for _ in range(3):
        g.step()
    n = sc.n
    pos = np.empty((n, 3), np.float32)
    rad = np.empty(n, np.float32)
    rgba = np.empty((n, 4), np.float32)
    f = g.readback_async(pos, rad, rgba)
    g.fence_wait(f)
    assert np.array_equal(pos, g.get("FLUID_POS"))
    assert np.array_equal(rad, g.get("FLUID_RADIUS"))
    assert np.array_equal(rgba, g.get("FLUID_COLOR"))


def test_readback_overlaps_the_next_step():
    """Frame N is packed on the compute stream and copied on the context's copy stream while step
    N+1 runs: on the 1M dam break, frame N's data arrive intact although the next step moves
    every particle, two frames are in flight at once (two staging slots), the renderer's double
    layout equals the float one widened, and readback + step take less wall time than readback,
    wait, then step (minimum over five repetitions of each)."""
    import time
    sc = scenes.dam_break_1m()
    g = HipSim(scenes.make_config(sc))
    scenes.load_into(g, sc)
    for _ in range(3):
        g.step()
    n = sc.n
    bufs = [(g.host_array((n, 3), np.float32), g.host_array((n,), np.float32), g.host_array((n, 4), np.float32))
            for _ in range(2)]
    want = []
    fences = []
    for k in range(2):
        want.append((g.get("FLUID_POS"), g.get("FLUID_RADIUS"), g.get("FLUID_COLOR")))
        fences.append(g.readback_async(*bufs[k]))
        g.step()
    for k in range(2):
        g.fence_wait(fences[k])
        for got, w in zip(bufs[k], want[k]):
            assert np.array_equal(got, w), k
    assert not np.array_equal(want[0][0], want[1][0])
    d = (g.host_array((n, 3), np.float64), g.host_array((n,), np.float64), g.host_array((n, 4), np.float32))
    g.fence_wait(g.readback_async_f64(*d))
    assert np.array_equal(d[0], g.get("FLUID_POS").astype(np.float64))
    assert np.array_equal(d[1], g.get("FLUID_RADIUS").astype(np.float64))
    assert np.array_equal(d[2], g.get("FLUID_COLOR"))

    def timed(fn):
        best = 1e30
        for _ in range(5):
            g.synchronize()
            t0 = time.perf_counter()
            fn()
            g.synchronize()
            best = min(best, time.perf_counter() - t0)
        return best

    t_copy = timed(lambda: g.fence_wait(g.readback_async(*bufs[0])))
    t_serial = timed(lambda: (g.fence_wait(g.readback_async(*bufs[0])), g.step()))
    t_overlap = timed(lambda: (lambda f: (g.step(), g.fence_wait(f)))(g.readback_async(*bufs[0])))
    say(f"[readback] 1M frame: copy {1e3 * t_copy:.2f} ms, copy then step {1e3 * t_serial:.2f} ms, "
        f"copy during step {1e3 * t_overlap:.2f} ms")
    assert t_overlap < t_serial - 0.3 * t_copy


def test_full_size_hash_and_neighbours_bit_exact():
    """The 1,000,000-particle workload: sorted ids, cell tables and all three neighbour
    lists identical to the oracle's, plus the first step's densities and alphas."""
    sc = scenes.dam_break_1m()
    g, o = make_pair(sc)
    say(f"[full] N={sc.n} B={g.get('BND_POS').shape[0]}")
    for ph in ("HASH", "NEIGHBORS", "BOUNDARY_VOLUME", "DENSITY", "ALPHA"):
        g.run_phase(ph)
        o.run_phase(ph)
    for name in INT_ARRAYS["HASH"] + INT_ARRAYS["NEIGHBORS"]:
        a, b = g.get(name), o.get(name)
        assert a.shape == b.shape and np.array_equal(a, b), name
    for name in ("BND_VOLUME", "FLUID_DENSITY", "FLUID_ALPHA"):
        assert rel_err(g.get(name), o.get(name)) < 1e-6, name
    # symmetry of the fluid lists (equal radii) as a size-independent property
    idx, nei = g.get("NEI_FF_IDX"), g.get("NEI_FF")
    rows = np.repeat(np.arange(sc.n), np.diff(idx))
    fwd = np.sort(rows.astype(np.int64) * sc.n + nei)
    bwd = np.sort(nei.astype(np.int64) * sc.n + rows)
    assert np.array_equal(fwd, bwd)
    say(f"[full] nnz_ff={len(nei)} identical and symmetric")


def test_full_size_step_parity():
    """One whole step of the 1,000,000-particle workload, phase by phase against the FP32
    oracle: integer arrays identical, every float array within 1e-5, and the DF, CG and CD
    loops ending at the same iteration (the halo-staged and LDS-staged kernels at their real
    sizes)."""
    import time
    sc = scenes.dam_break_1m()
    g, o = make_pair(sc)
    t0 = time.perf_counter()
    _phases_match(g, o, "full step", steps=1)
    say(f"[full] one 1M step phase by phase matches the oracle ({time.perf_counter() - t0:.1f} s)")


def test_graph_replay_matches_direct_launches():
    """Solver chunks replayed from cached HIP graphs give the same bits as direct launches,
    step after step (each replay follows eager hash / neighbour work and host syncs)."""
    sc = scenes.dam_break_cpu(n_side=10)
    a = HipSim(scenes.make_config(sc))
    cfg = scenes.make_config(sc)
    cfg.use_graphs = 1
    b = HipSim(cfg)
    scenes.load_into(a, sc)
    scenes.load_into(b, sc)
    for s in range(30):
        sa, sb = a.step(), b.step()
        assert (sa.l_df, sa.l_cg, sa.l_cd) == (sb.l_df, sb.l_cg, sb.l_cd), s
        assert np.array_equal(a.get("FLUID_POS"), b.get("FLUID_POS")), s
        assert np.array_equal(a.get("FLUID_VEL"), b.get("FLUID_VEL")), s


def test_autotune_keeps_state_and_results():
    sc = scenes.dam_break_cpu(n_side=10)
    a, b = HipSim(scenes.make_config(sc)), HipSim(scenes.make_config(sc))
    scenes.load_into(a, sc)
    scenes.load_into(b, sc)
    for _ in range(3):
        a.step()
        b.step()
    res = b.autotune(2)
    say(f"[autotune] {res}")
    assert len(res["ms_per_step"]) == 8 and b.frame == a.frame
    for _ in range(5):
        a.step()
        b.step()
    assert np.array_equal(a.get("FLUID_POS"), b.get("FLUID_POS"))


@pytest.mark.parametrize("limit", [1, 300])
def test_viscosity_halo_paths_bit_identical(limit):
    """The pressure solves and the viscosity operator stage each 256-slot block's neighbour
    records in LDS; a block whose halo is over the limit gathers from global memory instead.
    Both paths visit the same pairs in the same order: limit 1 (every block global) and 300
    give the same bits as the default, step after step."""
    sc = scenes.dam_break_cpu(n_side=12)
    a = HipSim(scenes.make_config(sc))
    cfg = scenes.make_config(sc)
    cfg.halo_limit = limit
    b = HipSim(cfg)
    scenes.load_into(a, sc)
    scenes.load_into(b, sc)
    spilled = 0
    for s in range(12):
        sa, sb = a.step(), b.step()
        assert sa.halo_spill == 0 and 0 < sa.halo_max <= 1536 and sa.halo_blocks > 0, s
        assert sb.halo_blocks == sa.halo_blocks and sb.halo_max <= limit, s
        if limit == 1:
            assert sb.halo_spill == sb.halo_blocks, s
        spilled += sb.halo_spill
        assert (sa.l_df, sa.l_cg, sa.l_cd) == (sb.l_df, sb.l_cg, sb.l_cd), s
        assert np.array_equal(a.get("FLUID_POS"), b.get("FLUID_POS")), s
        assert np.array_equal(a.get("FLUID_VEL"), b.get("FLUID_VEL")), s
    say(f"[halo] limit {limit}: default max halo {sa.halo_max}, {spilled} block-steps on the global path")
    assert spilled > 0


@pytest.mark.parametrize("kernel", [capi.KERNEL_CUBIC, capi.KERNEL_POLY6_SPIKY])
def test_viscosity_operator_variants_bit_identical(kernel):
    """The viscosity operator either streams both pair coefficients or streams a_ij and
    recomputes b_ij from the staged positions, masses and 1/rho (cg_stream_coeffs 1 / 0), on the
    staged and the global path (halo limit 1): the same bits step after step, with unequal
    particle masses so that the 0.5 (m_i + m_j) factor is exercised."""
    sc = scenes.dam_break_cpu(n_side=12)
    sc.config["kernel"] = kernel
    mass = np.random.default_rng(4).uniform(0.8, 1.2, sc.n).astype(np.float32) * np.float32(
        1000.0 * (0.04 ** 3) / 48.0 * 3.141592)
    sims = []
    for stream, limit in ((1, 0), (0, 0), (0, 1)):
        cfg = scenes.make_config(sc)
        cfg.cg_stream_coeffs = stream
        cfg.halo_limit = limit
        g = HipSim(cfg)
        scenes.load_into(g, sc)
        g.set("FLUID_MASS", mass)
        sims.append(g)
    for s in range(10):
        st = [g.step() for g in sims]
        assert len({(t.l_df, t.l_cg, t.l_cd) for t in st}) == 1, s
        for g in sims[1:]:
            assert np.array_equal(sims[0].get("FLUID_VEL"), g.get("FLUID_VEL")), s
            assert np.array_equal(sims[0].get("FLUID_POS"), g.get("FLUID_POS")), s
    say(f"[operator] kernel {kernel}: streamed and recomputed b_ij identical over 10 steps, l_cg {st[0].l_cg}")


def test_viscosity_halo_depths_bit_identical():
    """The streamed viscosity operator stages block halos up to 1264 records (shallow) or 1536
    (deep), or picks the depth from the previous list build (cg_halo_depth 1 / 2 / 0). On a
    lattice compressed to 0.39 h spacing many blocks have halos between the two depths, which
    only the deep form stages; the others gather them from global memory. Same bits, step after
    step."""
    sc = scenes.dam_break_cpu(n_side=24)
    p0 = sc.pos.min(axis=0)
    sc.pos = (p0 + (sc.pos - p0) * 0.78).astype(np.float32).astype(np.float64)
    sims = []
    for depth in (1, 2, 0):
        cfg = scenes.make_config(sc)
        cfg.cg_halo_depth = depth
        g = HipSim(cfg)
        scenes.load_into(g, sc)
        sims.append(g)
    for s in range(3):
        st = [g.step() for g in sims]
        assert st[0].halo_max > 1264, (s, st[0].halo_max)
        assert len({(t.l_df, t.l_cg, t.l_cd) for t in st}) == 1, s
        for g in sims[1:]:
            assert np.array_equal(sims[0].get("FLUID_VEL"), g.get("FLUID_VEL")), s
            assert np.array_equal(sims[0].get("FLUID_POS"), g.get("FLUID_POS")), s
    say(f"[halo depth] largest halo {st[0].halo_max}, l_cg {st[0].l_cg}: shallow, deep and chosen identical over 3 steps")


def test_checkpoint_resume_is_bit_identical():
    """Snapshot at frame 30 (after the Z-sort of frame 1) through the C ABI, restore into a fresh
    context and run across the Z-sort of frame 51: the same positions and loop counts as the
    uninterrupted run."""
    from tests.resume import run_resume
    sc = scenes.dam_break_cpu(n_side=10)

    def make(load=True):
        g = HipSim(scenes.make_config(sc))
        if load:
            scenes.load_into(g, sc)
        return g

    say("[features] checkpoint/resume")
    xa, xb, la, lb = run_resume(make, sc, 30, 56)
    assert la == lb and any(z for *_, z in la)
    assert np.array_equal(xa, xb)


def test_full_size_free_run():
    """The 1,000,000-particle workload free-running for 20 steps on the GPU and in the FP32
    oracle (16 host threads) from the same state: equal DF, CD and CG loop counts every step,
    positions and velocities within 1e-5 after the last one."""
    import time
    from oracle.binding import set_threads
    set_threads(0)
    sc = scenes.dam_break_1m()
    g, o = make_pair(sc)
    t0 = time.perf_counter()
    loops = []
    for s in range(20):
        sg, so = g.step(), o.step()
        assert (sg.l_df, sg.l_cd, sg.l_cg, sg.zsorted) == (so.l_df, so.l_cd, so.l_cg, so.zsorted), s
        loops.append(sg.l_cg)
    errs = {n: rel_err(g.get(n), o.get(n)) for n in ("FLUID_POS", "FLUID_VEL")}
    same = all(np.array_equal(g.get(n), o.get(n)) for n in ("FLUID_POS", "FLUID_VEL"))
    say(f"[full] 20 free steps of the 1M workload in {time.perf_counter() - t0:.1f} s: CG loops {loops}, "
        f"rel errs {errs}, bitwise equal {same}")
    assert max(errs.values()) <= 1e-5, errs
