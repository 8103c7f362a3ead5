"""Spatial hash and neighbour lists of the oracle (MPSSpatialHash.cuh:9-271,
MPSSpatialHash.cu:53-218) pinned by hand-computed keys and brute force."""
import numpy as np
import pytest

from mps_cuda_amd import _capi as capi
from oracle.binding import OracleSim, cell_keys

EMPTY = 0xFFFFFFFF


def test_cell_key_known_answers():
    H = [64, 64, 64]
    # x: trunc(1.25) = 1 (+32); y: 0 (+32); z: trunc(-1.25) = -1 toward zero (+32)
    k = cell_keys(64, [[0.05, 0.0, -0.05]], 0.04, H)[0]
    assert k == (31 * 64 + 32) * 64 + 33
    # wrap: trunc(2.0/0.04)=50 -> 82 & 63 = 18
    k = cell_keys(64, [[2.0, 0.0, 0.0]], 0.04, H)[0]
    assert k == (32 * 64 + 32) * 64 + 18
    # negative beyond one cell: trunc(-2.5) = -2 -> 30
    k = cell_keys(64, [[-0.1, 0.0, 0.0]], 0.04, H)[0]
    assert k == (32 * 64 + 32) * 64 + 30


def _interleave(x, y, z, bits=10):
    out = 0
    for b in range(bits):
        out |= ((x >> b) & 1) << (3 * b) | ((y >> b) & 1) << (3 * b + 1) | ((z >> b) & 1) << (3 * b + 2)
    return out


def test_morton_key_known_answers():
    H = [64, 64, 64]
    # grid pos (33, 32, 31); Morton adds H/2 again: (65&63, 64&63, 63&63) = (1, 0, 63)
    k = cell_keys(64, [[0.05, 0.0, -0.05]], 0.04, H, morton=True)[0]
    assert k == _interleave(1, 0, 63)
    # the x == 1024 clamp (only reachable with H = 2048): coordinate 1024 -> 1023
    H2 = [2048, 2048, 2048]
    # grid pos x = trunc(0.0) + 1024 = 1024; Morton x = (1024 + 1024) & 2047 = 0. Use x such
    # that the Morton coordinate is 1024: grid x = 0 -> x = -1024 cells
    k = cell_keys(64, [[-1024 * 0.04 - 0.001, 0.0, 0.0]], 0.04, H2, morton=True)[0]
    assert (k & 0x09249249) == _interleave(1023, 0, 0)


def _scene_sim(pos, radius, H=(16, 16, 16), cell=0.04, prec=32, boundary=None):
    cfg = capi.default_config()
    for i in range(3):
        cfg.hash_size[i] = H[i]
    cfg.cell_size = cell
    o = OracleSim(cfg, prec)
    o.set_physics(1e-3, (0, -9.81, 0), (-10, -10, -10), (10, 10, 10))
    o.set_fluid_material(cell, 1000.0, 0.01, 0.1, 0.05)
    if boundary is not None:
        o.upload_boundary(*boundary)
    o.upload_fluid(pos, radius)
    o.run_phase("HASH")
    o.run_phase("NEIGHBORS")
    return o


def test_hash_table_matches_brute_force():
    rng = np.random.default_rng(1)
    pos = rng.uniform(-0.2, 0.2, size=(700, 3)).astype(np.float32).astype(np.float64)
    o = _scene_sim(pos, 0.01)
    keys, ids = o.get("FHASH_KEY"), o.get("FHASH_ID")
    start, end = o.get("FHASH_START"), o.get("FHASH_END")
    ref = cell_keys(32, pos, 0.04, [16, 16, 16])
    assert np.array_equal(np.sort(ids), np.arange(len(pos)))
    assert np.array_equal(keys, ref[ids])
    # stable: keys non-decreasing, ties by ascending id
    order = np.lexsort((np.arange(len(pos)), ref))
    assert np.array_equal(ids, order)
    for c in range(16 ** 3):
        members = np.nonzero(keys == c)[0]
        if members.size:
            assert start[c] == members[0] and end[c] == members[-1] + 1
        else:
            assert start[c] == EMPTY and end[c] == EMPTY


def _brute_lists(pos32, rad32, cell, H, same=True, rpos32=None, rrad32=None):
    """Reference neighbour lists: set by the float32 glm-order distance test, order by the
    27-cell (z, y, x) visit order of the candidate's wrapped cell, then ascending id."""
    if rpos32 is None:
        rpos32, rrad32 = pos32, rad32
    kc = lambda p: np.trunc(p.astype(np.float64) * (1.0 / np.float64(np.float32(cell)))).astype(np.int64) + np.array(H) // 2
    ci, cj = kc(pos32), kc(rpos32)
    out = []
    for i in range(len(pos32)):
        d = pos32[i][None, :] - rpos32
        dd = (d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2]
        dist = np.sqrt(dd.astype(np.float32))
        thr = (rad32[i] * np.float32(4) + rrad32 * np.float32(4)) * np.float32(0.5)
        cand = np.nonzero(dist < thr)[0]
        if same:
            cand = cand[cand != i]
        # only the 27 (wrapped) cells around i's cell are searched: offsets -1, 0, +1 -> 0, 1, 2
        rel = [tuple(((cj[j] - ci[i] + 1) % np.array(H))[::-1]) for j in cand]  # (dz, dy, dx)
        keep = [t for t in range(len(cand)) if max(rel[t]) <= 2]
        order = sorted(keep, key=lambda t: (rel[t], cand[t]))
        out.append([int(cand[t]) for t in order])
    return out


@pytest.mark.parametrize("H", [(16, 16, 16), (4, 4, 4)])
def test_neighbour_lists_match_brute_force(H):
    rng = np.random.default_rng(2)
    pos = rng.uniform(-0.15, 0.15, size=(400, 3)).astype(np.float32)
    rad = np.full(400, 0.01, np.float32)
    rad[::7] = 0.0125  # some unequal supports
    o = _scene_sim(pos.astype(np.float64), rad.astype(np.float64), H=H)
    idx, nei = o.get("NEI_FF_IDX"), o.get("NEI_FF")
    want = _brute_lists(pos, rad, 0.04, H)
    for i in range(len(pos)):
        assert list(nei[idx[i]:idx[i + 1]]) == want[i], i


def test_fluid_boundary_lists_and_symmetry():
    rng = np.random.default_rng(3)
    pos = rng.uniform(-0.1, 0.1, size=(300, 3)).astype(np.float32)
    bpos = rng.uniform(-0.12, 0.12, size=(200, 3)).astype(np.float32)
    rad = np.full(300, 0.01, np.float32)
    brad = np.full(200, 0.01, np.float32)
    o = _scene_sim(pos.astype(np.float64), rad.astype(np.float64), boundary=(bpos.astype(np.float64), brad.astype(np.float64)))
    idx, nei = o.get("NEI_FB_IDX"), o.get("NEI_FB")
    want = _brute_lists(pos, rad, 0.04, (16, 16, 16), same=False, rpos32=bpos, rrad32=brad)
    for i in range(len(pos)):
        assert list(nei[idx[i]:idx[i + 1]]) == want[i]
    # equal radii: j in N(i) <=> i in N(j)
    fi, fn = o.get("NEI_FF_IDX"), o.get("NEI_FF")
    sets = [set(fn[fi[i]:fi[i + 1]].tolist()) for i in range(len(pos))]
    for i in range(len(pos)):
        for j in sets[i]:
            assert i in sets[j]
