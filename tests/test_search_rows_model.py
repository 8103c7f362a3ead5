"""CPU model of the staged fluid search's indexing (k_neighbors_fluid in hash.hip): the block's
rows (y, z) around its cell runs, each row's cells numbered densely in x order over the x range
its runs need, candidates staged cell after cell, and each particle's 27 cells walked as 9 row
runs. The runs must give the reference's 27-cell walk (z, y, x offsets, then slot order within a
cell) for every particle of a staged block, including when the hash wraps around in x, y or z.
The distance test is the reference's; what is checked is which candidates are visited, in which
order, and that the particle's own staged copy is the one the kernel skips."""
import numpy as np
import pytest


def _morton(v):
    v = v.astype(np.uint64) & np.uint64(0x3FF)
    v = (v | (v << np.uint64(16))) & np.uint64(0x030000FF)
    v = (v | (v << np.uint64(8))) & np.uint64(0x0300F00F)
    v = (v | (v << np.uint64(4))) & np.uint64(0x030C30C3)
    v = (v | (v << np.uint64(2))) & np.uint64(0x09249249)
    return v


def _scene(n_side, H, seed):
    """A jittered lattice at 0.39 h spacing (cell = h = 1), slots sorted by the Morton code of the
    wrapped cell, the hash's cell start/end tables; returns unwrapped cells, keys and tables."""
    rng = np.random.default_rng(seed)
    g = np.stack(np.meshgrid(*(np.arange(n_side) * 0.39,) * 3, indexing="ij"), -1).reshape(-1, 3)
    x = (g - g.mean(0) + rng.uniform(-0.1, 0.1, g.shape)).astype(np.float32)
    H = np.array(H)
    P = np.trunc(x.astype(np.float64)).astype(np.int64) + H // 2   # trunc(x / cell) + H / 2
    W = P & (H - 1)
    key = (W[:, 2] * H[1] + W[:, 1]) * H[0] + W[:, 0]
    m = _morton(W[:, 0]) | (_morton(W[:, 1]) << np.uint64(1)) | (_morton(W[:, 2]) << np.uint64(2))
    o = np.lexsort((np.arange(len(x)), m))
    x, P, key = x[o], P[o], key[o]
    C = int(np.prod(H))
    start = np.full(C, -1, np.int64)
    end = np.full(C, -1, np.int64)
    for s, k in enumerate(key):
        if start[k] < 0:
            start[k] = s
        end[k] = s + 1
    return x, P, H, start, end


def _gidx(H, xx, yy, zz):
    return ((zz & (H[2] - 1)) * H[1] + (yy & (H[1] - 1))) * H[0] + (xx & (H[0] - 1))


def _cell_slots(start, end, c):
    return np.arange(start[c], end[c]) if start[c] >= 0 else np.zeros(0, np.int64)


def _reference_order(P, H, start, end, i):
    """The candidates of particle i in the reference's 27-cell order (self included)."""
    parts = []
    for t in range(27):
        c = _gidx(H, P[i, 0] + t % 3 - 1, P[i, 1] + (t // 3) % 3 - 1, P[i, 2] + t // 9 - 1)
        parts.append(_cell_slots(start, end, c))
    return np.concatenate(parts)


def _block_model(P, H, start, end, ids, caps=(96, 384, 1920)):
    """The kernel's staging for one block of slots `ids`; None when over a cap (global path).
    Returns (candidate slot per staged copy, per particle: its 9 row runs and own copy)."""
    rows = {}
    for i in ids:
        if i > ids[0] and tuple(P[i - 1]) == tuple(P[i]):
            continue  # not the first particle of its cell run
        for r in range(9):
            k = int(_gidx(H, 0, P[i, 1] + r % 3 - 1, P[i, 2] + r // 3 - 1))
            lo, hi = rows.get(k, (1 << 40, -(1 << 40)))
            rows[k] = (min(lo, P[i, 0] - 1), max(hi, P[i, 0] + 1))
    if len(rows) > caps[0]:
        return None
    rbase, cells = {}, []
    for k, (lo, hi) in rows.items():
        rbase[k] = len(cells)
        cells += [k + (int(xx) & (H[0] - 1)) for xx in range(lo, hi + 1)]
    if len(cells) > caps[1]:
        return None
    eo, es, cand = [0], [], []
    for c in cells:
        s = _cell_slots(start, end, c)
        es.append(start[c])
        cand.append(s)
        eo.append(eo[-1] + len(s))
    cand = np.concatenate(cand)
    if len(cand) > caps[2]:
        return None
    walks = {}
    for i in ids:
        def cell_of(dx, y, z):
            k = int(_gidx(H, 0, y, z))
            return rbase[k] + (P[i, 0] + dx - rows[k][0])
        eh = cell_of(0, P[i, 1], P[i, 2])
        own = eo[eh] + (i - es[eh])
        runs = []
        for r in range(9):
            e = cell_of(-1, P[i, 1] + r % 3 - 1, P[i, 2] + r // 3 - 1)
            runs.append((eo[e], eo[e + 3]))
        walks[i] = (runs, own)
    return cand, walks


@pytest.mark.parametrize("n_side,H,seed", [
    (24, (64, 64, 64), 0),   # no wrap-around, negative coordinates
    (24, (8, 16, 8), 1),     # x and z wrap around
    (24, (16, 4, 16), 2),    # y wraps around: rows alias and merge
])
def test_row_runs_reproduce_the_reference_walk(n_side, H, seed):
    x, P, H, start, end = _scene(n_side, H, seed)
    N = len(x)
    staged = checked = 0
    for b0 in range(0, N, 256):
        ids = list(range(b0, min(N, b0 + 256)))
        model = _block_model(P, H, start, end, ids)
        if model is None:
            continue
        staged += 1
        cand, walks = model
        for i in ids:
            runs, own = walks[i]
            assert cand[own] == i
            got = np.concatenate([cand[lo:hi] for lo, hi in runs])
            np.testing.assert_array_equal(got, _reference_order(P, H, start, end, i))
            checked += 1
    assert staged > 0 and checked > 0
