"""CPU tests of the ghost refresh: the library's host-side patch table
(sgr_ghost_table_stats, no GPU needed) and the oracle's statement of the
refresh (oracle/ghost_oracle.cpp).  The reference repository has no ghost
fill, so the oracle is pinned by what the mathematics fixes: exact copies on
aligned grids, exactness on polynomials of degree 5 per axis, an independent
per-node restatement (Vandermonde solves, floating-point geometry) and
sixth-order convergence on smooth data."""
import numpy as np
import pytest

import pyoracle as po
import sympygr_amd as sgr
from sympygr_amd import data

from helpers import coords, ghost_mask, poly5, small_octree


def _expected_nodes(bl):
    tot = 0
    for b in range(len(bl)):
        m, f = bl.sz[b], bl.bflag[b]
        box = 1
        for d in range(3):
            box *= m[d] - 3 * ((f >> (2 * d)) & 1) - 3 * ((f >> (2 * d + 1)) & 1)
        tot += box - (m[0] - 6) * (m[1] - 6) * (m[2] - 6)
    return tot


def test_table_aligned_tiling_is_all_copies():
    bl = data.uniform_tiling(nb=3, n=16, lo=-1.0, hi=1.0, faces=True)
    st = sgr.ghost_table_stats(bl)
    assert st["interpolated"] == 0
    assert st["nodes"] == _expected_nodes(bl)
    # 27 blocks: each has one patch per existing neighbour (26 for the centre)
    assert st["patches"] == sum(
        sum(1 for dz in (-1, 0, 1) for dy in (-1, 0, 1) for dx in (-1, 0, 1)
            if (dx, dy, dz) != (0, 0, 0) and 0 <= bx + dx < 3 and 0 <= by + dy < 3 and 0 <= bz + dz < 3)
        for bz in range(3) for by in range(3) for bx in range(3))


def test_table_octree_is_complete():
    bl = data.octree_blocks()
    st = sgr.ghost_table_stats(bl)
    assert st["nodes"] == _expected_nodes(bl)
    assert 0 < st["interpolated"] < st["nodes"]


def test_table_of_the_benchmark_octree():
    """The 3032-leaf octree of the benchmark: the table is complete and every
    work box is within the refresh kernels' limits (checked while the table
    is built; a violation is reported as an error).  Counts pinned."""
    bl = data.octree_blocks()
    st = sgr.ghost_table_stats(bl)
    assert st == {"patches": 68952, "nodes": 65162112, "interpolated": 49837515}


def test_table_refuses_unowned_ghosts_and_overlaps():
    bl = data.uniform_tiling(nb=2, n=16, lo=-1.0, hi=1.0, faces=False)
    with pytest.raises(sgr.SgrError, match="UNSUPPORTED.*owning block"):
        sgr.ghost_table_stats(bl)
    two = data.BlockList()
    for _ in range(2):
        two.add((22, 22, 22), (-1.0,) * 3, (1.1,) * 3, 63)
    with pytest.raises(sgr.SgrError, match="overlap"):
        sgr.ghost_table_stats(two)
    odd = data.BlockList()
    odd.add((22, 22, 22), (0.0,) * 3, (2.1,) * 3, 63)
    odd.add((22, 22, 22), (2.1 + np.pi * 1e-3,) * 3, (4.2 + np.pi * 1e-3,) * 3, 63)
    with pytest.raises(sgr.SgrError, match="lattice"):
        sgr.ghost_table_stats(odd)


def _scrambled(u, bl, seed=0):
    mk = ghost_mask(bl, inside_only=False)
    v = u.copy()
    v[:, mk] = np.random.default_rng(seed).standard_normal((24, int(mk.sum())))
    return v


def test_oracle_copies_on_aligned_tiling():
    bl = data.uniform_tiling(nb=2, n=16, lo=-1.0, hi=1.0, faces=True)
    u = _scrambled(data.fill_blocks(bl, kind="perturbed", seed=3, amp=1e-2, lo=-1.0, L=2.0), bl)
    np.testing.assert_array_equal(po.ghost_fill(u, bl), po.ghost_fill_uniform(u, bl, 2))


def test_oracle_exact_on_degree5_polynomials():
    bl = small_octree()
    X = coords(bl)
    u = np.stack([poly5(X, f) for f in range(24)])
    v = _scrambled(u, bl)
    g = po.ghost_fill(v, bl)
    inside = ghost_mask(bl)
    beyond = ghost_mask(bl, inside_only=False) & ~inside
    np.testing.assert_array_equal(g[:, ~ghost_mask(bl, inside_only=False)], u[:, ~ghost_mask(bl, inside_only=False)])
    np.testing.assert_array_equal(g[:, beyond], v[:, beyond])
    err = np.abs(g[:, inside] - u[:, inside]).max() / np.abs(u).max()
    assert err < 1e-13, err
    # degree 6 along x is not reproduced: the stencil has exactly 6 nodes
    w = np.stack([(X[0] / 2.0) ** 6 for _ in range(24)])
    e6 = np.abs(po.ghost_fill(_scrambled(w, bl), bl)[:, inside] - w[:, inside]).max()
    assert e6 > 1e-9


def _owner_and_value(bl, u, x):
    """Restatement of one ghost node from the stated rule with floating-point
    geometry and Vandermonde solves (no shared code with the oracle)."""
    owner = None
    for s in range(len(bl)):
        h = [(bl.pmax[s][d] - bl.pmin[s][d]) / (bl.sz[s][d] - 1) for d in range(3)]
        lo = [bl.pmin[s][d] + 2.5 * h[d] for d in range(3)]
        hi = [bl.pmax[s][d] - 2.5 * h[d] for d in range(3)]
        if all(lo[d] <= x[d] < hi[d] for d in range(3)):
            if any(min(abs(x[d] - lo[d]), abs(x[d] - hi[d])) < 1e-9 for d in range(3)):
                return None, None  # on a box face: skip
            owner = (s, h)
    if owner is None:
        return None, None
    s, h = owner
    m = bl.sz[s]
    nodes = []
    for d in range(3):
        p = (x[d] - bl.pmin[s][d]) / h[d]
        if abs(p - round(p)) < 1e-9:
            nodes.append(([int(round(p))], p))
        else:
            q = int(np.floor(p))
            p0 = min(max(q - 2, 3), m[d] - 9)
            nodes.append((list(range(p0, p0 + 6)), p))
    V = u[:, bl.offsets[s]:bl.offsets[s] + m[0] * m[1] * m[2]].reshape(24, m[2], m[1], m[0])
    cube = V[:, nodes[2][0]][:, :, nodes[1][0]][:, :, :, nodes[0][0]]

    def interp(vals, idx, p):  # along the last axis
        if len(idx) == 1:
            return vals[..., 0]
        A = np.vander(np.array(idx, dtype=float) - p, 6, increasing=True)
        return np.linalg.solve(A, vals.reshape(-1, 6).T)[0].reshape(vals.shape[:-1])
    r = interp(cube, nodes[0][0], nodes[0][1])          # (24, nz, ny)
    r = interp(r, nodes[1][0], nodes[1][1])             # (24, nz)
    r = interp(r, nodes[2][0], nodes[2][1])             # (24,)
    return s, r


def test_oracle_matches_independent_restatement():
    bl = small_octree()
    u = data.fill_blocks(bl, kind="perturbed", seed=5, amp=0.3, lo=-2.0, L=4.0)
    g = po.ghost_fill(_scrambled(u, bl), bl)
    X = coords(bl)
    idx = np.flatnonzero(ghost_mask(bl))
    rng = np.random.default_rng(1)
    scale = np.abs(u).max(axis=1)
    checked = 0
    for i in rng.choice(idx, 300, replace=False):
        s, val = _owner_and_value(bl, u, X[:, i])
        if s is None:
            continue
        checked += 1
        assert np.all(np.abs(g[:, i] - val) <= 1e-11 * scale), (i, np.abs(g[:, i] - val) / scale)
    assert checked > 250


def _two_blocks(na, nb):
    """[0,1]x[0,1]^2 with na^3 interior nodes beside [1,2]x[0,1]^2 with nb^3."""
    bl = data.BlockList()
    for lo, n, flag in ((0.0, na, 1 | 4 | 8 | 16 | 32), (1.0, nb, 2 | 4 | 8 | 16 | 32)):
        h = 1.0 / n
        pmin = (lo - 2.5 * h, -2.5 * h, -2.5 * h)
        bl.add((n + 6,) * 3, pmin, tuple(p + (n + 5) * h for p in pmin), flag)
    return bl


def _face_error(na, nb):
    bl = _two_blocks(na, nb)
    X = coords(bl)
    u = np.stack([np.sin(1.3 * X[0] + 0.7 * X[1] - 0.4 * X[2] + 0.1 * f) for f in range(24)])
    g = po.ghost_fill(_scrambled(u, bl), bl)
    mk = ghost_mask(bl)
    mk[bl.offsets[1]:] = False  # block 0's ghost nodes inside block 1
    return np.abs(g[:, mk] - u[:, mk]).max()


def test_oracle_converges_at_sixth_order():
    e1, e2 = _face_error(16, 24), _face_error(32, 48)
    assert e1 < 1e-6
    assert 40.0 < e1 / e2 < 110.0, (e1, e2)
