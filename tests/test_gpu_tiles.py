"""RHS tile shapes other than 8^3 (the general stencil kernel, chosen per
block size by the tile table or forced with SGR_TILE_SHAPE).  The stencils
are the same expressions in the same order whatever the tile, so every
result must equal the 8^3 path's bit for bit; the 8^3 path itself is checked
against the oracle and the reference fixtures elsewhere.  Blocks of
Dendro's sizes (13, 17, 25 interior points) with shapes that fit them, and
shapes whose last tile is shifted back (13 with 6-point tiles)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

CASES = [  # interior size, tile shape
    (13, (13, 13, 3)),
    (13, (6, 6, 6)),
    (17, (17, 6, 5)),
    (25, (25, 4, 5)),
    (9, (9, 9, 6)),
    (12, (4, 12, 8)),
]


def _block(n_int, seed, bflag):
    from sympygr_amd import data
    n = n_int + 6
    pmin = (-0.6, -0.5, -0.7)
    pmax = tuple(p + 0.05 * (n - 1) for p in pmin)
    bl = data.BlockList()
    bl.add((n, n, n), pmin, pmax, bflag)
    u = data.perturbed_flat_block(seed=seed, n=n, amp=1e-2, pmin=pmin, pmax=pmax)
    return bl, u


def _rhs(bl, u, prm, monkeypatch, shape):
    import torch
    import sympygr_amd as sgr
    if shape is None:
        monkeypatch.delenv("SGR_TILE_SHAPE", raising=False)
    else:
        monkeypatch.setenv("SGR_TILE_SHAPE", ",".join(str(x) for x in shape))
    plan = sgr.Plan(bl, prm)
    du = torch.from_numpy(np.ascontiguousarray(u)).cuda()
    out = torch.full_like(du, 5.0)
    plan.rhs(out, du)
    torch.cuda.synchronize()
    tiles = plan.info().tiles
    del plan
    monkeypatch.delenv("SGR_TILE_SHAPE", raising=False)
    return out.cpu().numpy(), tiles


@pytest.mark.parametrize("n_int,shape", CASES)
def test_tile_shape_equals_cube_tiles(n_int, shape, monkeypatch):
    import sympygr_amd as sgr
    for bflag, prec, adv in ((0, "fp64", "upwind"), (63, "fp64", "upwind"), (42, "mixed", "upwind"),
                             (21, "fp64", "centred")):
        print(f"n_int={n_int} shape={shape} bflag={bflag} {prec} {adv}", flush=True)
        bl, u = _block(n_int, 11 + n_int, bflag)
        prm = sgr.params(adv=adv, precision=prec, ko_sigma=0.1)
        ref, _ = _rhs(bl, u, prm, monkeypatch, (8, 8, 8))
        got, tg = _rhs(bl, u, prm, monkeypatch, shape)
        nt = [-(-n_int // s) for s in shape]
        assert tg == nt[0] * nt[1] * nt[2], (tg, nt)
        assert np.array_equal(got, ref), (n_int, shape, bflag, prec, np.abs(got - ref).max())


def test_tile_shape_ccz4_and_rochester(monkeypatch):
    import sympygr_amd as sgr
    from sympygr_amd import data
    n_int, shape = 13, (13, 13, 3)
    n = n_int + 6
    bl = data.BlockList()
    bl.add((n, n, n), (-0.95,) * 3, (0.85,) * 3, 5)
    u = data.fill_blocks(bl, kind="ccz4", seed=4, amp=1e-2, lo=-0.95, L=1.8)
    prm = sgr.params(adv="upwind", formulation="ccz4", kappa=(0.02, 0.0, 0.5))
    ref, _ = _rhs(bl, u, prm, monkeypatch, (8, 8, 8))
    got, _ = _rhs(bl, u, prm, monkeypatch, shape)
    assert np.array_equal(got, ref)
    bl, u = _block(17, 3, 63)
    prm = sgr.params(adv="upwind", gauge="rochester", eta="func", xi=(1, 1, 1))
    ref, _ = _rhs(bl, u, prm, monkeypatch, (8, 8, 8))
    got, _ = _rhs(bl, u, prm, monkeypatch, (17, 6, 5))
    assert np.array_equal(got, ref)


def test_tile_shape_rk4_stages_update_each_point_once(monkeypatch):
    """A shape whose last tile overlaps its neighbour on every axis (13 with
    6-point tiles): the stage-1 and stage-2 read-modify-write of the
    accumulator must apply once per point, as on 8^3 tiles."""
    import torch
    import sympygr_amd as sgr
    bl, u0 = _block(13, 21, 0)
    prm = sgr.params(adv="upwind", ko_sigma=0.1)
    res = []
    for shape in ("8,8,8", "6,6,6"):
        monkeypatch.setenv("SGR_TILE_SHAPE", shape)
        plan = sgr.Plan(bl, prm)
        du0 = torch.from_numpy(u0).cuda()
        acc = torch.full_like(du0, 3.0)
        nxt = torch.zeros_like(du0)
        stage_in = du0.clone()
        for stage in range(4):
            plan.rk4_stage(stage, 0.01, du0, stage_in, acc, nxt)
            if stage < 3:
                stage_in = nxt.clone()
        torch.cuda.synchronize()
        res.append((acc.cpu().numpy(), nxt.cpu().numpy()))
        del plan
    monkeypatch.delenv("SGR_TILE_SHAPE", raising=False)
    assert np.array_equal(res[0][0], res[1][0]) and np.array_equal(res[0][1], res[1][1])


def test_mixed_shapes_in_one_plan(monkeypatch):
    """One plan over blocks of several sizes with the tile table's shapes
    (where it gives some sizes other shapes than 8^3: both stencil kernels in
    one pass, tiles of both kinds in one pointwise launch), and with every
    block forced to a general shape, equal the plan with 8^3 tiles
    everywhere."""
    import torch
    import sympygr_amd as sgr
    from sympygr_amd import data
    bl = data.BlockList()
    for k, m in enumerate((13, 16, 17, 25, 32)):
        n = m + 6
        lo = -1.0 + 0.3 * k
        bl.add((n, n, n), (lo, -0.5, -0.7), (lo + 0.04 * (n - 1), -0.5 + 0.04 * (n - 1), -0.7 + 0.04 * (n - 1)),
               (1 << k) & 63)
    u = np.zeros((24, bl.total))
    for b in range(len(bl)):
        n = bl.sz[b][0]
        u[:, bl.offsets[b]:bl.offsets[b] + n ** 3] = data.perturbed_flat_block(
            seed=30 + b, n=n, amp=1e-2, pmin=bl.pmin[b], pmax=bl.pmax[b])
    prm = sgr.params(adv="upwind", ko_sigma=0.1)
    du = torch.from_numpy(u).cuda()
    outs = {}
    for name, shape in (("cube", "8,8,8"), ("table", None), ("forced", "13,13,3")):
        if shape:
            monkeypatch.setenv("SGR_TILE_SHAPE", shape)
        else:
            monkeypatch.delenv("SGR_TILE_SHAPE", raising=False)
        print(f"plan with {name} tile shapes", flush=True)
        outs[name] = torch.full_like(du, 5.0)
        sgr.Plan(bl, prm).rhs(outs[name], du)
        torch.cuda.synchronize()
    monkeypatch.delenv("SGR_TILE_SHAPE", raising=False)
    assert torch.equal(outs["table"], outs["cube"])
    assert torch.equal(outs["forced"], outs["cube"])


def test_bad_tile_shapes_are_refused(monkeypatch):
    import sympygr_amd as sgr
    bl, _ = _block(13, 1, 0)
    for bad in ("1,8,8", "16,16,4", "30,8,2", "8,8", "a,b,c"):
        monkeypatch.setenv("SGR_TILE_SHAPE", bad)
        with pytest.raises(sgr.SgrError, match="SGR_TILE_SHAPE"):
            sgr.Plan(bl, sgr.params())
    monkeypatch.delenv("SGR_TILE_SHAPE", raising=False)


@pytest.mark.parametrize("shape", [None, "13,13,3"])
def test_stencil_queue_equals_one_workgroup_per_tile(shape, monkeypatch):
    """The stencil pass as a persistent grid fed by the per-XCD tile queues
    and as one workgroup per tile give the same RHS bit for bit, also on RK4
    stages and over several launches (workspace chunks)."""
    import torch
    import sympygr_amd as sgr
    from sympygr_amd import data
    bl = data.uniform_tiling(nb=3, n=13 if shape else 16, lo=-1.0, hi=1.0, stagger=0.5, faces=True)
    u = data.fill_blocks(bl, backend="torch", device="cuda", kind="perturbed", seed=8, amp=1e-2, lo=-1.0, L=2.0)
    if shape:
        monkeypatch.setenv("SGR_TILE_SHAPE", shape)
    outs = []
    for q, gib in (("0", None), ("1", None), ("1", "0.02")):
        monkeypatch.setenv("SGR_STENCIL_QUEUE", q)
        if gib:
            monkeypatch.setenv("SGR_WS_GIB", gib)
        print(f"queue={q} ws_gib={gib} shape={shape}", flush=True)
        plan = sgr.Plan(bl, sgr.params(adv="upwind", ko_sigma=0.1))
        out = torch.full_like(u, 5.0)
        plan.rhs(out, u)
        acc = torch.full_like(u, 3.0)
        nxt = torch.zeros_like(u)
        plan.rk4_stage(1, 0.01, u, u, acc, nxt)
        torch.cuda.synchronize()
        outs.append((out, acc, nxt, plan.info().chunks))
        del plan
        monkeypatch.delenv("SGR_WS_GIB", raising=False)
    monkeypatch.delenv("SGR_STENCIL_QUEUE", raising=False)
    monkeypatch.delenv("SGR_TILE_SHAPE", raising=False)
    assert outs[2][3] > 1
    for o in outs[1:]:
        for k in range(3):
            assert torch.equal(o[k], outs[0][k]), k
