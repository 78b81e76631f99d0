"""The pipelined chunk schedule: a plan whose workspace needs more than one
chunk runs its chunks over two workspace buffers, each chunk's stencil pass
on the caller's stream overlapping the previous chunk's pointwise pass on a
second stream.  Every point is computed by the same kernels in the same
order, so results equal consecutive chunks and a single chunk bit for bit:
plain RHS, RK4 stages, BSSN and CCZ4, both precisions, TVB dissipation, and
a captured RK4 step replayed between eager work.  A pipelined plan spreads
blocks with face tiles over the tile order (SGR_TILE_ORDER forces either
order); results do not depend on the order either."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _plan(bl, prm, monkeypatch, pipe=None, gib=None, order=None):
    import sympygr_amd as sgr
    env = (("SGR_PIPE_CHUNKS", pipe), ("SGR_WS_GIB", gib), ("SGR_TILE_ORDER", order))
    for k, v in env:
        if v is None:
            monkeypatch.delenv(k, raising=False)
        else:
            monkeypatch.setenv(k, v)
    plan = sgr.Plan(bl, prm)
    for k, _ in env:
        monkeypatch.delenv(k, raising=False)
    return plan


@pytest.mark.parametrize("kind", ["fp64", "mixed", "tvb5", "ccz4"])
def test_pipelined_equals_consecutive_and_single_chunk(kind, monkeypatch):
    import torch
    import sympygr_amd as sgr
    from sympygr_amd import data
    bl = data.uniform_tiling(nb=3, n=16, lo=-1.0, hi=1.0, stagger=0.5, faces=True)
    kw = dict(adv="upwind", ko_sigma=0.1)
    if kind == "mixed":
        kw["precision"] = "mixed"
    if kind == "tvb5":
        kw["dissipation"] = "tvb5"
    if kind == "ccz4":
        kw["formulation"] = "ccz4"
        u = data.fill_blocks(bl, backend="torch", device="cuda", kind="ccz4", seed=4, amp=1e-2, lo=-1.0, L=2.0)
    else:
        u = data.fill_blocks(bl, backend="torch", device="cuda", kind="perturbed", seed=6, amp=1e-2, lo=-1.0, L=2.0)
    prm = sgr.params(**kw)
    res = {}
    for name, pipe, gib, order in (("single", None, None, None), ("consecutive", "0", "0.01", None),
                                   ("pipelined", "5", None, None), ("budget", None, "0.01", None),
                                   ("single spread", None, None, "spread"), ("pipelined block", "5", None, "block")):
        plan = _plan(bl, prm, monkeypatch, pipe, gib, order)
        info = plan.info()
        print(f"{kind} {name}: chunks={info.chunks} pipelined={info.pipelined}", flush=True)
        out = torch.full_like(u, 5.0)
        plan.rhs(out, u)
        acc = torch.full_like(u, 3.0)
        nxt = torch.zeros_like(u)
        plan.rk4_stage(1, 0.01, u, u, acc, nxt)
        torch.cuda.synchronize()
        res[name] = (out, acc, nxt, info.chunks, info.pipelined)
        del plan
    assert res["single"][3] == 1 and not res["single"][4]
    assert res["consecutive"][3] > 1 and not res["consecutive"][4]
    assert res["pipelined"][3] >= 5 and res["pipelined"][4]
    # a workspace budget that needs several chunks selects the pipelined
    # schedule, with twice as many chunks
    assert res["budget"][4] and res["budget"][3] >= 2 * res["consecutive"][3] - 1
    for name in ("consecutive", "pipelined", "budget", "single spread", "pipelined block"):
        for k in range(3):
            assert torch.equal(res[name][k], res["single"][k]), (kind, name, k)


def test_pipelined_graph_replays_match_eager(monkeypatch):
    """An RK4 step of a pipelined plan captured into a graph (the second
    stream joins the capture through the fork and join events), replayed
    three times with eager work and synchronizations in between, equals eager
    steps of a single-chunk plan bit for bit."""
    import torch
    import sympygr_amd as sgr
    from sympygr_amd import data
    bl = data.uniform_tiling(nb=2, n=16, lo=-1.0, hi=1.0, stagger=0.5, faces=True)
    u = data.fill_blocks(bl, kind="perturbed", seed=3, amp=1e-2, lo=-1.0, L=2.0)
    prm = sgr.params(adv="upwind", ko_sigma=0.1)
    eager = _plan(bl, prm, monkeypatch)
    piped = _plan(bl, prm, monkeypatch, pipe="4")
    assert piped.info().pipelined and piped.info().chunks == 4
    dt = 0.25 * (2.0 / 32)
    de = torch.from_numpy(u).cuda()
    dg = de.clone()
    bufs_e = [torch.zeros_like(de) for _ in range(3)]
    bufs_g = [torch.zeros_like(de) for _ in range(3)]
    g = piped.rk4_graph(dt, dg, *bufs_g)
    for rep in range(3):
        eager.rk4_step(dt, de, *bufs_e)
        torch.cuda.synchronize()
        _ = (de * 2.0).sum().item()  # unrelated eager work between replays
        g.launch()
        torch.cuda.synchronize()
        print(f"replay {rep}", flush=True)
        assert torch.equal(dg, de), f"replay {rep} differs"
    g.close()
    # eager steps of the pipelined plan as well
    dp = torch.from_numpy(u).cuda()
    bufs_p = [torch.zeros_like(dp) for _ in range(3)]
    for _ in range(3):
        piped.rk4_step(dt, dp, *bufs_p)
    torch.cuda.synchronize()
    assert torch.equal(dp, de)


def test_large_plan_without_tile_queues_runs_pipelined(monkeypatch):
    """A plan of at least 8192 tiles whose stencil pass runs one workgroup
    per tile (no flagged faces, so no tile queues) runs pipelined in one
    workspace budget, and equals the consecutive schedule bit for bit."""
    import torch
    import sympygr_amd as sgr
    from sympygr_amd import data
    bl = data.uniform_tiling(nb=8, n=24, lo=-1.0, hi=1.0, stagger=0.5, faces=False)
    u = data.fill_blocks(bl, backend="torch", device="cuda", kind="perturbed", seed=2, amp=1e-2, lo=-1.0, L=2.0)
    prm = sgr.params(adv="upwind", ko_sigma=0.1)
    outs = []
    for pipe in (None, "0"):
        plan = _plan(bl, prm, monkeypatch, pipe)
        info = plan.info()
        print(f"SGR_PIPE_CHUNKS={pipe}: tiles={info.tiles} chunks={info.chunks} pipelined={info.pipelined} "
              f"queue={info.tile_queue}", flush=True)
        assert info.tile_queue == 0 and info.tiles >= 8192
        assert bool(info.pipelined) == (pipe is None)
        out = torch.full_like(u, 5.0)
        plan.rhs(out, u)
        torch.cuda.synchronize()
        outs.append(out)
        del plan
    assert torch.equal(outs[0], outs[1])
