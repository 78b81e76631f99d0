"""GPU tests of the same-level ghost refresh, the fused RK4 stages and the
captured RK4-step graph, against the oracle RK4 (oracle RHS, textbook
tableau, independent ghost fill)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

NB, N = 2, 16
DT = 0.25 * (2.0 / (NB * N))


def _setup(n=N):
    import torch
    from sympygr_amd import data
    bl = data.uniform_tiling(nb=NB, n=n, lo=-1.0, hi=1.0, stagger=0.5, faces=True)
    u = data.fill_blocks(bl, kind="perturbed", seed=3, amp=1e-2, lo=-1.0, L=2.0)
    return torch, bl, u


def _interior_mask(bl):
    mk = np.zeros(bl.total, dtype=bool)
    for b in range(len(bl)):
        s = bl.sz[b]
        v = np.zeros((s[2], s[1], s[0]), dtype=bool)
        v[3:-3, 3:-3, 3:-3] = True
        mk[bl.offsets[b]:bl.offsets[b] + v.size] = v.reshape(-1)
    return mk


def _err(a, b, mk):
    a, b = a[:, mk], b[:, mk]
    return max(np.abs(a[f] - b[f]).max() / max(np.abs(b[f]).max(), 1e-300) for f in range(24))


def test_ghost_fill_matches_oracle():
    import pyoracle as po
    import sympygr_amd as sgr
    torch, bl, u = _setup()
    rng = np.random.default_rng(0)
    noisy = u + 0.0
    mk = _interior_mask(bl)
    noisy[:, ~mk] = rng.standard_normal((24, (~mk).sum()))  # scramble all ghosts
    plan = sgr.Plan(bl, sgr.params(adv="upwind"))
    du = torch.from_numpy(noisy).cuda()
    plan.ghost_fill(du)
    torch.cuda.synchronize()
    ref = po.ghost_fill_uniform(noisy, bl, NB)
    np.testing.assert_array_equal(du.cpu().numpy(), ref)


@pytest.mark.parametrize("adv,n", [("upwind", 16), ("centred", 16), ("upwind", 17), ("upwind", 20)])
def test_rk4_step_matches_oracle(adv, n):
    """Two RK4 steps against the oracle stepper.  Interiors of 17 and 20
    points are not multiples of the 8-point tile: the last tile along each
    axis is shifted back and overlaps its neighbour, and the accumulator of
    stages 1 and 2 (acc += dt/3 k) must still be updated once per point."""
    import pyoracle as po
    import sympygr_amd as sgr
    torch, bl, u = _setup(n)
    print(f"rk4 step adv={adv} interior={n}", flush=True)
    plan = sgr.Plan(bl, sgr.params(adv=adv, ko_sigma=0.1))
    du = torch.from_numpy(u).cuda()
    sa, sb, acc = torch.zeros_like(du), torch.zeros_like(du), torch.zeros_like(du)
    P = po.make_params(upwind=adv == "upwind", ko_sigma=0.1)
    ref = u
    mk = _interior_mask(bl)
    for step in range(2):
        plan.rk4_step(DT, du, sa, sb, acc)
        torch.cuda.synchronize()
        ref = po.rk4_step(ref, bl, NB, DT, P, nthreads=4)
        e = _err(du.cpu().numpy(), ref, mk)
        print(f"{adv} n={n} step {step} err {e:.2e}", flush=True)
        assert e <= 1e-12


def test_rk4_graph_replays_match_eager():
    """The captured step, replayed three times with eager work and syncs in
    between, equals three eager steps bit for bit."""
    import sympygr_amd as sgr
    torch, bl, u = _setup()
    plan = sgr.Plan(bl, sgr.params(adv="upwind", ko_sigma=0.1))
    de = torch.from_numpy(u).cuda()
    dg = de.clone()
    bufs_e = [torch.zeros_like(de) for _ in range(3)]
    bufs_g = [torch.zeros_like(de) for _ in range(3)]
    g = plan.rk4_graph(DT, dg, *bufs_g)
    for rep in range(3):
        plan.rk4_step(DT, de, *bufs_e)
        torch.cuda.synchronize()
        _ = (de * 2.0).sum().item()  # unrelated eager work between replays
        g.launch()
        torch.cuda.synchronize()
        assert torch.equal(dg, de), f"replay {rep} differs"
    g.close()


@pytest.mark.parametrize("n_int", [13, 25])
def test_rk4_stages_on_a_shifted_tile_block(n_int):
    """sgr_rk4_stage 0..3 one by one on a single block of 13 or 25 interior
    points (Dendro's eleOrder*2^k+1 sizes, every axis with an overlapping last
    tile), each stage against the tableau applied to the oracle RHS."""
    import pyoracle as po
    import sympygr_amd as sgr
    from sympygr_amd import data
    torch = __import__("torch")
    n = n_int + 6
    pmin, pmax = (-0.6, -0.5, -0.7), (-0.6 + 0.05 * (n - 1), -0.5 + 0.05 * (n - 1), -0.7 + 0.05 * (n - 1))
    bl = data.BlockList()
    bl.add((n, n, n), pmin, pmax, 0)
    u0 = data.perturbed_flat_block(seed=9, n=n, amp=1e-2, pmin=pmin, pmax=pmax)
    P = po.make_params(upwind=True, ko_sigma=0.1)
    plan = sgr.Plan(bl, sgr.params(adv="upwind", ko_sigma=0.1))
    dt = 0.01
    mk = _interior_mask(bl)
    du0 = torch.from_numpy(u0).cuda()
    acc = torch.full_like(du0, 3.0)
    nxt = torch.zeros_like(du0)
    stage_in = du0.clone()
    ref_acc = np.full_like(u0, 3.0)
    ref_stage = u0.copy()
    for stage in range(4):
        print(f"n_int={n_int} stage {stage}", flush=True)
        k = po.rhs_block(ref_stage, (n, n, n), pmin, pmax, 0, P)
        w = dt / (6.0 if stage in (0, 3) else 3.0)
        cn = (dt / 2.0, dt / 2.0, dt)[stage] if stage < 3 else None
        if stage < 3:
            ref_acc[:, mk] = (u0[:, mk] if stage == 0 else ref_acc[:, mk]) + w * k[:, mk]
            new_stage = u0.copy()
            new_stage[:, mk] = u0[:, mk] + cn * k[:, mk]
        else:
            new_stage = ref_stage.copy()
            new_stage[:, mk] = ref_acc[:, mk] + w * k[:, mk]
        plan.rk4_stage(stage, dt, du0, stage_in, acc, nxt)
        torch.cuda.synchronize()
        got_next = nxt.cpu().numpy()
        e = _err(got_next, new_stage, mk)
        ea = _err(acc.cpu().numpy(), ref_acc, mk) if stage < 3 else 0.0
        print(f"  next err {e:.2e} acc err {ea:.2e}", flush=True)
        assert e <= 1e-12 and ea <= 1e-12
        # next stage reads this stage's output (ghosts unchanged: one block, no refresh)
        ref_stage = new_stage.copy()
        ref_stage[:, ~mk] = u0[:, ~mk]
        stage_in = torch.from_numpy(ref_stage).cuda()


def test_rk4_stage_rejects_bad_arguments():
    import sympygr_amd as sgr
    torch, bl, u = _setup()
    plan = sgr.Plan(bl, sgr.params())
    du = torch.from_numpy(u).cuda()
    with pytest.raises(sgr.SgrError):
        plan.rk4_stage(4, DT, du, du, du, du)
    with pytest.raises(sgr.SgrError):
        plan.rk4_stage(0, float("nan"), du, du, du, du)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_constraints_match_oracle(precision):
    """GPU constraint kernel vs the oracle on perturbed multi-block data (the
    mixed plan stages derivatives in FP32, hence its looser bound)."""
    import pyoracle as po
    import sympygr_amd as sgr
    torch, bl, u = _setup()
    plan = sgr.Plan(bl, sgr.params(adv="upwind", precision=precision))
    du = torch.from_numpy(u).cuda()
    c = torch.zeros((4, bl.total), dtype=torch.float64, device="cuda")
    plan.constraints(du, c)
    torch.cuda.synchronize()
    got = c.cpu().numpy()
    for b in range(len(bl)):
        o, s = bl.offsets[b], bl.sz[b]
        n = s[0] * s[1] * s[2]
        ref = po.constraints_block(u[:, o:o + n], s, bl.pmin[b], bl.pmax[b], bl.bflag[b])
        g = got[:, o:o + n].reshape(4, s[2], s[1], s[0])[:, 3:-3, 3:-3, 3:-3].reshape(4, -1)
        r = ref.reshape(4, s[2], s[1], s[0])[:, 3:-3, 3:-3, 3:-3].reshape(4, -1)
        e = max(np.abs(g[q] - r[q]).max() / np.abs(r[q]).max() for q in range(4))
        assert e <= (1e-11 if precision == "fp64" else 1e-5), (b, e)


def test_chunked_workspace_matches_single_launch(monkeypatch):
    """A plan whose workspace budget holds 2 tiles (32 launches per pass)
    gives bit-identical RHS and constraints to the single-launch plan."""
    import sympygr_amd as sgr
    torch, bl, u = _setup()
    du = torch.from_numpy(u).cuda()
    monkeypatch.delenv("SGR_WS_GIB", raising=False)
    one = sgr.Plan(bl, sgr.params(adv="upwind"))
    assert one.info().chunks == 1
    monkeypatch.setenv("SGR_WS_GIB", "0.002")
    many = sgr.Plan(bl, sgr.params(adv="upwind"))
    assert many.info().chunks > 8
    a, b = torch.zeros_like(du), torch.zeros_like(du)
    one.rhs(a, du)
    many.rhs(b, du)
    ca = torch.zeros((4, bl.total), dtype=torch.float64, device="cuda")
    cb = torch.zeros_like(ca)
    one.constraints(du, ca)
    many.constraints(du, cb)
    torch.cuda.synchronize()
    assert torch.equal(a, b)
    assert torch.equal(ca, cb)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_octree_uneven_blocks_match_oracle(precision):
    """Uneven octree leaves (interior 16/24/32/48, refined toward the
    punctures, Brill-Lindquist data) in one plan.  FP64: the oracle on one
    block of each size plus the first boundary block of each size, 1e-12
    normwise per field and block.  Mixed precision (FP32 pointwise algebra),
    the stated criterion: the per-field normwise error over the whole block
    set against the FP64 path (itself checked against the oracle above) at
    most 1e-6; and each sampled block against the oracle within 1e-5 (a far
    block's own maximum is small, so its own ratio is larger)."""
    import pyoracle as po
    import sympygr_amd as sgr
    from sympygr_amd import data
    torch = __import__("torch")
    bl = data.octree_blocks(lmax=4)
    u = data.fill_blocks(bl, backend="torch", device="cuda")
    out = torch.zeros_like(u)
    plan = sgr.Plan(bl, sgr.params(adv="upwind", precision=precision))
    plan.rhs(out, u)
    torch.cuda.synchronize()
    if precision == "mixed":
        ref64 = torch.zeros_like(u)
        sgr.Plan(bl, sgr.params(adv="upwind", precision="fp64")).rhs(ref64, u)
        torch.cuda.synchronize()
        mk = torch.from_numpy(_interior_mask(bl)).cuda()
        d = (out - ref64)[:, mk].abs().amax(dim=1)
        r = ref64[:, mk].abs().amax(dim=1).clamp_min(1e-300)
        whole = (d / r).max().item()
        print(f"octree mixed vs fp64, whole set: max per-field normwise {whole:.2e}", flush=True)
        assert whole <= 1e-6
    pick, seen = [], set()
    for b in range(len(bl)):
        key = (bl.sz[b][0], bl.bflag[b] != 0)
        if key not in seen:
            seen.add(key)
            pick.append(b)
    P = po.make_params(upwind=True)
    for b in pick:
        o, s = bl.offsets[b], bl.sz[b]
        n = s[0] * s[1] * s[2]
        ref = po.rhs_block(u[:, o:o + n].cpu().numpy(), s, bl.pmin[b], bl.pmax[b], bl.bflag[b], P)
        got = out[:, o:o + n].cpu().numpy()
        g = got.reshape(24, s[2], s[1], s[0])[:, 3:-3, 3:-3, 3:-3].reshape(24, -1)
        r = ref.reshape(24, s[2], s[1], s[0])[:, 3:-3, 3:-3, 3:-3].reshape(24, -1)
        e = max(np.abs(g[f] - r[f]).max() / max(np.abs(r[f]).max(), 1e-300) for f in range(24))
        print(f"octree {precision} block {b} size {s[0]} bflag {bl.bflag[b]} err {e:.2e}", flush=True)
        assert e <= (1e-12 if precision == "fp64" else 1e-5)


def test_large_single_block_matches_sub_blocks():
    """One 192^3-interior block (198^3 padded, 13824 tiles in one launch)
    against 38^3 sub-blocks cut from the same data at three places (the big
    block's face closures included through bflag=63 corners)."""
    import sympygr_amd as sgr
    from sympygr_amd import data
    torch = __import__("torch")
    n = 198
    lo, h = -4.0, 8.0 / 197
    big = data.BlockList()
    big.add((n, n, n), (lo,) * 3, (lo + (n - 1) * h,) * 3, 63)
    u = data.fill_blocks(big, backend="torch", device="cuda", kind="perturbed", seed=7, amp=1e-2, lo=lo, L=8.0)
    out = torch.zeros_like(u)
    plan = sgr.Plan(big, sgr.params(adv="upwind"))
    plan.rhs(out, u)
    torch.cuda.synchronize()
    U = u.view(24, n, n, n)
    O = out.view(24, n, n, n)
    for (i0, j0, k0) in ((0, 0, 0), (80, 64, 96), (160, 160, 160)):
        m = 38
        sub = data.BlockList()
        flags = (1 if i0 == 0 else 0) | (4 if j0 == 0 else 0) | (16 if k0 == 0 else 0) | \
                (2 if i0 + m == n else 0) | (8 if j0 + m == n else 0) | (32 if k0 + m == n else 0)
        pmin = (lo + i0 * h, lo + j0 * h, lo + k0 * h)
        sub.add((m, m, m), pmin, tuple(p + (m - 1) * h for p in pmin), flags)
        su = U[:, k0:k0 + m, j0:j0 + m, i0:i0 + m].contiguous().view(24, -1)
        so = torch.zeros_like(su)
        sgr.Plan(sub, sgr.params(adv="upwind")).rhs(so, su)
        torch.cuda.synchronize()
        a = so.view(24, m, m, m)[:, 3:-3, 3:-3, 3:-3]
        b = O[:, k0 + 3:k0 + m - 3, j0 + 3:j0 + m - 3, i0 + 3:i0 + m - 3]
        err = ((a - b).abs().amax(dim=(1, 2, 3)) / b.abs().amax(dim=(1, 2, 3)).clamp_min(1e-300)).max().item()
        print(f"sub-block at {(i0, j0, k0)} flags {flags} err {err:.2e}", flush=True)
        assert err <= 1e-12


def test_check_finite_reports_an_injected_nan():
    """A NaN in one ghost node of one block is reported as SGR_E_NONFINITE;
    the clean input passes; an Inf is reported as well."""
    import sympygr_amd as sgr
    torch, bl, u = _setup()
    plan = sgr.Plan(bl, sgr.params())
    du = torch.from_numpy(u).cuda()
    assert plan.check_finite(du) == 0
    ghost = int(np.flatnonzero(~_interior_mask(bl))[123])
    du[7, ghost] = float("nan")
    assert plan.check_finite(du) == -5
    assert "non-finite" in sgr.last_error()
    du[7, ghost] = u[7, ghost]
    du[20, bl.offsets[-1] + 5] = float("inf")
    assert plan.check_finite(du) == -5


def test_short_or_foreign_tensors_are_rejected_on_the_host():
    import sympygr_amd as sgr
    torch, bl, u = _setup()
    plan = sgr.Plan(bl, sgr.params())
    du = torch.from_numpy(u).cuda()
    with pytest.raises(ValueError, match="needs"):
        plan.ghost_fill(du[:, :-1].contiguous())
    with pytest.raises(ValueError, match="float64"):
        plan.rhs(du.float(), du)
    with pytest.raises(ValueError, match="cuda:0"):
        plan.check_finite(torch.from_numpy(u))
    c = torch.zeros((4, bl.total - 1), dtype=torch.float64, device="cuda")
    with pytest.raises(ValueError, match="needs"):
        plan.constraints(du, c)
