"""GPU tests of sgr_constraints_psi4 (GR/constraints.py:46-173) against the
oracle's oracle_constraints_psi4 on perturbed 8-block data with boundary
faces and a different origin per block (so the coordinates pmin + i h
enter), in FP64 and mixed precision, chunked and not."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _setup():
    import torch
    from sympygr_amd import data
    bl = data.uniform_tiling(nb=2, n=16, lo=-1.0, hi=1.0, stagger=0.5, faces=True)
    u = data.fill_blocks(bl, kind="perturbed", seed=5, amp=1e-2, lo=-1.0, L=2.0)
    return torch, bl, u


def _run(plan, du, total, psi4=True):
    import torch
    c = torch.full((6 if psi4 else 4, total), 7.0, dtype=torch.float64, device="cuda")
    (plan.constraints_psi4 if psi4 else plan.constraints)(du, c)
    torch.cuda.synchronize()
    return c.cpu().numpy()


def _interior(a, s):
    return a.reshape(a.shape[0], s[2], s[1], s[0])[:, 3:-3, 3:-3, 3:-3].reshape(a.shape[0], -1)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_psi4_matches_oracle(precision):
    """Every interior node, each row against its own maximum: FP64 within
    1e-11, mixed (FP32-staged derivatives) within 1e-5; nodes outside the
    interior are not written."""
    import pyoracle as po
    import sympygr_amd as sgr
    torch, bl, u = _setup()
    plan = sgr.Plan(bl, sgr.params(adv="upwind", precision=precision))
    got = _run(plan, torch.from_numpy(u).cuda(), bl.total)
    tol = 1e-11 if precision == "fp64" else 1e-5
    for b in range(len(bl)):
        o, s = bl.offsets[b], bl.sz[b]
        n = s[0] * s[1] * s[2]
        ref = _interior(po.constraints_block(u[:, o:o + n], s, bl.pmin[b], bl.pmax[b], bl.bflag[b], psi4=True), s)
        g = _interior(got[:, o:o + n], s)
        for q in range(6):
            scale = np.abs(ref[q]).max()
            assert scale > 0, (b, q)
            e = np.abs(g[q] - ref[q]).max() / scale
            assert e <= tol, (b, q, e)
        full = got[:, o:o + n].reshape(6, s[2], s[1], s[0])
        full[:, 3:-3, 3:-3, 3:-3] = 7.0
        assert np.all(full == 7.0), b


def test_psi4_rows_0_to_3_are_the_constraints():
    """Rows 0..3 agree with sgr_constraints (the two kernels share the
    derivative pass; their algebra is generated separately, so agreement is
    to rounding)."""
    import sympygr_amd as sgr
    torch, bl, u = _setup()
    plan = sgr.Plan(bl, sgr.params(adv="upwind"))
    du = torch.from_numpy(u).cuda()
    a = _run(plan, du, bl.total)
    c = _run(plan, du, bl.total, psi4=False)
    for b in range(len(bl)):
        o, s = bl.offsets[b], bl.sz[b]
        n = s[0] * s[1] * s[2]
        x, y = _interior(a[:4, o:o + n], s), _interior(c[:, o:o + n], s)
        for q in range(4):
            assert np.abs(x[q] - y[q]).max() <= 1e-12 * np.abs(y[q]).max(), (b, q)


def test_psi4_chunked_workspace_is_bit_identical(monkeypatch):
    import sympygr_amd as sgr
    torch, bl, u = _setup()
    du = torch.from_numpy(u).cuda()
    monkeypatch.delenv("SGR_WS_GIB", raising=False)
    one = sgr.Plan(bl, sgr.params(adv="upwind"))
    monkeypatch.setenv("SGR_WS_GIB", "0.002")
    many = sgr.Plan(bl, sgr.params(adv="upwind"))
    assert many.info().chunks > 8
    assert np.array_equal(_run(one, du, bl.total), _run(many, du, bl.total))


def test_psi4_refuses_null_rows():
    import ctypes
    import sympygr_amd as sgr
    torch, bl, u = _setup()
    plan = sgr.Plan(bl, sgr.params())
    du = torch.from_numpy(u).cuda()
    c = torch.zeros((6, bl.total), dtype=torch.float64, device="cuda")
    rows = (ctypes.POINTER(ctypes.c_double) * 6)(*[ctypes.cast(c[q].data_ptr(), ctypes.POINTER(ctypes.c_double))
                                                   for q in range(6)])
    rows[5] = ctypes.POINTER(ctypes.c_double)()
    rc = sgr.lib().sgr_constraints_psi4(plan._h, plan._f(du), rows, None)
    assert rc != 0
    assert "null" in sgr.last_error()


def test_psi4_on_octree_leaves_of_other_sizes_and_spacings():
    """22 octree leaves of interior 16 and 24 on three levels (six different
    spacings), perturbed data: each leaf's coordinates pmin + i h enter with
    its own h.  FP64 GPU vs oracle within 1e-11 per row."""
    import torch
    import pyoracle as po
    import sympygr_amd as sgr
    from helpers import small_octree
    from sympygr_amd import data
    bl = small_octree()
    assert len({tuple(s) for s in bl.sz}) > 1
    u = data.fill_blocks(bl, kind="perturbed", seed=9, amp=1e-2, lo=-2.0, L=4.0)
    plan = sgr.Plan(bl, sgr.params(adv="upwind"))
    got = _run(plan, torch.from_numpy(u).cuda(), bl.total)
    for b in range(len(bl)):
        o, s = bl.offsets[b], bl.sz[b]
        n = s[0] * s[1] * s[2]
        ref = _interior(po.constraints_block(u[:, o:o + n], s, bl.pmin[b], bl.pmax[b], bl.bflag[b], psi4=True), s)
        g = _interior(got[:, o:o + n], s)
        for q in range(6):
            assert np.abs(g[q] - ref[q]).max() <= 1e-11 * np.abs(ref[q]).max(), (b, q)
