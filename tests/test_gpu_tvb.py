"""TVB dissipation (DISSIPATION_TYPE 1 and 2 of the reference's bssnrhs_sep,
bssn/src/rhs.cpp:729-757) on the GPU: against the compiled reference's
fixtures at every node where the reference's result is defined, and against
the oracle (which states the closures where it is not) on every face-flag
set, on blocks of several sizes and tile shapes, in both precisions, in RK4
stages, over several blocks and workspace chunks, and through the C++
drop-in bssnrhs_sep reading bssn::DISSIPATION_TYPE.

Tolerances: FP64 1e-12 normwise relative per field; mixed precision 1e-6
(FP32 algebra; the dissipation itself is FP64 in both)."""
import ctypes

import numpy as np
import pytest

from helpers import case_gauge, case_input, case_params, fixture_mask, golden, interior, manifest, normwise

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-12
MIXED_TOL = 1e-6
TVB_CASES = ["tvb3_staged_b63", "tvb3_staged_b42", "tvb5_staged_b63", "tvb5_staged_b42"]


def _run(bl, u, **pkw):
    import torch
    import sympygr_amd as sgr
    plan = sgr.Plan(bl, sgr.params(**pkw))
    du = torch.from_numpy(np.ascontiguousarray(u)).cuda()
    out = torch.full_like(du, 7.0)
    plan.rhs(out, du)
    torch.cuda.synchronize()
    return out.cpu().numpy()


def _one(sz, pmin, pmax, bflag):
    from sympygr_amd.data import BlockList
    bl = BlockList()
    bl.add(sz, pmin, pmax, bflag)
    return bl


@pytest.mark.parametrize("name", TVB_CASES)
def test_tvb_matches_reference_fixture(name):
    import pyoracle as po
    u, sz, pmin, pmax, c = case_input(name)
    p, _ = case_params(c)
    kw = case_gauge(c)
    print(f"case {name} {kw}", flush=True)
    m = fixture_mask(c)
    bl = _one(sz, pmin, pmax, c["bflag"])
    out = _run(bl, u, adv="upwind", ko_sigma=p["ko_sigma"], **kw)
    err = normwise(interior(out, sz)[:, m], golden(name)[:, m])
    assert err.max() <= FP64_TOL, (name, err)
    ref = po.rhs_block(u, sz, pmin, pmax, c["bflag"], po.make_params(upwind=True, ko_sigma=p["ko_sigma"], **kw))
    assert normwise(interior(out, sz), interior(ref, sz)).max() <= FP64_TOL
    outm = _run(bl, u, adv="upwind", precision="mixed", ko_sigma=p["ko_sigma"], **kw)
    em = normwise(interior(outm, sz), interior(ref, sz)).max()
    print(f"{name} mixed err {em:.2e}", flush=True)
    assert em <= MIXED_TOL


@pytest.mark.parametrize("diss", ["tvb3", "tvb5"])
def test_tvb_matches_oracle_any_bflag(diss):
    import pyoracle as po
    u, sz, pmin, pmax, _ = case_input("fake_upwind_b0")
    for bflag in (0, 1, 2, 12, 48, 63, 37):
        for adv in ("upwind", "centred"):
            print(f"{diss} bflag {bflag} adv {adv}", flush=True)
            out = _run(_one(sz, pmin, pmax, bflag), u, adv=adv, dissipation=diss)
            ref = po.rhs_block(u, sz, pmin, pmax, bflag, po.make_params(upwind=adv == "upwind", dissipation=diss))
            err = normwise(interior(out, sz), interior(ref, sz))
            assert err.max() <= FP64_TOL, (diss, bflag, adv, err)
            assert np.all(out.reshape(24, sz[2], sz[1], sz[0])[:, :3] == 7.0)  # ghost nodes untouched


@pytest.mark.parametrize("n_int", [7, 9, 13, 17])
def test_tvb_block_sizes_and_tile_shapes(n_int, monkeypatch):
    """Interior sizes 7 to 17 (tile table shapes at 9, 13, 17) against the
    oracle, and the 8^3 tiling bit-equal to the table's shape."""
    import pyoracle as po
    from sympygr_amd import data
    n = n_int + 6
    pmin = (-0.7, -0.4, -0.9)
    pmax = tuple(p + h * (n - 1) for p, h in zip(pmin, (0.07, 0.06, 0.08)))
    u = data.perturbed_flat_block(seed=7, n=n, amp=1e-2, pmin=pmin, pmax=pmax)
    for diss, bflag in (("tvb3", 0), ("tvb5", 63), ("tvb5", 21)):
        print(f"n_int={n_int} {diss} bflag={bflag}", flush=True)
        bl = _one((n, n, n), pmin, pmax, bflag)
        out = _run(bl, u, adv="upwind", dissipation=diss)
        ref = po.rhs_block(u, (n, n, n), pmin, pmax, bflag, po.make_params(upwind=True, dissipation=diss))
        assert normwise(interior(out, (n, n, n)), interior(ref, (n, n, n))).max() <= FP64_TOL
        monkeypatch.setenv("SGR_TILE_SHAPE", "8,8,8")
        cube = _run(bl, u, adv="upwind", dissipation=diss)
        monkeypatch.delenv("SGR_TILE_SHAPE")
        assert np.array_equal(out, cube)


def test_tvb_several_blocks_chunks_and_rk4_stage(monkeypatch):
    """A 27-block tiling with flagged outer faces against the oracle; the same
    plan split into workspace chunks gives the same result bit for bit; the
    fused RK4 stage 0 stores u0 + dt/6 k and u0 + dt/2 k of that RHS."""
    import torch
    import pyoracle as po
    import sympygr_amd as sgr
    from sympygr_amd import data
    bl = data.uniform_tiling(nb=3, n=16, lo=-1.0, hi=1.0, stagger=0.5, faces=True)
    u = data.fill_blocks(bl, kind="perturbed", seed=8, amp=1e-2, lo=-1.0, L=2.0)
    ref = po.rhs_blocks(u, bl, po.make_params(upwind=True, dissipation="tvb5"), nthreads=8)
    du = torch.from_numpy(u).cuda()
    outs = []
    for gib in (None, "0.005"):
        if gib:
            monkeypatch.setenv("SGR_WS_GIB", gib)
        plan = sgr.Plan(bl, sgr.params(adv="upwind", dissipation="tvb5"))
        print(f"ws_gib={gib} chunks={plan.info().chunks}", flush=True)
        out = torch.zeros_like(du)
        plan.rhs(out, du)
        acc = torch.zeros_like(du)
        nxt = torch.zeros_like(du)
        plan.rk4_stage(0, 0.01, du, du, acc, nxt)
        torch.cuda.synchronize()
        outs.append((out.cpu().numpy(), acc.cpu().numpy(), nxt.cpu().numpy(), plan.info().chunks))
        del plan
        monkeypatch.delenv("SGR_WS_GIB", raising=False)
    assert outs[1][3] > 1
    for k in range(3):
        assert np.array_equal(outs[0][k], outs[1][k]), k
    out, acc, nxt, _ = outs[0]
    for b in range(len(bl)):
        o, s = bl.offsets[b], bl.sz[b]
        sl = slice(o, o + s[0] * s[1] * s[2])
        g = interior(out[:, sl], s)
        assert normwise(g, interior(ref[:, sl], s)).max() <= FP64_TOL, b
        k = g.reshape(24, -1)
        u0 = interior(u[:, sl], s).reshape(24, -1)
        assert np.allclose(interior(acc[:, sl], s).reshape(24, -1), u0 + 0.01 / 6 * k, rtol=1e-15, atol=1e-15)
        assert np.allclose(interior(nxt[:, sl], s).reshape(24, -1), u0 + 0.01 / 2 * k, rtol=1e-15, atol=1e-15)


def test_cpp_dropin_bssnrhs_sep_reads_dissipation_type():
    """bssnrhs_sep with bssn::DISSIPATION_TYPE = 1 and 2 against the
    reference's fixtures; bssnrhs ignores it (Kreiss-Oliger, as the
    reference's bssnrhs); an unknown type fails the call (NaN interior, the
    error kept for sgr_dropin_status)."""
    import sympygr_amd as sgr
    L = ctypes.CDLL(sgr.LIB_PATH)
    PD = ctypes.POINTER(ctypes.c_double)
    argt = [ctypes.POINTER(PD), ctypes.POINTER(PD), ctypes.POINTER(ctypes.c_uint), PD, PD,
            ctypes.POINTER(ctypes.c_uint), ctypes.POINTER(ctypes.c_uint)]
    sep = getattr(L, "_Z11bssnrhs_sepPPdPPKdRKjS2_S2_PS4_S5_")
    plain = getattr(L, "_Z7bssnrhsPPdPPKdRKjS2_S2_PS4_S5_")
    for fn in (sep, plain):
        fn.argtypes = argt
        fn.restype = None
    dtype = ctypes.c_uint.in_dll(L, "_ZN4bssn16DISSIPATION_TYPEE")
    sigma = ctypes.c_double.in_dll(L, "_ZN4bssn13KO_DISS_SIGMAE")
    old_sigma = sigma.value
    sigma.value = manifest()["params"]["ko_sigma"]
    L.sgr_dropin_configure(1, 0, 0)
    try:
        def call(fn, name, bflag):
            u, sz, pmin, pmax, c = case_input(name)
            out = np.zeros_like(u)
            pi = (PD * 24)(*[u[f].ctypes.data_as(PD) for f in range(24)])
            po_ = (PD * 24)(*[out[f].ctypes.data_as(PD) for f in range(24)])
            off, bf = ctypes.c_uint(0), ctypes.c_uint(bflag)
            fn(po_, pi, ctypes.byref(off), (ctypes.c_double * 3)(*pmin), (ctypes.c_double * 3)(*pmax),
               (ctypes.c_uint * 3)(*sz), ctypes.byref(bf))
            return interior(out, sz)

        for t, name in ((1, "tvb3_staged_b63"), (2, "tvb5_staged_b63")):
            dtype.value = t
            print(f"DISSIPATION_TYPE={t}", flush=True)
            assert normwise(call(sep, name, 63), golden(name)).max() <= FP64_TOL
            assert normwise(call(plain, "gauge_standard_b63", 63), golden("gauge_standard_b63")).max() <= FP64_TOL
        assert L.sgr_dropin_status(None, 0) == 0
        dtype.value = 7
        got = call(sep, "tvb3_staged_b63", 63)
        assert np.isnan(got).all()
        assert L.sgr_dropin_status(None, 0) == -1  # SGR_E_INVALID
    finally:
        dtype.value = 0
        sigma.value = old_sigma
        L.sgr_dropin_configure(0, 0, 0)
