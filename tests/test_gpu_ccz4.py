"""GPU CCZ4 RHS (25 fields, Theta last) against the CPU oracle's CCZ4
(oracle/bssn_oracle.cpp with the generated oracle/gen/ccz4_algebra.inc,
pinned by the closed forms and the symbolic reduction to BSSN of
tests/test_ccz4_cpu.py).  Parity with the reference's own CCZ4 is unpinned:
the reference has no compiled CCZ4."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

KAPPA = (0.02, 0.0, 0.5)


def _block(n, lo, h, bflag):
    from sympygr_amd import data
    bl = data.BlockList()
    bl.add((n, n, n), (lo,) * 3, (lo + (n - 1) * h,) * 3, bflag)
    return bl


def _gpu(bl, u, precision="fp64", adv="upwind"):
    import torch
    import sympygr_amd as sgr
    plan = sgr.Plan(bl, sgr.params(adv=adv, precision=precision, formulation="ccz4", kappa=KAPPA))
    du = torch.from_numpy(np.ascontiguousarray(u)).cuda()
    out = torch.full_like(du, 7.0)
    plan.rhs(out, du)
    torch.cuda.synchronize()
    return out.cpu().numpy()


def _oracle(bl, u, adv="upwind"):
    import pyoracle as po
    P = po.make_params(upwind=adv == "upwind", ko_sigma=0.1, ccz4=True, kappa=KAPPA)
    return po.rhs_block(u, bl.sz[0], bl.pmin[0], bl.pmax[0], bl.bflag[0], P)


def _interior(a, n):
    return a.reshape(25, n, n, n)[:, 3:-3, 3:-3, 3:-3].reshape(25, -1)


def _err(a, b):
    return np.array([np.abs(a[f] - b[f]).max() / max(np.abs(b[f]).max(), 1e-300) for f in range(25)])


@pytest.mark.parametrize("bflag,adv", [(0, "upwind"), (63, "upwind"), (21, "centred")])
def test_ccz4_fp64_matches_oracle(bflag, adv):
    from sympygr_amd import data
    n = 22
    bl = _block(n, -1.05, 0.1, bflag)
    u = data.fill_blocks(bl, kind="ccz4", seed=4, amp=1e-2, lo=-1.05, L=2.1)
    got, ref = _gpu(bl, u, adv=adv), _oracle(bl, u, adv=adv)
    e = _err(_interior(got, n), _interior(ref, n))
    print(f"ccz4 fp64 bflag={bflag} {adv}: max normwise err {e.max():.2e} (Theta {e[24]:.2e})", flush=True)
    assert e.max() <= 1e-12, e
    ghost = np.ones((n, n, n), dtype=bool)
    ghost[3:-3, 3:-3, 3:-3] = False
    assert (got.reshape(25, n, n, n)[:, ghost] == 7.0).all()


def test_ccz4_mixed_error_bound():
    from sympygr_amd import data
    n = 22
    bl = _block(n, -1.05, 0.1, 0)
    u = data.fill_blocks(bl, kind="ccz4", seed=4, amp=1e-2, lo=-1.05, L=2.1)
    e = _err(_interior(_gpu(bl, u, precision="mixed"), n), _interior(_oracle(bl, u), n))
    print(f"ccz4 mixed: max normwise err {e.max():.2e}", flush=True)
    assert e.max() <= 1e-6, e


def test_ccz4_linear_gamma_hat_closed_form_on_gpu():
    """Flat metric, Theta = 0, Gamma-hat^i = a^i + b^i_j x^j: K = tr b,
    Theta = tr b / 2, At = (b + b^T)/2 - delta tr b / 3, Gamma-hat and B =
    -kappa1 Gamma-hat (tests/test_ccz4_cpu.py states the derivation)."""
    from sympygr_amd import data
    n = 17
    bl = _block(n, -0.8, 0.1, 0)
    x, y, z = data.block_coords(data._Np(), bl.sz[0], bl.pmin[0], bl.pmax[0])
    X = np.stack([x.reshape(-1), y.reshape(-1), z.reshape(-1)])
    u = np.zeros((25, n ** 3))
    for f in (0, 1, 12, 15, 17):
        u[f] = 1.0
    a = np.array([0.05, -0.02, 0.03])
    b = np.array([[0.3, -0.1, 0.2], [0.05, -0.4, 0.1], [-0.2, 0.15, 0.25]])
    for i in range(3):
        u[3 + i] = a[i] + b[i] @ X
    r = _interior(_gpu(bl, u), n)
    tr = np.trace(b)
    np.testing.assert_allclose(r[2], tr, rtol=1e-12, atol=1e-13)
    np.testing.assert_allclose(r[24], 0.5 * tr, rtol=1e-12, atol=1e-13)
    G = _interior(u, n)[3:6]
    for i in range(3):
        np.testing.assert_allclose(r[3 + i], -KAPPA[0] * G[i], rtol=1e-12, atol=1e-15)
