"""GPU tests of the ghost refresh across blocks of other sizes, spacings and
levels (6-point Lagrange from the owning block), and of RK4 steps on such a
block set, against the oracle."""
import numpy as np
import pytest

from helpers import coords, ghost_mask, poly5, small_octree

pytestmark = pytest.mark.gpu


def _scramble(u, bl, seed=0):
    mk = ghost_mask(bl, inside_only=False)
    v = u.copy()
    v[:, mk] = np.random.default_rng(seed).standard_normal((24, int(mk.sum())))
    return v


def test_octree_ghost_fill_matches_oracle_bitwise():
    import torch
    import pyoracle as po
    import sympygr_amd as sgr
    from sympygr_amd import data
    bl = small_octree()
    u = _scramble(data.fill_blocks(bl, kind="perturbed", seed=5, amp=0.3, lo=-2.0, L=4.0), bl)
    plan = sgr.Plan(bl, sgr.params(adv="upwind"))
    du = torch.from_numpy(u).cuda()
    plan.ghost_fill(du)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(du.cpu().numpy(), po.ghost_fill(u, bl))


def test_octree_ghost_fill_exact_on_polynomials_at_scale():
    """848-leaf octree over [-16,16]^3 (16.9M ghost nodes, 77% interpolated):
    every ghost node inside the domain reproduces a degree-5 polynomial."""
    import torch
    import sympygr_amd as sgr
    from sympygr_amd import data
    bl = data.octree_blocks(lmax=4)
    plan = sgr.Plan(bl, sgr.params(adv="upwind"))
    X = torch.from_numpy(coords(bl) / 8.0).cuda()
    u = torch.stack([poly5(X, f) for f in range(24)])
    del X
    inside = torch.from_numpy(ghost_mask(bl)).cuda()
    allg = torch.from_numpy(ghost_mask(bl, inside_only=False)).cuda()
    v = u.clone()
    v[:, allg] = 12345.0
    plan.ghost_fill(v)
    torch.cuda.synchronize()
    err = ((v[:, inside] - u[:, inside]).abs().amax() / u.abs().amax()).item()
    print(f"polynomial ghost error {err:.2e}", flush=True)
    assert err < 1e-12
    assert torch.equal(v[:, ~allg], u[:, ~allg])
    assert bool((v[:, allg & ~inside] == 12345.0).all())


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_rk4_steps_on_octree_match_oracle(precision):
    import torch
    import pyoracle as po
    import sympygr_amd as sgr
    from sympygr_amd import data
    bl = small_octree()
    u = data.fill_blocks(bl, kind="perturbed", seed=3, amp=1e-2, lo=-2.0, L=4.0)
    hmin = min((bl.pmax[b][0] - bl.pmin[b][0]) / (bl.sz[b][0] - 1) for b in range(len(bl)))
    dt = 0.25 * hmin
    plan = sgr.Plan(bl, sgr.params(adv="upwind", precision=precision, ko_sigma=0.1))
    du = torch.from_numpy(u).cuda()
    bufs = [torch.zeros_like(du) for _ in range(3)]
    P = po.make_params(upwind=True, ko_sigma=0.1)
    ref = u
    mk = ~ghost_mask(bl, inside_only=False)
    # error of the change over the steps relative to the change (per field):
    # FP64 agrees to rounding; mixed stages its algebra in FP32
    tol = 1e-10 if precision == "fp64" else 1e-5
    for step in range(2):
        plan.rk4_step(dt, du, *bufs)
        torch.cuda.synchronize()
        ref = po.rk4_step(ref, bl, None, dt, P, nthreads=8)
        got = du.cpu().numpy()
        # compare the change over the step, per field, relative to its size
        d_ref = ref[:, mk] - u[:, mk]
        d_got = got[:, mk] - u[:, mk]
        e = max(np.abs(d_got[f] - d_ref[f]).max() / max(np.abs(d_ref[f]).max(), 1e-300) for f in range(24))
        print(f"{precision} step {step} err {e:.2e}", flush=True)
        assert e <= tol
