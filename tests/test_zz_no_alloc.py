"""The allocation-free single-block entry (sgr_rhs_block_noalloc) against
the allocating one (sgr_rhs_block), through the C ABI: two calls in a row on
the same device, the second with a smaller block (the scratch keeps the first
call's contents), each bit-equal to the allocating entry; plus the size
limit and a CCZ4 block."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _case(n, seed, off):
    from sympygr_amd import data
    pmin = (-0.7, -0.6, -0.9)
    pmax = tuple(p + 0.05 * (n - 1) for p in pmin)
    u = data.perturbed_flat_block(seed=seed, n=n, amp=1e-2, pmin=pmin, pmax=pmax)
    big = np.zeros((24, u.shape[1] + off))
    big[:, off:] = u
    return big, pmin, pmax


def test_noalloc_entry_equals_allocating_entry_twice_then_smaller():
    import torch
    import sympygr_amd as sgr
    for k, (n, bflag, prec) in enumerate([(38, 63, "fp64"), (22, 5, "fp64"), (38, 0, "mixed"), (19, 42, "mixed")]):
        off = 100 * (k + 1)
        big, pmin, pmax = _case(n, 3 + k, off)
        sz = (n, n, n)
        prm = sgr.params(adv="upwind", precision=prec)
        du = torch.from_numpy(big).cuda()
        ref = torch.full_like(du, 3.0)
        sgr.rhs_block(ref, du, off, pmin, pmax, sz, bflag, prm)
        got = torch.full_like(du, 3.0)
        print(f"noalloc call {k}: n={n} bflag={bflag} {prec}", flush=True)
        sgr.rhs_block_noalloc(got, du, off, pmin, pmax, sz, bflag, prm)
        torch.cuda.synchronize()
        assert torch.equal(got, ref), k


def test_noalloc_ccz4_block_and_size_limit():
    import torch
    import sympygr_amd as sgr
    from sympygr_amd import data
    n = 22
    bl = data.BlockList()
    bl.add((n, n, n), (-1.05,) * 3, (1.05,) * 3, 63)
    u = data.fill_blocks(bl, backend="torch", device="cuda", kind="ccz4", seed=4, amp=1e-2, lo=-1.05, L=2.1)
    prm = sgr.params(adv="upwind", formulation="ccz4")
    ref = torch.zeros_like(u)
    sgr.Plan(bl, prm).rhs(ref, u)
    got = torch.zeros_like(u)
    sgr.rhs_block_noalloc(got, u, 0, (-1.05,) * 3, (1.05,) * 3, (n, n, n), 63, prm)
    torch.cuda.synchronize()
    assert torch.equal(got, ref)
    m = 54  # 48^3 interior: 216 tiles, over the scratch
    big = torch.zeros((24, m ** 3), dtype=torch.float64, device="cuda")
    with pytest.raises(sgr.SgrError, match="scratch"):
        sgr.rhs_block_noalloc(big, big.clone(), 0, (0.0,) * 3, (1.0,) * 3, (m, m, m), 0, sgr.params())
