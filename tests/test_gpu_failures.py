"""Failure paths of the C ABI on the GPU, forced from the host (no kernel
is made to fault): a device allocation that fails part-way through plan
creation (SGR_FAULT_ALLOC=k fails the k-th allocation with out-of-memory),
after which the library keeps working."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_allocation_failure_at_every_step_is_reported_and_cleaned_up(monkeypatch):
    import torch
    import sympygr_amd as sgr
    from sympygr_amd import data
    bl = data.uniform_tiling(nb=2, n=16, lo=-1.0, hi=1.0, stagger=0.5, faces=True)
    u = data.fill_blocks(bl, backend="torch", device="cuda", kind="perturbed", seed=3, amp=1e-2, lo=-1.0, L=2.0)
    ref = torch.zeros_like(u)
    sgr.Plan(bl, sgr.params()).rhs(ref, u)
    torch.cuda.synchronize()
    free0 = torch.cuda.mem_get_info()[0]
    # blocks, RHS tile map, constraint tile map, workspace, flag, tile queue, ghost patches, ghost boxes
    for k in range(1, 9):
        monkeypatch.setenv("SGR_FAULT_ALLOC", str(k))
        print(f"fault at allocation {k}", flush=True)
        with pytest.raises(sgr.SgrError) as e:
            sgr.Plan(bl, sgr.params())
        assert e.value.code == -4 and "out of device memory" in str(e.value), str(e.value)
    monkeypatch.setenv("SGR_FAULT_ALLOC", "9")  # past the last allocation: no fault
    plan = sgr.Plan(bl, sgr.params())
    monkeypatch.delenv("SGR_FAULT_ALLOC")
    out = torch.zeros_like(u)
    plan.rhs(out, u)
    torch.cuda.synchronize()
    assert torch.equal(out, ref)
    del plan
    torch.cuda.synchronize()
    # the refused plans released what they had allocated (allow the
    # allocator's granularity)
    assert torch.cuda.mem_get_info()[0] >= free0 - (64 << 20)
