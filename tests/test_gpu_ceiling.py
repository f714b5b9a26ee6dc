"""The matrix-core ceiling probe (mgx_mfma_i8_ceiling) through the C ABI:
with every operand byte 1 each v_mfma_i32_16x16x64_i8 adds 64 to each of
the 16 x 16 outputs, so every thread's sum over its 32 accumulator
fragments (4 elements each) after 32 * iters MFMAs is 32 * 4 * 64 * iters;
random operands (re-read from LDS every round) give a deterministic result."""
import pytest

from tests.gpu_util import announce

pytestmark = pytest.mark.gpu


def test_constant_operands_closed_form(mgx):
    import torch
    blocks, iters = 4, 10
    announce(f"mfma ceiling constant blocks={blocks} iters={iters}")
    out = torch.full((blocks * 512,), -1, dtype=torch.int32, device="cuda")
    mgx._lib.call("mgx_mfma_i8_ceiling", out.data_ptr(), blocks, iters, 0, None)
    torch.cuda.synchronize()
    assert bool(torch.all(out == 32 * 4 * 64 * iters))


def test_random_operands_deterministic(mgx):
    import torch
    blocks, iters = 8, 6
    announce(f"mfma ceiling random blocks={blocks} iters={iters}")
    a = torch.zeros(blocks * 512, dtype=torch.int32, device="cuda")
    b = torch.ones(blocks * 512, dtype=torch.int32, device="cuda")
    mgx._lib.call("mgx_mfma_i8_ceiling", a.data_ptr(), blocks, iters, 1, None)
    mgx._lib.call("mgx_mfma_i8_ceiling", b.data_ptr(), blocks, iters, 1, None)
    torch.cuda.synchronize()
    assert torch.equal(a, b)
    # operands differ per lane and per workgroup; waves w and w + 4 of a
    # workgroup read the same LDS regions, so about half the sums repeat
    assert int(torch.unique(a).numel()) > 0.4 * blocks * 512
