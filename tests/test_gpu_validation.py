"""Device front end rejects mismatched cotangents, carries and checkpoints on
the host, before any kernel runs (the reference's validate_matrix_output_shape /
validate_vjp_data_shape, src/composer/backends.hpp:124-180): the C ABI reads
these buffers by the problem's shape, so a wrong one must never reach it."""
from __future__ import annotations

import numpy as np
import pytest

from helpers import fourier_spacecurve

pytestmark = pytest.mark.gpu


def _dense(B=2, S=17, d=3):
    import torch

    t = np.linspace(0.0, 1.0, S)
    A = np.stack([np.sin((b + 1) * t)[:, None, None] * np.ones((d, d)) + 0.1 * np.eye(d) for b in range(B)])
    return torch.from_numpy(A).cuda()


def test_matrix_vjp_rejects_bad_cotangent_and_carry():
    import torch

    from magnus import torch_ops

    A = _dense()
    n = 4
    out, carry = torch_ops.forward(n, A, op="sum", record_vjp=True)
    good = torch.ones_like(out)
    torch_ops.vjp(n, A, good, carry, op="sum")  # the valid call runs
    with pytest.raises(ValueError, match="GPU tensor"):
        torch_ops.vjp(n, A, good.cpu(), carry, op="sum")
    with pytest.raises(TypeError, match="float64"):
        torch_ops.vjp(n, A, good.float(), carry, op="sum")
    with pytest.raises(ValueError, match="shape"):  # (B, n, d, d) is op='many''s cotangent
        torch_ops.vjp(n, A, torch.ones((2, n, 3, 3), dtype=torch.float64, device=A.device), carry, op="sum")
    with pytest.raises(ValueError, match="shape"):
        torch_ops.vjp(n, A, good[:1], carry, op="sum")
    with pytest.raises(ValueError, match="shape"):  # carry of another order
        torch_ops.vjp(n, A, good, carry[:, :, :2].contiguous(), op="sum")
    with pytest.raises(ValueError, match="GPU tensor"):
        torch_ops.vjp(n, A, good, carry.cpu(), op="sum")
    with pytest.raises(ValueError, match="shape"):
        torch_ops.vjp(n, A, torch.ones((2, 3, 3), dtype=torch.float64, device=A.device), None, op="many")
    torch.cuda.synchronize()


def test_spacecurve_vjp_rejects_bad_cotangent_carry_and_checkpoint():
    import torch

    from magnus import torch_ops

    a = torch.from_numpy(np.stack([fourier_spacecurve(np.random.default_rng([41, b]), 129) for b in range(3)])).cuda()
    n = 5
    out, carry = torch_ops.sc_forward(n, a, op="many", integrator="Boole", record_vjp=True)
    cot = torch.ones_like(out)
    torch_ops.sc_vjp(n, a, cot, carry, op="many", integrator="Boole")
    with pytest.raises(ValueError, match="shape"):
        torch_ops.sc_vjp(n, a, cot[:, 0], carry, op="many", integrator="Boole")
    with pytest.raises(ValueError, match="shape"):
        torch_ops.sc_vjp(n, a, cot, carry[:2], op="many", integrator="Boole")
    with pytest.raises(ValueError, match="GPU tensor"):
        torch_ops.sc_vjp(n, a, cot.cpu(), None, op="many", integrator="Boole")
    _, ckpt = torch_ops.sc_forward_ckpt(n, a, op="many", integrator="Boole")
    torch_ops.sc_vjp_ckpt(n, a, cot, ckpt, op="many", integrator="Boole")
    # a checkpoint of a larger batch has another chunk layout: rejected, not misread
    big = torch.cat([a, a, a, a])
    _, ckpt_big = torch_ops.sc_forward_ckpt(n, big, op="many", integrator="Boole")
    assert ckpt_big.numel() != ckpt.numel()
    with pytest.raises(ValueError, match="exactly"):
        torch_ops.sc_vjp_ckpt(n, a, cot, ckpt_big, op="many", integrator="Boole")
    with pytest.raises(ValueError, match="exactly"):
        torch_ops.sc_vjp_ckpt(n, a, cot, ckpt[:-1], op="many", integrator="Boole")
    torch.cuda.synchronize()
