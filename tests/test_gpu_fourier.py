"""GPU: Fourier-form input to the d = 8 single-pass sweeps.

magnus_fourier_forward_basis_f64 / _vjp_basis_f64 / _forward_orders_from_f64
form each sample of A(t) = C_0 + sum_k (sin(2 pi k t) C_{2k-1} + cos(2 pi k t)
C_{2k}) inside the sweeps' loader wavefront instead of reading a sample array.
They must give the results of sampling first (magnus_sample_fourier_f64) and
running the sample-array entry points, bit for bit: outputs, the kept basis
state, dA and every order's sum; one block up to 512 blocks, 0 to 8 modes,
orders 2..8, ops many and sum, a grid not starting at 0.  Problems off the
single-pass path and too many modes are refused before any launch.
"""
from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _coeffs(B, modes, seed, d=8):
    import torch

    rng = np.random.default_rng(seed)
    c = rng.standard_normal((B, 2 * modes + 1, d, d)) / np.sqrt(d)
    return torch.from_numpy(c).cuda()


@pytest.mark.parametrize("n,S,modes,op", [(2, 5, 3, "sum"), (4, 9, 8, "many"), (8, 129, 5, "sum"), (6, 2049, 8, "sum"),
                                          (3, 33, 0, "many"), (8, 257, 1, "many")])
def test_fourier_sweeps_equal_sampling_first_bitwise(n, S, modes, op):
    import torch

    from magnus import torch_ops

    B = 5
    t0, tf = 0.25, 1.75
    c = _coeffs(B, modes, [83, n, S, modes])
    A = torch_ops.sample_fourier(c, S, t0, tf)
    shape = (B, n, 8, 8) if op == "many" else (B, 8, 8)
    cot = torch.from_numpy(np.random.default_rng([83, n, S, 1]).standard_normal(shape)).cuda()
    o1, st1 = torch_ops.forward_basis(n, A, t0, tf, op=op, integrator="Boole")
    o2, st2 = torch_ops.fourier_forward_basis(n, c, S, t0, tf, op=op, integrator="Boole")
    torch.cuda.synchronize()
    assert torch.equal(o1, o2)
    # the state's written parts: the kept (even) levels [B][(n-1)//2][S][64],
    # sums [B][n][64] rows 1..n-1, Z_0 [B][64], totals [B][n][64] rows 0..n-2
    lv = B * ((n - 1) // 2) * S * 64
    assert torch.equal(st1[:lv], st2[:lv])
    z1, z2 = st1[lv:lv + B * n * 64].view(B, n, 64), st2[lv:lv + B * n * 64].view(B, n, 64)
    assert torch.equal(z1[:, 1:], z2[:, 1:])
    o = lv + B * n * 64
    assert torch.equal(st1[o:o + B * 64], st2[o:o + B * 64])
    o += B * 64
    t1, t2 = st1[o:o + B * n * 64].view(B, n, 64), st2[o:o + B * n * 64].view(B, n, 64)
    assert torch.equal(t1[:, :n - 1], t2[:, :n - 1])
    g1 = torch_ops.vjp_basis(n, A, cot, st1, t0, tf, op=op, integrator="Boole")
    g2 = torch_ops.fourier_vjp_basis(n, c, S, cot, st2, t0, tf, op=op, integrator="Boole")
    torch.cuda.synchronize()
    assert torch.equal(g1, g2)
    for first in sorted({1, 2, n}):
        r1 = torch_ops.forward_orders(n, A, t0, tf, integrator="Boole", first=first)
        r2 = torch_ops.fourier_forward_orders(n, c, S, t0, tf, integrator="Boole", first=first)
        torch.cuda.synchronize()
        assert torch.equal(r1, r2), first


def test_fourier_sweeps_refuse_other_problems():
    import torch

    from magnus import torch_ops

    c = _coeffs(2, 3, 89)
    with pytest.raises(Exception):  # order 9 is beyond the single-pass sweeps
        torch_ops.fourier_forward_orders(9, c, 33, integrator="Boole")
    with pytest.raises(Exception):  # Simpson
        torch_ops.fourier_forward_basis(3, c, 33, integrator="Simpson")
    with pytest.raises(Exception):  # dim 16
        torch_ops.fourier_forward_basis(3, _coeffs(2, 3, 89, d=16), 33, integrator="Boole")
    with pytest.raises(Exception):  # nine modes
        torch_ops.fourier_forward_basis(3, _coeffs(2, 9, 89), 33, integrator="Boole")
    torch.cuda.synchronize()
