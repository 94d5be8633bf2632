"""GPU: the fused level walkers of the per-node recursion (d = 16, 32, Boole).

A record_vjp forward, the VJP from its carry and the VJP::none recompute run
one workgroup per (instance, GL node) that walks each level's samples once
(kernels_walk.hip).  Every output -- forward, carry, dA -- is checked against
the oracle at 1e-12 relative Frobenius, over every op, orders from 2 (the top
level is also the last) to 8, one to eight Boole blocks and batches whose job
count is not a multiple of the eight XCDs; the launch counts show that the
walkers, not the product + scan level kernels, ran."""
from __future__ import annotations

import numpy as np
import pytest

import oracle as O
from helpers import fourier_matrices, rel_fro

pytestmark = pytest.mark.gpu

TOL = 1e-12


@pytest.fixture(scope="module", autouse=True)
def numpy_gl_table():
    total = 64 * 65 // 2
    w = np.empty(total)
    x = np.empty(total)
    off = 0
    for order in range(1, 65):
        xr, wr = np.polynomial.legendre.leggauss(order)
        x[off:off + order] = (xr + 1) / 2
        w[off:off + order] = wr / 2
        off += order
    O.gl_replace(64, w, x)
    yield
    O.gl_reset()


@pytest.mark.parametrize("d", [16, 32])
@pytest.mark.parametrize("op", ["one", "many", "sum"])
@pytest.mark.parametrize("n,S,B", [(2, 5, 3), (3, 33, 1), (5, 17, 2), (8, 9, 3)])
def test_walkers_match_oracle(d, op, n, S, B):
    import torch

    from magnus import capi, torch_ops

    A = np.stack([fourier_matrices(np.random.default_rng([61, b, d, n]), S, d, modes=3) for b in range(B)])
    out_shape = (n, d, d) if op == "many" else (d, d)
    cot = np.stack([np.random.default_rng([61, b, d, n, 1]).standard_normal(out_shape) for b in range(B)])
    At, ct = torch.from_numpy(A).cuda(), torch.from_numpy(cot).cuda()
    capi.timing_reset()
    capi.timing_enable(True)
    try:
        out, carry = torch_ops.forward(n, At, 0.0, 1.0, op=op, integrator="Boole", record_vjp=True)
        dA = torch_ops.vjp(n, At, ct, carry, 0.0, 1.0, op=op, integrator="Boole")
        torch.cuda.synchronize()
    finally:
        capi.timing_enable(False)
    Q = (n + 3) // 2
    assert capi.timing_read("walk_fwd")[0] == n
    assert capi.timing_read("walk_rev")[0] == n - 1
    assert capi.timing_read("lm_fwd_prod")[0] == 0 and capi.timing_read("lm_vjp_prod")[0] == 0
    by, fl = capi.timing_counts("walk_rev")
    per = 8 * S * d * d * B * Q
    assert by == per * (4 + 6 * (n - 2))
    out, carry, dA = (x.cpu().numpy() for x in (out, carry, dA))
    for b in range(B):
        ref, ref_carry = O.forward(n, A[b], 0.0, 1.0, op=op, integrator="Boole", record_vjp=True)
        ref_g = O.vjp(n, A[b], cot[b], 0.0, 1.0, op=op, integrator="Boole", vjp_data=ref_carry)
        assert rel_fro(out[b], ref) <= TOL, b
        assert rel_fro(carry[b], ref_carry) <= TOL, b
        assert rel_fro(dA[b], ref_g) <= TOL, b


@pytest.mark.parametrize("d", [16, 32])
def test_walker_recompute_vjp_matches_carry_vjp(d):
    """VJP::none with op 'one' keeps the per-node recursion: its recompute runs
    the forward walkers without outputs; the result equals the VJP from a
    recorded carry bit for bit (same kernels on the same carry)."""
    import torch

    from magnus import torch_ops

    B, S, n = 2, 41, 6
    A = np.stack([fourier_matrices(np.random.default_rng([67, b, d]), S, d, modes=4) for b in range(B)])
    cot = np.stack([np.random.default_rng([67, b, d, 1]).standard_normal((d, d)) for b in range(B)])
    At, ct = torch.from_numpy(A).cuda(), torch.from_numpy(cot).cuda()
    _, carry = torch_ops.forward(n, At, 0.0, 1.0, op="one", integrator="Boole", record_vjp=True)
    g_carry = torch_ops.vjp(n, At, ct, carry, 0.0, 1.0, op="one", integrator="Boole")
    g_free = torch_ops.vjp(n, At, ct, None, 0.0, 1.0, op="one", integrator="Boole")
    assert torch.equal(g_carry, g_free)
    for b in range(B):
        ref = O.vjp(n, A[b], cot[b], 0.0, 1.0, op="one", integrator="Boole")
        assert rel_fro(g_free[b].cpu().numpy(), ref) <= TOL
