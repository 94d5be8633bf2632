"""GPU: the dense W-basis VJP with time-segment checkpoints (kernels_walkseg.hip).

magnus_vjp_segmented_f64 keeps the basis levels for one segment of samples at
a time.  Its dA is checked against the oracle at 1e-12 (relative Frobenius)
for dims 16, 32 and 64, ops many and sum, orders 2..8, one segment up to one
segment per Boole block; any segment count gives the same dA bit for bit; and
its scratch at S = 4097 and 8193 is a few percent of the n - 1 full levels the
unsegmented VJP keeps."""
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


@pytest.mark.parametrize("d", [16, 32, 64])
@pytest.mark.parametrize("op", ["many", "sum"])
@pytest.mark.parametrize("n,S,B", [(2, 9, 2), (4, 65, 3), (6, 33, 1), (8, 21, 2)])
def test_segmented_vjp_matches_oracle_for_every_segment_count(d, op, n, S, B):
    import torch

    from magnus import capi, torch_ops

    A = np.stack([fourier_matrices(np.random.default_rng([71, b, d, n]), S, d, modes=3) for b in range(B)])
    out_shape = (n, d, d) if op == "many" else (d, d)
    cot = np.stack([np.random.default_rng([71, b, d, n, 1]).standard_normal(out_shape) for b in range(B)])
    At, ct = torch.from_numpy(A).cuda(), torch.from_numpy(cot).cuda()
    blocks = (S - 1) // 4
    ref_dA = None
    for segs in sorted({1, 2, 3, blocks, 0}):
        dA = torch_ops.vjp_segmented(n, At, ct, 0.0, 1.0, op=op, integrator="Boole", segments=segs)
        assert capi.last_kernel() == "walkseg"
        if ref_dA is None:
            ref_dA = dA
        else:
            assert torch.equal(dA, ref_dA), segs
    got = ref_dA.cpu().numpy()
    for b in range(B):
        ref = O.vjp(n, A[b], cot[b], 0.0, 1.0, op=op, integrator="Boole")
        assert rel_fro(got[b], ref) <= TOL, b


@pytest.mark.parametrize("d,S,n", [(32, 4097, 6), (16, 8193, 6), (64, 8193, 6)])
def test_segmented_vjp_full_length_matches_the_unsegmented_vjp(d, S, n):
    """At the dense configurations' lengths: the segmented dA against the
    W-basis VJP that keeps every level (1e-12), and its scratch under 7 % of
    the n - 1 full levels."""
    import torch

    from magnus import torch_ops

    B = 2
    A = np.stack([fourier_matrices(np.random.default_rng([73, b, d]), S, d, modes=4) for b in range(B)])
    cot = np.stack([np.random.default_rng([73, b, d, 1]).standard_normal((d, d)) for b in range(B)])
    At, ct = torch.from_numpy(A).cuda(), torch.from_numpy(cot).cuda()
    dA = torch_ops.vjp_segmented(n, At, ct, 0.0, 1.0, op="sum", integrator="Boole").cpu().numpy()
    full = torch_ops.vjp(n, At, ct, None, 0.0, 1.0, op="sum", integrator="Boole").cpu().numpy()
    for b in range(B):
        assert rel_fro(dA[b], full[b]) <= TOL, b
    levels = B * (n - 1) * S * d * d * 8
    used = torch_ops.vjp_segmented_scratch_bytes(n, B, S, d, op="sum", integrator="Boole")
    assert 0 < used <= 0.07 * levels, (used, levels)


def test_carry_free_vjp_keeps_segments_when_the_levels_do_not_fit():
    """With the library's scratch bounded below the n - 1 full levels, the
    carry-free VJP falls back to the segmented one (same dA bit for bit as a
    direct segmented call, 1e-12 of the full-level VJP)."""
    import torch

    from magnus import capi, torch_ops

    d, S, n, B = 16, 257, 5, 2
    A = np.stack([fourier_matrices(np.random.default_rng([79, b]), S, d, modes=3) for b in range(B)])
    cot = np.stack([np.random.default_rng([79, b, 1]).standard_normal((d, d)) for b in range(B)])
    At, ct = torch.from_numpy(A).cuda(), torch.from_numpy(cot).cuda()
    full = torch_ops.vjp(n, At, ct, None, 0.0, 1.0, op="sum", integrator="Boole")
    assert capi.last_kernel() == "wbd"
    seg = torch_ops.vjp_segmented(n, At, ct, 0.0, 1.0, op="sum", integrator="Boole")
    need = torch_ops.vjp_segmented_scratch_bytes(n, B, S, d, op="sum", integrator="Boole")
    levels = B * (n - 1) * S * d * d * 8
    limit = need + need // 2
    assert limit < levels
    torch.cuda.synchronize()
    capi.set_scratch_limit(limit)
    try:
        got = torch_ops.vjp(n, At, ct, None, 0.0, 1.0, op="sum", integrator="Boole")
        assert capi.last_kernel() == "walkseg"
    finally:
        capi.set_scratch_limit(0)
    assert torch.equal(got, seg)
    g, f = got.cpu().numpy(), full.cpu().numpy()
    for b in range(B):
        assert rel_fro(g[b], f[b]) <= TOL, b


def test_segmented_vjp_rejects_other_problems():
    import torch

    from magnus import torch_ops

    A = torch.zeros((1, 9, 8, 8), dtype=torch.float64, device="cuda")
    with pytest.raises(ValueError):
        torch_ops.vjp_segmented(3, A, torch.zeros((1, 8, 8), dtype=torch.float64, device="cuda"), integrator="Boole")
    A = torch.zeros((1, 9, 16, 16), dtype=torch.float64, device="cuda")
    with pytest.raises(ValueError):
        torch_ops.vjp_segmented(3, A, torch.zeros((1, 16, 16), dtype=torch.float64, device="cuda"),
                                integrator="Simpson")
    assert torch_ops.vjp_segmented_scratch_bytes(3, 1, 9, 16, op="one", integrator="Boole") == 0
