"""GPU: the learned resolution controller -- all-orders forward, features,
MLP gradients and optimizer step, and what the trained network gains.

The all-orders forward is checked against the oracle at 1e-12 for every
order; the features against a numpy restatement (exact SVD norms, numpy FFT);
the MLP's gradients against a float64 numpy backprop through the same
bfloat16-rounded weights, and its Adam step against the closed form."""
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


@pytest.mark.parametrize("integrator", ["Boole", "Simpson", "Trapezoid", "Riemann"])
@pytest.mark.parametrize("dim", [3, 8])
def test_forward_orders_matches_oracle_per_order(dim, integrator):
    import torch

    from magnus import torch_ops

    B, S, n = 3, 129, 6
    A = np.stack([fourier_matrices(np.random.default_rng([31, b, dim]), S, dim, modes=4) for b in range(B)])
    out = torch_ops.forward_orders(n, torch.from_numpy(A).cuda(), 0.0, 1.0, integrator=integrator).cpu().numpy()
    assert out.shape == (B, n, dim, dim)
    for b in range(B):
        for k in range(1, n + 1):
            ref = O.forward(k, A[b], 0.0, 1.0, op="sum", integrator=integrator)
            assert rel_fro(out[b, k - 1], ref) <= TOL, (b, k)


def _features_numpy(A, T):
    G = A.shape[0]
    N = G - 1
    spec = np.linalg.norm(A, 2, axis=(1, 2))
    fro = np.linalg.norm(A, axis=(1, 2))
    f = np.empty(24)
    f[0] = np.log(1e-30 + spec.max() * T)
    f[1] = np.log(1e-30 + spec.mean() * T)
    f[2] = np.log(1e-30 + fro.max() * T)
    w = np.ones(G)
    w[0] = w[-1] = 0.5
    f[3] = np.log(1e-30 + (w * fro).sum() / N * T)
    X = np.fft.fft(A[:N], axis=0)
    E = (np.abs(X) ** 2).sum(axis=(1, 2))
    tot = N * (A[:N] ** 2).sum()
    for k in range(1, 17):
        f[3 + k] = np.log(1e-12 + E[k] / (1e-30 + tot))
    ks = np.arange(1, 17)
    f[20] = np.log(1e-12 + (ks * E[1:17]).sum() / (1e-30 + E[1:17].sum()))
    pairs = [(0, N // 4), (0, N // 2), (N // 4, 3 * N // 4), (N // 8, 5 * N // 8)]
    c = max(np.linalg.norm(A[i] @ A[j] - A[j] @ A[i]) for i, j in pairs)
    f[21] = np.log(1e-30 + c * T * T)
    f[22] = np.log(1e-30 + np.linalg.norm(A[:N].mean(0)) * T)
    f[23] = np.log(1e-30 + c / (1e-30 + fro.max() ** 2))
    return f


def test_features_match_numpy():
    import torch

    from magnus import controller

    B, G, d = 6, 65, 8
    A = np.stack([fourier_matrices(np.random.default_rng([37, b]), G, d, modes=1 + b, tf=1.5) for b in range(B)])
    got = controller.features(torch.from_numpy(A).cuda(), 0.0, 1.5).cpu().numpy()
    for b in range(B):
        ref = _features_numpy(A[b], 1.5)
        np.testing.assert_allclose(got[b], ref, rtol=0, atol=2e-3)


def _bf16(x):
    u = np.ascontiguousarray(x, np.float32).view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32)


def _mlp_numpy(p, x, y, dims):
    """float64 forward + backward through the bfloat16-rounded weights."""
    i, h, o = dims
    offs = np.cumsum([0, h * i, h, h * h, h, o * h, o])
    W1, b1, W2, b2, W3, b3 = (p[offs[k]:offs[k + 1]].astype(np.float64) for k in range(6))
    W1 = _bf16(W1.astype(np.float32)).astype(np.float64).reshape(h, i)
    W2 = _bf16(W2.astype(np.float32)).astype(np.float64).reshape(h, h)
    W3 = _bf16(W3.astype(np.float32)).astype(np.float64).reshape(o, h)
    x = x.astype(np.float64)
    h1 = np.maximum(x @ W1.T + b1, 0)
    h2 = np.maximum(h1 @ W2.T + b2, 0)
    z = h2 @ W3.T + b3
    z = z - z.max(1, keepdims=True)
    pz = np.exp(z) / np.exp(z).sum(1, keepdims=True)
    B = x.shape[0]
    loss = -np.log(pz[np.arange(B), y])
    dz = pz.copy()
    dz[np.arange(B), y] -= 1
    dz /= B
    g = [dz.T @ h2, dz.sum(0)]
    dh2 = (dz @ W3) * (h2 > 0)
    g = [dh2.T @ h1, dh2.sum(0)] + g
    dh1 = (dh2 @ W2) * (h1 > 0)
    g = [dh1.T @ x, dh1.sum(0)] + g
    return loss, np.concatenate([t.ravel() for t in g])


def test_mlp_gradients_and_adam_step_match_fp64_oracle():
    import torch

    from magnus.controller import MLP

    dims = (24, 128, 63)
    net = MLP(*dims, max_batch=512, seed=3)
    p0 = net.params()
    rng = np.random.default_rng(41)
    x = rng.standard_normal((300, dims[0])).astype(np.float32)
    y = rng.integers(0, dims[2], 300).astype(np.int32)
    loss = net.train_step(torch.from_numpy(x).cuda(), torch.from_numpy(y).cuda(), 1e-3).cpu().numpy()
    g = net.grads()
    ref_loss, ref_g = _mlp_numpy(p0, x, y, dims)
    np.testing.assert_allclose(loss, ref_loss, rtol=1e-4, atol=1e-5)
    i, h, o = dims
    offs = np.cumsum([0, h * i, h, h * h, h, o * h, o])
    for k in range(6):
        a, b = g[offs[k]:offs[k + 1]], ref_g[offs[k]:offs[k + 1]]
        assert np.linalg.norm(a - b) <= 1e-4 * np.linalg.norm(b) + 1e-9, k
    # first Adam step: m = 0.1 g, v = 0.001 g^2, bias corrections 0.1 and 0.001
    p1 = net.params()
    gg = g.astype(np.float64)
    step = 1e-3 * (0.1 * gg / 0.1) / (np.sqrt(0.001 * gg * gg / 0.001) + 1e-8)
    np.testing.assert_allclose(p1, p0 - step, rtol=0, atol=2e-6)
    net.close()


def test_controller_gains_over_the_judge_without_it():
    """A network trained on a small labelled set starts the judge at (or near)
    the resolution it accepts, where the judge alone starts from (65, 2): under
    half the verification rounds, under 70% of their sweep work, and at least 97% of the
    accepted results within the tolerance of the S = 32769, n = 8 reference."""
    import torch

    from magnus import controller as C
    from magnus import torch_ops

    tol = 1e-8
    coeffs = C.make_coefficients(2600, seed=5)
    feats = C.features(torch_ops.sample_fourier(coeffs, C.FEATURE_GRID), 0.0, 1.0)
    labels, minimal, reach, errs, refs = C.label(coeffs, tol)
    tr, ev = slice(0, 2304), slice(2304, 2600)
    ctl = C.Controller(feats.shape[1], seed=1)
    hist = ctl.fit(feats[tr], torch.from_numpy(labels[tr]).cuda(), epochs=60, batch=256, lr=3e-3)
    assert hist[-1] < hist[0]
    pred = ctl.predict(feats[ev]).cpu().numpy()
    start = [C.CLASSES[k] for k in pred]
    res_l, out_l, st_l = C.accept(coeffs[ev], start, tol)
    res_0, out_0, st_0 = C.accept(coeffs[ev], [(65, 2)] * len(start), tol)
    work_l, work_0 = st_l["work"], st_0["work"]
    hit_l = float((C.rel_err(out_l, refs[ev]) <= tol).float().mean())
    hit_0 = float((C.rel_err(out_0, refs[ev]) <= tol).float().mean())
    assert work_l < 0.7 * work_0, (work_l, work_0)
    assert st_l["rounds"].mean() < 0.5 * st_0["rounds"].mean()
    assert hit_l >= 0.97 and hit_l >= hit_0 - 0.02, (hit_l, hit_0)


def test_judge_delivers_its_class_sum_and_its_vjp_from_the_kept_state():
    """accept(..., cot=...) delivers, per instance, the sum of its accepted
    class (S, n) -- the coarser member of the doubling pair -- and the VJP of
    that sweep from its kept basis state: both equal the one-problem forward
    and VJP at the delivered resolution bit for bit (the VJP of a rejected
    round is replaced by the accepting round's)."""
    import torch

    from magnus import controller as C
    from magnus import torch_ops

    tol = 1e-8
    B = 24
    coeffs = C.make_coefficients(B, seed=7)
    cot = torch.from_numpy(np.random.default_rng(7).standard_normal((B, 8, 8))).cuda()
    start = [(129, 3)] * (B // 2) + [(1025, 8)] * (B - B // 2)  # low starts escalate
    res, out, st = C.accept(coeffs, start, tol, cot=cot)
    assert st["rounds"].max() > 1
    for i, (S, n) in enumerate(res):
        A = torch_ops.sample_fourier(coeffs[i:i + 1], S)
        assert torch.equal(out[i], torch_ops.forward(n, A, 0.0, 1.0, op="sum", integrator="Boole")[0]), i
        ref = torch_ops.vjp(n, A, cot[i:i + 1], None, 0.0, 1.0, op="sum", integrator="Boole")[0]
        assert torch.equal(C.dA_of(st, i), ref), i
