"""GPU: uneven batches (magnus_uneven_*) -- many matrix problems of different
dim, samples, order, op and integrator in one enqueue or one hipGraph launch.

A bucketed instance's forward output and dA must equal a per-instance call of
the one-problem entry points bit for bit (same kernels, same per-job
arithmetic); an instance of the grouped walker pipeline (dense dim 16, 32,
64, Boole, op many or sum) must agree with it at 1e-12 relative Frobenius (same
recurrences, matrix-core products in another summation order) and with the
oracle.  Every replay of the captured graph repeats the eager results bit for
bit; product-path variants chosen by the tuner are checked against the oracle
(1e-12).
"""
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


def _mixed_problems(seed, count):
    rng = np.random.default_rng(seed)
    dims = [2, 3, 4, 8, 16, 32]
    samples = [33, 65, 129]
    integs = ["Boole", "Simpson", "Trapezoid", "Riemann"]
    ops = ["sum", "many", "one"]
    probs = []
    for i in range(count):
        probs.append(dict(n=int(rng.integers(1, 6)), dim=int(rng.choice(dims)), samples=int(rng.choice(samples)),
                          t0=0.0, tf=1.0, op=str(rng.choice(ops)), integrator=str(rng.choice(integs))))
    # repeated shapes so that buckets hold several instances, and one bucket of
    # instances too large to share a small workspace
    big = dict(n=5, dim=32, samples=129, t0=0.0, tf=1.0, op="sum", integrator="Simpson")
    return probs + [dict(p) for p in probs[: count // 2]] + [dict(big) for _ in range(3)]


def _inputs(probs, seed):
    import torch

    A = [torch.from_numpy(fourier_matrices(np.random.default_rng([seed, i]), p["samples"], p["dim"], modes=3)).cuda()
         for i, p in enumerate(probs)]
    cot = []
    for i, p in enumerate(probs):
        shape = (p["n"], p["dim"], p["dim"]) if p["op"] == "many" else (p["dim"], p["dim"])
        cot.append(torch.from_numpy(np.random.default_rng([seed, i, 1]).standard_normal(shape)).cuda())
    return A, cot


def _per_instance(probs, A, cot):
    from magnus import torch_ops

    outs, dAs = [], []
    for p, a, c in zip(probs, A, cot):
        kw = dict(op=p["op"], integrator=p["integrator"])
        outs.append(torch_ops.forward(p["n"], a[None], p["t0"], p["tf"], **kw)[0])
        dAs.append(torch_ops.vjp(p["n"], a[None], c[None], None, p["t0"], p["tf"], **kw)[0])
    return outs, dAs


def _same(plan, i, got, ref):
    import torch

    if plan.in_pipeline(i):
        return rel_fro(got.cpu().numpy(), ref.cpu().numpy()) <= TOL
    return torch.equal(got, ref)


@pytest.mark.parametrize("branches,workspace", [(1, 0), (3, 0), (4, 96 << 20)])
def test_uneven_matches_per_instance_calls(branches, workspace):
    """Eager and graph modes against per-instance calls; the small workspace
    (what the pipeline needs, then under 24 MB per branch against about 9 MB
    per bucketed dim-32 instance) forces sub-batches and arena reuse."""
    import torch

    from magnus.uneven import UnevenBatch

    probs = _mixed_problems(3, 24)
    A, cot = _inputs(probs, 3)
    ref_out, ref_dA = _per_instance(probs, A, cot)
    plan = UnevenBatch(probs, workspace_bytes=workspace, branches=branches)
    info = plan.info()
    assert info["buckets"] <= len(probs) and info["sub_batches"] >= info["buckets"]
    assert info["pipeline_instances"] == sum(plan.in_pipeline(i) for i in range(len(probs))) > 0
    if workspace:
        assert info["sub_batches"] > info["buckets"]
    first = None
    for use_graph in (False, True):
        for rep in range(3 if use_graph else 1):
            outs = plan.forward(A, use_graph=use_graph)
            dAs = plan.vjp(A, cot, use_graph=use_graph)
            torch.cuda.synchronize()
            for i in range(len(probs)):
                assert _same(plan, i, outs[i], ref_out[i]), (use_graph, rep, i, probs[i])
                assert _same(plan, i, dAs[i], ref_dA[i]), (use_graph, rep, i, probs[i])
            if first is None:
                first = ([x.clone() for x in outs], [x.clone() for x in dAs])
            else:
                assert all(torch.equal(x, y) for x, y in zip(outs, first[0])), (use_graph, rep)
                assert all(torch.equal(x, y) for x, y in zip(dAs, first[1])), (use_graph, rep)
            # eager work of the program between graph replays
            _ = sum(float(x.sum()) for x in outs)
    plan.close()


@pytest.mark.parametrize("workspace", [0, 1 << 30])
def test_uneven_pipeline_matches_oracle_within_workspace(workspace):
    """The grouped walker pipeline: dims 16, 32 and 64, one segment (5, 9
    samples), a short last segment (133 samples: 128 + 4), several segments,
    orders 2..6, ops many and sum.  With room for all, every instance runs in
    the pipeline; in 1 GiB the shortest chains join it and the rest run in
    buckets.  Against per-instance calls (1e-12, bit for bit when bucketed),
    the oracle (1e-12, the smaller instances) and across graph replays (bit for
    bit)."""
    import torch

    from magnus.uneven import UnevenBatch

    probs = []
    for d in (16, 32, 64):
        for S in (5, 9, 133, 261, 1025):
            for n in (2, 3, 6):
                probs.append(dict(n=n, dim=d, samples=S, t0=0.0, tf=1.0, op="sum" if (S + n) % 2 else "many",
                                  integrator="Boole"))
    A, cot = _inputs(probs, 13)
    ref_out, ref_dA = _per_instance(probs, A, cot)
    plan = UnevenBatch(probs, workspace_bytes=workspace, branches=2)
    info = plan.info()
    inside = [plan.in_pipeline(i) for i in range(len(probs))]
    assert info["pipeline_instances"] == sum(inside) and info["pipeline_waves"] == 1
    if workspace:
        assert 0 < sum(inside) < len(probs) and info["buckets"] > 0
        # the shortest chains join first: no bucketed instance has a shorter one than a pipeline instance
        chain = [(p["samples"] + (p["n"] - 1) * 128) * p["dim"] ** 3 for p in probs]
        assert max(c for c, x in zip(chain, inside) if x) <= min(c for c, x in zip(chain, inside) if not x)
    else:
        assert all(inside) and info["buckets"] == 0
    outs = plan.forward(A, use_graph=False)
    dAs = plan.vjp(A, cot, use_graph=False)
    torch.cuda.synchronize()
    outs = [x.clone() for x in outs]
    dAs = [x.clone() for x in dAs]
    for i, p in enumerate(probs):
        assert _same(plan, i, outs[i], ref_out[i]), (i, p)
        assert _same(plan, i, dAs[i], ref_dA[i]), (i, p)
        if p["samples"] <= 261 and (p["dim"] == 16 or p["n"] <= 3):
            a = A[i].cpu().numpy()
            ref = O.forward(p["n"], a, 0.0, 1.0, op=p["op"], integrator="Boole")
            ref_g = O.vjp(p["n"], a, cot[i].cpu().numpy(), 0.0, 1.0, op=p["op"], integrator="Boole")
            assert rel_fro(outs[i].cpu().numpy(), ref) <= TOL, (i, p)
            assert rel_fro(dAs[i].cpu().numpy(), ref_g) <= TOL, (i, p)
    for rep in range(3):
        g_out = plan.forward(A, use_graph=True)
        g_dA = plan.vjp(A, cot, use_graph=True)
        torch.cuda.synchronize()
        assert all(torch.equal(x, y) for x, y in zip(g_out, outs)), rep
        assert all(torch.equal(x, y) for x, y in zip(g_dA, dAs)), rep
        _ = sum(float(x.sum()) for x in g_out)
    plan.close()


@pytest.mark.parametrize("dim,tile", [(16, 1), (32, 1), (32, 2), (32, 3), (64, 1), (64, 0)])
def test_uneven_product_paths_match_oracle(dim, tile):
    """Every product path the tuner can pick, forced through the tuning table."""
    import torch

    from magnus.uneven import UnevenBatch

    probs = [dict(n=4, dim=dim, samples=65, t0=0.0, tf=1.0, op="many", integrator="Simpson") for _ in range(3)]
    probs.append(dict(n=3, dim=dim, samples=33, t0=0.0, tf=1.0, op="sum", integrator="Simpson"))
    A, cot = _inputs(probs, 5)
    plan = UnevenBatch(probs, branches=2)
    plan.set_tuning(f"{dim} 65 4 {tile}\n{dim} 33 3 {tile}\n")
    assert all(line.split()[3] == str(tile) for line in plan.tuning.splitlines())
    outs = plan.forward(A)
    dAs = plan.vjp(A, cot)
    torch.cuda.synchronize()
    for i, p in enumerate(probs):
        a = A[i].cpu().numpy()
        ref = O.forward(p["n"], a, 0.0, 1.0, op=p["op"], integrator=p["integrator"])
        ref_g = O.vjp(p["n"], a, cot[i].cpu().numpy(), 0.0, 1.0, op=p["op"], integrator=p["integrator"])
        assert rel_fro(outs[i].cpu().numpy(), ref) <= TOL
        assert rel_fro(dAs[i].cpu().numpy(), ref_g) <= TOL
    plan.close()


def test_uneven_autotune_picks_measured_paths_and_keeps_parity():
    import torch

    from magnus.uneven import UnevenBatch

    probs = [dict(n=3, dim=d, samples=129, t0=0.0, tf=1.0, op="sum", integrator="Boole")
             for d in (4, 16, 32, 32, 64)]
    A, cot = _inputs(probs, 7)
    plan = UnevenBatch(probs, branches=2)
    plan.autotune(A, cot)
    lines = [ln.split() for ln in plan.tuning.splitlines() if not ln.startswith("pipeline")]
    pipe = [ln.split() for ln in plan.tuning.splitlines() if ln.startswith("pipeline")]
    bucketed = {p["dim"] for i, p in enumerate(probs) if not plan.in_pipeline(i)}
    assert len(lines) == len(bucketed) and all(float(ln[4]) > 0 for ln in lines)
    # the pipeline's segment length is chosen among 64, 128, 256 by its measured VJP
    assert len(pipe) == 1 and int(pipe[0][1]) in (64, 128, 256) and float(pipe[0][2]) > 0
    outs = plan.forward(A)
    dAs = plan.vjp(A, cot)
    torch.cuda.synchronize()
    for i, p in enumerate(probs):
        a = A[i].cpu().numpy()
        assert rel_fro(outs[i].cpu().numpy(), O.forward(3, a, 0.0, 1.0, integrator="Boole")) <= TOL
        ref_g = O.vjp(3, a, cot[i].cpu().numpy(), 0.0, 1.0, integrator="Boole")
        assert rel_fro(dAs[i].cpu().numpy(), ref_g) <= TOL
    # a second plan takes the table as text
    plan2 = UnevenBatch(probs, branches=1)
    plan2.set_tuning(plan.tuning)
    lines2 = [ln.split() for ln in plan2.tuning.splitlines()]
    assert [ln[3] for ln in lines2 if ln[0] != "pipeline"] == [ln[3] for ln in lines]
    assert [ln[1] for ln in lines2 if ln[0] == "pipeline"] == [pipe[0][1]]
    outs2 = plan2.forward(A)
    torch.cuda.synchronize()
    assert all(rel_fro(x.cpu().numpy(), y.cpu().numpy()) <= TOL for x, y in zip(outs2, outs))
    plan.close()
    plan2.close()


def test_uneven_argument_errors():
    import torch

    from magnus import capi
    from magnus.uneven import UnevenBatch

    with pytest.raises(Exception, match="instance 1"):
        UnevenBatch([dict(n=2, dim=4, samples=33), dict(n=2, dim=4, samples=32, integrator="Boole")])
    with pytest.raises(Exception):
        UnevenBatch([dict(n=6, dim=64, samples=8193)], workspace_bytes=1 << 20)
    plan = UnevenBatch([dict(n=2, dim=4, samples=33)])
    with pytest.raises(ValueError):
        plan.forward([torch.zeros((33, 4, 4), dtype=torch.float64, device="cuda")] * 2)
    plan.close()
    assert capi.lib().magnus_uneven_info(None, 0) == 0
