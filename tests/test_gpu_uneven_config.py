"""GPU: the uneven batch at its configuration's own shapes.

Instances drawn by the benchmark's generator (d in {2, 4, 8, 16, 32, 64},
samples in {257, 1025, 4097, 8193}, n in 2..6, Boole, op 'sum', Fourier-form
A(t) rescaled so that ||A||_2 T ~ U[0.1, 3]), every (d, samples) bucket
present, d = 64 at 8193 samples with n = 6 included.  The plan runs in graph
mode with a workspace small enough that the bucketed d = 64 / 8193-sample
instances are cut into one-instance sub-batches and the grouped walker
pipeline (d = 16, 32) into waves; three replays are checked against
per-instance calls of the one-problem entry points (bucketed instances bit for
bit, pipeline instances at 1e-12) and against each other (bit for bit), and
one instance per (d, samples) pair (its smallest order) against the oracle at
1e-12, both the per-instance call and the batch.  The reference
solves one problem per call (src/composer/backends.hpp:108-122), so the
per-instance calls are the drop-in's own.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import sys

import numpy as np
import pytest

import oracle as O
from helpers import rel_fro

pytestmark = pytest.mark.gpu

TOL = 1e-12
COUNT = 256


def _say(msg):
    print(f"[uneven-config] {msg}", file=sys.stderr, flush=True)


@pytest.fixture(scope="module")
def batch():
    import torch

    import bench
    from magnus import torch_ops
    from magnus.uneven import UnevenBatch

    probs = bench._uneven_problems(2048)[:COUNT]
    present = {(p["dim"], p["samples"]) for p in probs}
    for d in bench.UNEVEN_DIMS:
        for S in bench.UNEVEN_SAMPLES:
            if (d, S) not in present:
                probs.append(dict(n=2, dim=d, samples=S, t0=0.0, tf=1.0, op="sum", integrator="Boole"))
    for _ in range(2):  # the largest instances, at the largest order
        probs.append(dict(n=6, dim=64, samples=8193, t0=0.0, tf=1.0, op="sum", integrator="Boole"))
    for d in (32, 64):  # an n = 2 instance of every large bucket keeps the oracle's share short
        for S in bench.UNEVEN_SAMPLES:
            probs.append(dict(n=2, dim=d, samples=S, t0=0.0, tf=1.0, op="sum", integrator="Boole"))
    _say(f"{len(probs)} instances, generating inputs")
    A, cot = bench._uneven_inputs(torch, probs, torch.device("cuda:0"))
    torch.cuda.synchronize()
    _say("per-instance calls")
    ref_out, ref_dA = [], []
    for p, a, c in zip(probs, A, cot):
        ref_out.append(torch_ops.forward(p["n"], a[None], 0.0, 1.0, op="sum", integrator="Boole")[0])
        ref_dA.append(torch_ops.vjp(p["n"], a[None], c[None], None, 0.0, 1.0, op="sum", integrator="Boole")[0])
    torch.cuda.synchronize()
    branches = 4
    plan = UnevenBatch(probs, workspace_bytes=branches * (3 << 30), branches=branches)
    _say(f"plan {plan.info()}")
    yield probs, A, cot, ref_out, ref_dA, plan
    plan.close()


def test_config_shapes_graph_replays_match_per_instance_calls(batch):
    import torch

    probs, A, cot, ref_out, ref_dA, plan = batch
    info = plan.info()
    bucketed = {(p["dim"], p["samples"], p["n"]) for i, p in enumerate(probs) if not plan.in_pipeline(i)}
    assert info["buckets"] == len(bucketed)
    assert info["pipeline_instances"] == sum(1 for i in range(len(probs)) if plan.in_pipeline(i)) > 0
    big = [i for i, p in enumerate(probs) if p["dim"] == 64 and p["samples"] == 8193 and p["n"] == 6]
    assert len(big) >= 2
    if not plan.in_pipeline(big[0]):
        assert info["sub_batches"] >= info["buckets"] + len(big) - 1  # one instance per sub-batch there
    first = None
    for rep in range(3):
        outs = plan.forward(A, use_graph=True)
        dAs = plan.vjp(A, cot, use_graph=True)
        torch.cuda.synchronize()
        for i, p in enumerate(probs):
            if plan.in_pipeline(i):
                assert rel_fro(outs[i].cpu().numpy(), ref_out[i].cpu().numpy()) <= TOL, (rep, i, p)
                assert rel_fro(dAs[i].cpu().numpy(), ref_dA[i].cpu().numpy()) <= TOL, (rep, i, p)
            else:
                assert torch.equal(outs[i], ref_out[i]), (rep, i, p)
                assert torch.equal(dAs[i], ref_dA[i]), (rep, i, p)
        if first is None:
            first = ([x.clone() for x in outs], [x.clone() for x in dAs])
        else:
            assert all(torch.equal(x, y) for x, y in zip(outs, first[0])), rep
            assert all(torch.equal(x, y) for x, y in zip(dAs, first[1])), rep
        _ = sum(float(x.abs().sum()) for x in outs)  # the program's own eager work between replays
        _say(f"replay {rep} checked")


def test_config_shapes_one_instance_per_bucket_matches_oracle(batch):
    import torch

    probs, A, cot, ref_out, ref_dA, plan = batch
    total = 64 * 65 // 2
    w, x = np.empty(total), np.empty(total)
    off = 0
    for order in range(1, 65):
        xr, wr = np.polynomial.legendre.leggauss(order)
        x[off:off + order] = (xr + 1) / 2
        w[off:off + order] = wr / 2
        off += order
    O.gl_replace(64, w, x)
    pick = {}
    for i, p in enumerate(probs):
        key = (p["dim"], p["samples"])
        if key not in pick or p["n"] < probs[pick[key]]["n"]:
            pick[key] = i
    assert len(pick) == 24
    outs = plan.forward(A, use_graph=False)
    dAs = plan.vjp(A, cot, use_graph=False)
    torch.cuda.synchronize()
    cases = [(i, A[i].cpu().numpy(), cot[i].cpu().numpy(), ref_out[i].cpu().numpy(), ref_dA[i].cpu().numpy(),
              outs[i].cpu().numpy(), dAs[i].cpu().numpy()) for i in pick.values()]

    def check(case):
        i, a, c, out, dA, b_out, b_dA = case
        n = probs[i]["n"]
        ref = O.forward(n, a, 0.0, 1.0, op="sum", integrator="Boole")
        ref_g = O.vjp(n, a, c, 0.0, 1.0, op="sum", integrator="Boole")
        return i, max(rel_fro(out, ref), rel_fro(b_out, ref)), max(rel_fro(dA, ref_g), rel_fro(b_dA, ref_g))

    try:
        workers = min(len(cases), max(1, len(os.sched_getaffinity(0))))
        with cf.ThreadPoolExecutor(workers) as ex:
            for i, e_out, e_dA in ex.map(check, sorted(cases, key=lambda c: -c[1].size)):
                p = probs[i]
                _say(f"d={p['dim']} S={p['samples']} n={p['n']}: out {e_out:.2e}, dA {e_dA:.2e}")
                assert e_out <= TOL and e_dA <= TOL, (i, p, e_out, e_dA)
    finally:
        O.gl_reset()
