"""Benchmark of the MI355X Magnus path.  Prints ONE JSON line.

Default workload (one GPU): the learned resolution controller, the metric's
"Magnus fwd+VJP solves/s at rel err <= 1e-8, learned vs plain".  200k training
and 10k held-out Fourier-form instances (d = 8, 3-8 modes, ||A||_2 T ~ U[0.1, 4]),
labelled with the minimal (samples in {65..16385}, n in {2..8}) that reaches
relative error <= 1e-8 against a samples = 32769, n = 8 run; the MLP is trained
on the device.  One step = the learned forward + VJP of the 10k held-out
instances (features, MLP prediction, the doubling judge of every accepted
result, the VJP at the accepted resolution); the plain fixed resolution
(16385, 8) and the judge without the network are measured with the same
protocol.  Labelling and training happen before the timed steps.

Other workloads (--workload):
  spacecurve-batch  4096 trajectories f(t) = sum_{m=1..4} a_m sin(2 pi m t + phi_m),
                    S = 16385, Simpson, compute_sc n = 6 'many' + compute_sc_vjp; the
                    forward keeps its compact per-chunk checkpoint and the VJP starts
                    from it (--sc-residual recompute: the reference's VJP::none)
  dense-d32         256 x d = 32, S = 4097, Boole, n = 4, sum, record carry + VJP
  plain-d4          one d = 4 instance, S = 1025, Boole, n = 4, sum, forward only
  uneven            2048 instances of d 2..64, S 257..8193, n 2..6 in one graph launch
  wbasis-d8         one resolution group of the controller's solve: 1000 d = 8 instances,
                    S = 2049, Boole, n = 8, sum, forward keeping the basis state + VJP from
                    it, samples formed in the sweeps from the Fourier coefficients

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME] [--no-cpu]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "magnus-dp_amd"))

MI355X_HBM_GBS = 8000.0
MI355X_FP64_TFLOPS = 78.6


def _sc_inputs(torch, B, S, modes, device, seed=2):
    """Seeded Fourier-form trajectories a_c(t) = sum_k amp sin(2 pi k t + phase),
    sampled on the device by the library (magnus_sample_sc_fourier_f64)."""
    import numpy as np

    from magnus import torch_ops

    m = np.arange(1, modes + 1, dtype=np.float64)
    amp = np.empty((B, modes, 3))
    phi = np.empty((B, modes, 3))
    for b in range(B):
        rng = np.random.default_rng([seed, b])
        amp[b] = rng.standard_normal((modes, 3)) / m[:, None]
        phi[b] = rng.uniform(0.0, 2 * np.pi, size=(modes, 3))
    out = torch_ops.sample_sc_fourier(torch.from_numpy(amp).to(device), torch.from_numpy(phi).to(device), S)
    cot = np.stack([np.random.default_rng([seed, b, 1]).standard_normal((6, 3)) for b in range(B)])
    return out, torch.from_numpy(cot).to(device)


def _dense_inputs(torch, B, S, d, modes, device, seed):
    """Seeded Fourier-form matrices A(t) = A0 + sum_k B_k sin(2 pi k t) + C_k cos(2 pi k t),
    sampled on the device by the library (magnus_sample_fourier_f64)."""
    import numpy as np

    from magnus import torch_ops

    coeffs = np.stack([np.random.default_rng([seed, b]).standard_normal((2 * modes + 1, d, d)) / np.sqrt(d)
                       for b in range(B)])
    out = torch_ops.sample_fourier(torch.from_numpy(coeffs).to(device), S)
    cot = np.stack([np.random.default_rng([seed, b, 1]).standard_normal((d, d)) for b in range(B)])
    return out, torch.from_numpy(cot).to(device)


def _counts(Q, K, S, m, m_in, P, n_out, op_one=False):
    """Algorithmic flops and bytes per solve (FMA = 2 flops, 8 B per value)."""
    F_fwd = Q * K * S * (P + 8 * m) + (0 if op_one else 2 * m * S)
    F_vjp = Q * K * S * (2 * P + 10 * m) + Q * S * m + 2 * m * S
    B_fwd = 8 * (S * m_in + n_out * m_in)
    B_carry = 8 * Q * K * S * m
    B_vjp = 8 * (2 * S * m_in + n_out * m_in)
    return F_fwd, F_vjp, B_fwd, B_carry, B_vjp


KERNELS = ("sample_fourier", "sample_sc_fourier", "lm_fwd_prod", "lm_fwd_scan", "lm_fwd_fused", "lm_vjp_init", "lm_vjp_prod", "lm_vjp_scan", "lm_vjp_final",
           "wb_fwd_pass1", "wb_combine", "wb_fwd_pass1_vjp", "wb_combine_vjp", "wb_rev", "wb_fixup", "wbd8_sweep", "wbd8_rev",
           "walk_fwd", "walk_rev", "walk_reduce")

# su(2) products on the single-pass path: pure sample x 4-vector = 4 mul + 8 FMA
# (20 flops), 4-vector x 4-vector = 4 mul + 12 FMA (28 flops).
SC_AMUL_FLOPS = 20
SC_QMUL_FLOPS = 28


def _wb_counts(K, S):
    """Algorithmic flops / HBM bytes per solve of each single-pass kernel.

    Chunk pass: K+1 local prefix levels (a block prefix step is 16 flops per
    sample) and K sample products.  Reverse pass, per level and sample: the
    backward rebuild of the local stream (one product, 16 flops of block
    rule), the dA accumulation (one 4-vector product and an add), the
    adjoint product with the sample and the transposed block rule with the
    suffix sums (18 flops).  Bytes: the samples read (and dA written) once;
    the per-chunk summaries are a few hundred bytes per chunk and left out."""
    m_in = 3
    f1 = S * (K * SC_AMUL_FLOPS + (K + 1) * 16)
    fr = S * (K * (2 * SC_AMUL_FLOPS + SC_QMUL_FLOPS + 4 + 16 + 18) + 8)
    flops = {"wb_fwd_pass1": f1, "wb_fwd_pass1_vjp": f1, "wb_rev": fr}
    by = {"wb_fwd_pass1": 8 * S * m_in, "wb_fwd_pass1_vjp": 8 * S * m_in, "wb_rev": 8 * S * 2 * m_in}
    return flops, by


def _lm_counts(B, Q, K, S, m, m_in, P):
    """Algorithmic flops / HBM bytes per launch of the level-parallel kernels
    (one launch per level; jobs = B*Q): a product launch reads the samples and
    the previous prefix level and writes the products; a scan launch reads the
    products and writes the prefix level (the carry); a fused level (products
    formed inside the chunk passes) reads the samples and the previous level
    and writes the level."""
    jobs = B * Q
    flops = {"lm_fwd_prod": jobs * S * P, "lm_vjp_prod": jobs * S * 2 * P,
             "lm_fwd_scan": jobs * S * 8 * m, "lm_vjp_scan": jobs * S * 8 * m,
             "lm_fwd_fused": jobs * S * (P + 8 * m)}
    by = {"lm_fwd_prod": 8 * (B * S * m_in + 2 * jobs * S * m), "lm_fwd_scan": 8 * 2 * jobs * S * m,
          "lm_fwd_fused": 8 * (B * S * m_in + 2 * jobs * S * m),
          "lm_vjp_prod": 8 * (B * S * m_in + 5 * jobs * S * m), "lm_vjp_scan": 8 * 2 * jobs * S * m}
    return by, flops


def _measured_traffic(kernel, config):
    """HBM bytes per launch of `kernel` measured with rocprofv3 --pmc (FETCH_SIZE and
    WRITE_SIZE, each scaled by its reading for a plain copy of known size), from the
    summary committed under profiles/; used only while the kernel's source file is
    unchanged since the measurement (its SHA-256 is recorded) and the workload matches."""
    import hashlib

    for f in sorted((ROOT / "profiles").glob("*_traffic.json")):
        rec = json.loads(f.read_text())
        if rec.get("kernel") != kernel or rec.get("workload") != config.get("workload"):
            continue
        src = ROOT / rec["source"]
        if not src.exists() or hashlib.sha256(src.read_bytes()).hexdigest() != rec["source_sha256"]:
            return None
        return rec["traffic_bytes_per_launch"]
    return None


def _measured_traffic_ratio(kernel):
    """Measured HBM bytes per stated algorithmic byte of `kernel` (rocprofv3 FETCH_SIZE
    and WRITE_SIZE against the byte counts its launch site states), from a record under
    profiles/ taken on the current source of the kernel (SHA-256), or None."""
    import hashlib

    for f in sorted((ROOT / "profiles").glob("*_traffic.json")):
        rec = json.loads(f.read_text())
        if rec.get("kernel") != kernel or "traffic_per_stated_byte" not in rec:
            continue
        src = ROOT / rec["source"]
        if src.exists() and hashlib.sha256(src.read_bytes()).hexdigest() == rec["source_sha256"]:
            return rec["traffic_per_stated_byte"]
    return None


def _cpu_baseline_sc(a_host, cot_host, n, S, budget_s):
    """Oracle (C restatement of the reference CPU path) on host cores: the faster of
    (1) instances in sequence with the reference Auto GL-lane rule, VJP serial, and
    (2) one instance per core.  Time-bounded sample of the same trajectories."""
    import numpy as np

    sys.path.insert(0, str(ROOT / "oracle"))
    import oracle as O

    cores = len(os.sched_getaffinity(0))
    Q, K = (n + 3) // 2, n - 1
    results = {}
    for mode in ("reference-faithful", "all-cores"):
        done, t0 = 0, time.perf_counter()
        chunk = 1 if mode == "reference-faithful" else cores
        pool = a_host.shape[0] - a_host.shape[0] % chunk  # the sample is cycled until the budget is used
        while time.perf_counter() - t0 < budget_s / 2:
            lo = done % pool
            a = np.ascontiguousarray(a_host[lo:lo + chunk])
            c = np.ascontiguousarray(cot_host[lo:lo + chunk])
            carry = np.zeros((chunk, Q, K, S, 2, 2))
            if mode == "reference-faithful":
                lanes = O.auto_lanes(n, S, cores)
                O.sc_forward_batch(n, a, 0.0, 1.0, op="many", integrator="Simpson", lanes=lanes, threads=1,
                                   carry=carry)
                O.sc_vjp_batch(n, a, c, 0.0, 1.0, op="many", integrator="Simpson", threads=1, carry=carry)
            else:
                O.sc_forward_batch(n, a, 0.0, 1.0, op="many", integrator="Simpson", lanes=1, threads=cores,
                                   carry=carry)
                O.sc_vjp_batch(n, a, c, 0.0, 1.0, op="many", integrator="Simpson", threads=cores, carry=carry)
            done += chunk
        el = time.perf_counter() - t0
        results[mode] = (done / el if el > 0 else 0.0, done, el)
    best = max(results, key=lambda k: results[k][0])
    v, done, el = results[best]
    return {
        "value": v,
        "unit": "solves/s",
        "cores": cores if best == "all-cores" else min(cores, Q),
        "kind": "port",
        "sample": f"{done} forward+VJP solves (full S={S}) cycling through the batch's first "
                  f"{a_host.shape[0]} trajectories in {el:.1f} s, mode={best}; "
                  f"other mode {[(k, round(r[0], 2)) for k, r in results.items() if k != best]}",
    }


def _cpu_baseline_dense(a_host, cot_host, n, S, budget_s):
    """Oracle on host cores for the dense record + VJP step (forward recording the
    reference carry, VJP from it): the faster of (1) instances in sequence with the
    reference Auto GL-lane rule (VJP serial) and (2) one instance per core, each
    mode for about budget_s / 2 seconds on the batch's first instances."""
    import numpy as np

    sys.path.insert(0, str(ROOT / "oracle"))
    import oracle as O

    cores = len(os.sched_getaffinity(0))
    Q, K = (n + 3) // 2, n - 1
    d = a_host.shape[-1]
    results = {}
    for mode in ("reference-faithful", "all-cores"):
        done, t0 = 0, time.perf_counter()
        chunk = 1 if mode == "reference-faithful" else min(cores, a_host.shape[0])
        pool = a_host.shape[0] - a_host.shape[0] % chunk
        while time.perf_counter() - t0 < budget_s / 2:
            lo = done % pool
            a = np.ascontiguousarray(a_host[lo:lo + chunk])
            c = np.ascontiguousarray(cot_host[lo:lo + chunk])
            carry = np.zeros((chunk, Q, K, S, d, d))
            lanes, threads = (O.auto_lanes(n, S, cores), 1) if mode == "reference-faithful" else (1, cores)
            O.forward_batch(n, a, 0.0, 1.0, op="sum", integrator="Boole", lanes=lanes, threads=threads, carry=carry)
            O.vjp_batch(n, a, c, 0.0, 1.0, op="sum", integrator="Boole", threads=threads, carry=carry)
            done += chunk
        el = time.perf_counter() - t0
        results[mode] = (done / el if el > 0 else 0.0, done, el,
                         O.auto_lanes(n, S, cores) if mode == "reference-faithful" else cores)
    best = max(results, key=lambda k: results[k][0])
    v, done, el, used = results[best]
    return {
        "value": v,
        "unit": "solves/s",
        "cores": used,
        "kind": "port",
        "sample": f"{done} forward (recording the carry) + VJP (from it) solves at full size (S={S}, d={d}) cycling "
                  f"through the batch's first {a_host.shape[0]} instances in {el:.1f} s, mode={best}; other mode "
                  f"{[(k, round(r[0], 3), r[3]) for k, r in results.items() if k != best]} (mode, solves/s, threads); "
                  f"host cores {cores}",
    }


UNEVEN_DIMS = (2, 4, 8, 16, 32, 64)
UNEVEN_SAMPLES = (257, 1025, 4097, 8193)


def _uneven_problems(count, seed=4):
    """Uneven batch: d ~ U{2..64}, samples ~ U{257..8193}, n ~ U{2..6}, Boole, op='sum'."""
    import numpy as np

    rng = np.random.default_rng([seed, 0])
    return [dict(n=int(rng.integers(2, 7)), dim=int(rng.choice(UNEVEN_DIMS)), samples=int(rng.choice(UNEVEN_SAMPLES)),
                 t0=0.0, tf=1.0, op="sum", integrator="Boole") for _ in range(count)]


def _uneven_inputs(torch, probs, device, modes=3, seed=4):
    """Fourier-form A(t) (3 modes, entries N(0,1)/sqrt(d)) rescaled so that the
    largest ||A(t)||_2 on a 65-point grid times T is U[0.1, 3]; sampled on the
    device per (dim, samples) group; cotangents N(0,1) (d, d)."""
    import numpy as np

    from magnus import torch_ops

    tg = np.linspace(0.0, 1.0, 65)
    k = np.arange(1, modes + 1)
    basis = np.concatenate([np.ones((65, 1)), np.stack([np.sin(2 * np.pi * k[None, :] * tg[:, None]),
                                                       np.cos(2 * np.pi * k[None, :] * tg[:, None])],
                                                      axis=2).reshape(65, 2 * modes)], axis=1)
    coeffs, cots = [], []
    for i, p in enumerate(probs):
        rng = np.random.default_rng([seed, i])
        d = p["dim"]
        c = rng.standard_normal((2 * modes + 1, d, d)) / np.sqrt(d)
        grid = np.einsum("tk,kij->tij", basis, c)
        target = rng.uniform(0.1, 3.0)
        c *= target / (np.linalg.norm(grid, 2, axis=(1, 2)).max() * (p["tf"] - p["t0"]))
        coeffs.append(c)
        cots.append(np.random.default_rng([seed, i, 1]).standard_normal((d, d)))
    A = [None] * len(probs)
    groups = {}
    for i, p in enumerate(probs):
        groups.setdefault((p["dim"], p["samples"]), []).append(i)
    for (d, S), ids in groups.items():
        out = torch_ops.sample_fourier(torch.from_numpy(np.stack([coeffs[i] for i in ids])).to(device), S)
        for j, i in enumerate(ids):
            A[i] = out[j]
    return A, [torch.from_numpy(c).to(device) for c in cots]


def _uneven_flops(p, nodes=None):
    """Flops of forward (outputs) + VJP with recompute for one instance: with
    nodes=None the reference algorithm's count (Q GL nodes per level); with
    nodes=1 the dense W-basis path's (one basis sweep serves every node)."""
    n, S, d = p["n"], p["samples"], p["dim"]
    Q, K, m, P = (n + 3) // 2, n - 1, d * d, 2 * d ** 3
    Q = Q if nodes is None else nodes
    F_fwd, F_vjp, _, _, _ = _counts(Q, K, S, m, m, P, 1)
    F_rec = Q * K * S * (P + 6 * m)
    return F_fwd + F_rec + F_vjp


def _cpu_baseline_uneven(probs, A, cot, budget_s):
    """Oracle on host cores over a time-bounded sample of the uneven batch,
    extrapolated to the whole batch by the flop model; the faster of
    reference-faithful (instances in sequence, Auto GL lanes) and all-cores
    (one instance per core)."""
    import concurrent.futures as cf

    import numpy as np

    sys.path.insert(0, str(ROOT / "oracle"))
    import oracle as O

    cores = len(os.sched_getaffinity(0))
    cost = [_uneven_flops(p) for p in probs]
    total_cost = sum(cost)
    order = sorted(range(len(probs)), key=lambda i: cost[i])  # cheapest first, so the sample is bounded

    def solve(i, lanes):
        p = probs[i]
        a = A[i].cpu().numpy()
        c = cot[i].cpu().numpy()
        O.forward(p["n"], a, 0.0, 1.0, op="sum", integrator="Boole", lanes=lanes)
        O.vjp(p["n"], a, c, 0.0, 1.0, op="sum", integrator="Boole")

    results = {}
    for mode in ("reference-faithful", "all-cores"):
        budget = budget_s / 2
        stride = max(1, len(order) // 64)
        sample = order[::stride]
        t0 = time.perf_counter()
        done_cost, done = 0.0, 0
        if mode == "reference-faithful":
            for i in sample:
                if time.perf_counter() - t0 > budget:
                    break
                solve(i, O.auto_lanes(probs[i]["n"], probs[i]["samples"], cores))
                done_cost += cost[i]
                done += 1
        else:
            with cf.ThreadPoolExecutor(cores) as ex:
                for w in range(0, len(sample), cores):
                    if time.perf_counter() - t0 > budget:
                        break
                    part = sample[w:w + cores]
                    list(ex.map(lambda i: solve(i, 1), part))
                    done_cost += sum(cost[i] for i in part)
                    done += len(part)
        el = time.perf_counter() - t0
        est_total = el * total_cost / done_cost if done_cost else float("inf")
        results[mode] = (len(probs) / est_total, done, el)
    best = max(results, key=lambda k: results[k][0])
    v, done, el = results[best]
    return {
        "value": v,
        "unit": "solves/s",
        "cores": cores if best == "all-cores" else min(cores, 4),
        "kind": "port",
        "sample": f"{done} forward+VJP solves of the batch (every {max(1, len(order) // 64)}th instance in "
                  f"order of cost, cheapest first) in {el:.1f} s, extrapolated to the {len(probs)} instances "
                  f"by algorithmic flops; mode={best}; other mode "
                  f"{[(k, round(r[0], 3)) for k, r in results.items() if k != best]}",
    }


def _bench_uneven(args, torch, dev):
    from magnus import capi, torch_ops
    from magnus.uneven import UnevenBatch

    count = args.batch or 2048
    probs = _uneven_problems(count)
    print(f"[bench] uneven: generating {count} instances", file=sys.stderr, flush=True)
    A, cot = _uneven_inputs(torch, probs, dev)
    torch.cuda.synchronize()
    ws = int(args.uneven_workspace_gb * (1 << 30))
    plan = UnevenBatch(probs, workspace_bytes=ws, branches=args.uneven_branches)
    t = time.perf_counter()
    plan.autotune(A, cot)
    torch.cuda.synchronize()
    tune_s = time.perf_counter() - t
    print(f"[bench] uneven: autotune {tune_s:.1f} s, plan {plan.info()}", file=sys.stderr, flush=True)

    def step(graph):
        outs = plan.forward(A, use_graph=graph)
        dAs = plan.vjp(A, cot, use_graph=graph)
        return outs, dAs

    timing = {}
    for graph in (False, True):
        for _ in range(args.warmup):
            step(graph)
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(args.steps):
            outs, dAs = step(graph)
        torch.cuda.synchronize()
        timing[graph] = (time.perf_counter() - t) / args.steps * 1e3
        print(f"[bench] uneven: {'graph' if graph else 'eager'} {timing[graph]:.2f} ms/step", file=sys.stderr,
              flush=True)

    # per-instance sequential launches of the one-problem entry points
    torch.cuda.synchronize()
    t = time.perf_counter()
    seq_out, seq_dA = [], []
    for i, p in enumerate(probs):
        seq_out.append(torch_ops.forward(p["n"], A[i][None], 0.0, 1.0, op="sum", integrator="Boole")[0])
        dA_i = torch_ops.vjp(p["n"], A[i][None], cot[i][None], None, 0.0, 1.0, op="sum", integrator="Boole")[0]
        if i < 16:
            seq_dA.append((i, dA_i))
    torch.cuda.synchronize()
    seq_ms = (time.perf_counter() - t) * 1e3

    # bucketed instances equal their per-instance calls bit for bit, pipeline
    # instances to rounding (another summation order of the same recurrences)
    def same(i, ref, got, what):
        if plan.in_pipeline(i):
            err = float(torch.linalg.norm(got - ref) / max(float(torch.linalg.norm(ref)), 1e-300))
            assert err <= 1e-12, f"uneven {what} differs from the per-instance call at {i}: {err:.2e}"
        else:
            assert torch.equal(got, ref), f"uneven {what} differs from the per-instance call at {i}"

    for i, dA_i in seq_dA:
        same(i, dA_i, dAs[i], "dA")
    for i in range(count):
        same(i, seq_out[i], outs[i], "output")
    print(f"[bench] uneven: per-instance sequential {seq_ms:.1f} ms", file=sys.stderr, flush=True)

    # per-kernel totals of one eager step (the branches overlap, so these add
    # up to more than the step)
    capi.timing_reset()
    capi.timing_enable(True)
    step(False)
    torch.cuda.synchronize()
    capi.timing_enable(False)
    per_kernel = {}
    for name in KERNELS + ("wbd_fold", "wbd_fold_vjp", "uneven_gather", "uneven_scatter", "gw_fwd", "gw_rev"):
        cnt, tot = capi.timing_read(name)
        if cnt:
            per_kernel[name] = {"launches": cnt, "total_ms": round(tot, 3)}
    print(f"[bench] uneven: kernels {per_kernel}", file=sys.stderr, flush=True)

    ref_flops = sum(_uneven_flops(p) for p in probs)
    flops = sum(_uneven_flops(p, nodes=1) for p in probs)  # executed: W-basis sweep
    ms = timing[True]
    cpu = None if args.no_cpu else _cpu_baseline_uneven(probs, A, cot, args.cpu_budget)
    ach = flops / (ms / 1e3) / 1e12
    dims = {d: sum(1 for p in probs if p["dim"] == d) for d in UNEVEN_DIMS}
    line = {
        "metric": "Magnus fwd+VJP solves/s (uneven batch)",
        "value": round(count / (ms / 1e3), 3),
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded Fourier-form A(t), norm-rescaled, sampled on the device by the library)",
        "config": {"workload": "uneven-batch", "instances": count, "dims": dims, "samples": list(UNEVEN_SAMPLES),
                   "n": "2..6", "integrator": "Boole", "op": "sum", "vjp": "no carry (recompute)",
                   "branches": args.uneven_branches, "plan": plan.info(), "graph": True},
        "roofline": {"bound": "fp64", "achieved": round(ach, 3), "peak": MI355X_FP64_TFLOPS, "unit": "TFLOP/s",
                     "frac": round(ach / MI355X_FP64_TFLOPS, 4), "traffic": None, "kernel": "whole step",
                     "algorithmic_flops_per_step": flops,
                     "reference_algorithm_flops_per_step": ref_flops,
                     "reference_algorithm_rate_tflops": round(ref_flops / (ms / 1e3) / 1e12, 3)},
        "kernels_eager_step": per_kernel,
        "cpu_baseline": cpu,
        "per_instance_sequential": {"ms": round(seq_ms, 2), "solves_per_s": round(count / (seq_ms / 1e3), 3)},
        "eager_bucketed": {"ms": round(timing[False], 3)},
        "graph_vs_sequential": round(seq_ms / ms, 3),
        "autotune_s": round(tune_s, 2),
        "tuning": plan.tuning.strip().splitlines(),
    }
    plan.close()
    print(json.dumps(line))


def _cpu_baseline_controller(coeffs, cot, budget_s):
    """The reference CPU path at the plain resolution (S = 16385, n = 8,
    forward op='sum' + VJP) on the host cores, over a time-bounded sample of the
    evaluation set, in two modes: reference-faithful (instances in sequence,
    the reference's Auto GL-lane rule for the forward, serial VJP) and
    all-cores (one instance per core); the faster is reported."""
    import concurrent.futures as cf

    sys.path.insert(0, str(ROOT / "oracle"))
    import oracle as O
    from magnus import controller as C, torch_ops

    S, n = C.PLAIN
    cores = len(os.sched_getaffinity(0))
    lanes = O.auto_lanes(n, S, cores)
    # the sample's inputs are drawn on the device before the clock starts
    k = min(coeffs.shape[0], 4 * cores)
    A = torch_ops.sample_fourier(coeffs[:k], S).cpu().numpy()
    G = cot[:k].cpu().numpy()

    def one(i, ln):
        a = A[i % k]
        O.forward(n, a, 0.0, 1.0, op="sum", integrator="Boole", lanes=ln)
        O.vjp(n, a, G[i % k], 0.0, 1.0, op="sum", integrator="Boole")

    results = {}
    t0 = time.perf_counter()
    done = 0
    while time.perf_counter() - t0 < budget_s / 2 and done < coeffs.shape[0]:
        one(done, lanes)
        done += 1
    results["reference-faithful"] = (done / (time.perf_counter() - t0), done, time.perf_counter() - t0, lanes)
    t0 = time.perf_counter()
    done = 0
    with cf.ThreadPoolExecutor(cores) as ex:
        while time.perf_counter() - t0 < budget_s / 2 and done + cores <= coeffs.shape[0]:
            list(ex.map(lambda i: one(i, 1), range(done, done + cores)))
            done += cores
    results["all-cores"] = (done / (time.perf_counter() - t0), done, time.perf_counter() - t0, cores)
    best = max(results, key=lambda m: results[m][0])
    v, done, el, used = results[best]
    return {"value": v, "unit": "solves/s", "cores": used, "kind": "port",
            "sample": f"{done} plain-resolution (S={S}, n={n}) forward+VJP solves of the evaluation set "
                      f"(cycling its first {k} instances) in {el:.1f} s, mode={best}; other mode "
                      f"{[(m, round(r[0], 3), r[3]) for m, r in results.items() if m != best]} "
                      f"(mode, solves/s, threads); host cores {cores}"}


def _roctx():
    """ROCm's marker library, the one rocprofv3 --marker-trace records (ranges are
    no-ops without a profiler), or None where it is absent."""
    import ctypes

    lib_dir = os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "lib")
    for path in ("librocprofiler-sdk-roctx.so", os.path.join(lib_dir, "librocprofiler-sdk-roctx.so")):
        try:
            lib = ctypes.CDLL(path)
        except OSError:
            continue
        lib.roctxRangePushA.argtypes = [ctypes.c_char_p]
        return lib
    return None


def _bench_controller(args, torch, dev):
    import numpy as np

    from magnus import controller as C
    from magnus import torch_ops

    tol = 1e-8
    n_tr, n_ev = args.ctrl_train, args.ctrl_eval
    log = lambda msg: print(f"[bench] controller: {msg}", file=sys.stderr, flush=True)  # noqa: E731
    t = time.perf_counter()
    c_tr = C.make_coefficients(n_tr, seed=11)
    c_ev = C.make_coefficients(n_ev, seed=12)
    f_tr = C.features(torch_ops.sample_fourier(c_tr, C.FEATURE_GRID), 0.0, 1.0)
    def progress(done, total):
        if done % 10240 < 512 or done == total:
            log(f"labelled {done}/{total} ({time.perf_counter() - t:.0f} s)")

    lab_tr, min_tr, reach_tr, _, _ = C.label(c_tr, tol, progress=progress)
    lab_ev, min_ev, reach_ev, errs_ev, refs_ev = C.label(c_ev, tol, progress=progress)
    torch.cuda.synchronize()
    label_s = time.perf_counter() - t
    hist = np.bincount(lab_tr, minlength=len(C.CLASSES))
    top = [(C.CLASSES[k], int(hist[k])) for k in np.argsort(-hist)[:8]]
    log(f"{n_tr}+{n_ev} instances labelled in {label_s:.1f} s; reachable {reach_tr.mean():.3f}; top labels {top}")
    ctl = C.Controller(f_tr.shape[1], seed=1)
    (loss_hist), train_s = C.timed(lambda: ctl.fit(f_tr, torch.from_numpy(lab_tr).to(dev), epochs=args.ctrl_epochs,
                                                   batch=1024, lr=2e-3))
    log(f"trained {args.ctrl_epochs} epochs in {train_s:.1f} s; loss {loss_hist[0]:.3f} -> {loss_hist[-1]:.3f}")
    f_ev = C.features(torch_ops.sample_fourier(c_ev, C.FEATURE_GRID), 0.0, 1.0)
    pred = ctl.predict(f_ev).cpu().numpy()
    costs = np.array([C.cost(*c) for c in C.CLASSES])
    flat = errs_ev.reshape(n_ev, -1)
    sufficient = flat[np.arange(n_ev), pred] <= tol
    acc = {"exact": float((pred == lab_ev).mean()), "sufficient": float(sufficient.mean()),
           "cost_ratio_vs_label": float(costs[pred].sum() / costs[lab_ev].sum()),
           "label_cost_vs_minimal": float(costs[lab_ev].sum() / costs[min_ev].sum())}
    log(f"eval prediction {acc}")
    cot = torch.from_numpy(np.stack([np.random.default_rng([12, i, 1]).standard_normal((8, 8))
                                     for i in range(n_ev)])).to(dev)

    def plain():
        return C.solve(c_ev, [C.PLAIN] * n_ev, cot)

    stages = {}

    def learned():
        # the network's class starts the judge; each round runs the VJP of its
        # (S, n) sweep from the kept basis state, so the accepted round delivers
        # the forward and the VJP
        start, t1 = C.timed(lambda: ctl.predict(
            C.features(torch_ops.sample_fourier(c_ev, C.FEATURE_GRID), 0.0, 1.0)).cpu().numpy())
        (res, out, st), t2 = C.timed(lambda: C.accept(c_ev, start, tol, cot=cot))
        stages["learned"] = {"predict_s": round(t1, 3), "judge_and_vjp_s": round(t2, 3), "judge_calls": st["calls"]}
        return res, out, st, 0

    def nonet():
        # the judge from (65, 2) without the network, then the VJP at the delivered resolutions
        (res, fwd, st), t2 = C.timed(lambda: C.accept(c_ev, [(C.S_GRID[0], C.N_GRID[0])] * n_ev, tol))
        (out, _, w), t3 = C.timed(lambda: C.solve(c_ev, C.delivered(res), cot, forward_out=fwd))
        stages["no_network"] = {"judge_s": round(t2, 3), "solve_s": round(t3, 3), "judge_calls": st["calls"]}
        return res, out, st, w

    from magnus import capi

    kernel_ms = {}
    kernel_counts = {}
    roctx = _roctx()

    def measure(fn):
        """W untimed warm-up steps, then exactly K timed steps between device
        synchronisations, with the library's per-launch HIP events on; returns the
        last step's result and the seconds per step.  Device memory cached by the
        previous mode (torch's allocator, the library's scratch) is returned first,
        outside the timed region, so the modes do not stack their working sets."""
        torch.cuda.synchronize()
        torch.cuda.empty_cache()
        capi.lib().magnus_release_scratch()
        for _ in range(args.warmup):
            fn()
        torch.cuda.synchronize()
        capi.timing_reset()
        capi.timing_enable(True)
        if roctx:
            roctx.roctxRangePushA(f"{fn.__name__} timed steps".encode())
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            res = fn()
        torch.cuda.synchronize()
        per_step = (time.perf_counter() - t0) / args.steps
        if roctx:
            roctx.roctxRangePop()
        capi.timing_enable(False)
        tot, cnts = {}, {}
        for name in KERNELS + ("wbd_fold", "wbd_fold_vjp"):
            cnt, ms = capi.timing_read(name)
            if cnt:
                tot[name] = (cnt, round(ms, 1))
                cnts[name] = (cnt, ms) + capi.timing_counts(name)
        kernel_ms[fn.__name__] = tot
        kernel_counts[fn.__name__] = cnts
        log(f"{fn.__name__} kernels {tot}")
        return res, per_step

    (out_p, _, w_p), t_p = measure(plain)
    log(f"plain {t_p:.2f} s")
    (res_l, out_l, st_l, w_l), t_l = measure(learned)
    log(f"learned {t_l:.2f} s, rounds {np.bincount(st_l['rounds'])}")
    (res_0, out_0, st_0, w_0), t_0 = measure(nonet)
    log(f"without network {t_0:.2f} s, rounds mean {st_0['rounds'].mean():.2f}")
    # roofline: one more learned step, untimed, with the groups run serially on one
    # stream, so that every launch's HIP event pair brackets that launch alone; the
    # family with the most device time in it, at the algorithmic bytes / flops its
    # launch sites state over its summed durations.  Serial launches cannot
    # overlap, so the family's time is at most the pass's wall time.
    stages_timed = {k: dict(v) for k, v in stages.items()}
    C.set_streams(1)
    learned()  # the single stream's scratch is allocated outside the measured pass
    capi.timing_reset()
    capi.timing_enable(True)
    torch.cuda.synchronize()
    if roctx:
        roctx.roctxRangePushA(b"learned serial pass")
    t_pass = time.perf_counter()
    learned()
    torch.cuda.synchronize()
    pass_s = time.perf_counter() - t_pass
    if roctx:
        roctx.roctxRangePop()
    capi.timing_enable(False)
    C.set_streams(4)
    stages.clear()
    stages.update(stages_timed)
    serial = {}
    for name in KERNELS + ("wbd_fold", "wbd_fold_vjp"):
        cnt, ms = capi.timing_read(name)
        if cnt:
            serial[name] = (cnt, ms) + capi.timing_counts(name)
    log(f"serial learned pass {pass_s * 1e3:.1f} ms: {({k: (v[0], round(v[1], 2)) for k, v in serial.items()})}")
    roofline = None
    counted = {k: v for k, v in serial.items() if v[2] > 0 or v[3] > 0}
    if counted:
        dom = max(counted, key=lambda k: counted[k][1])
        n_l, ms_l, by, fl = counted[dom]
        dur_s = ms_l / 1e3
        if by / (MI355X_HBM_GBS * 1e9) >= fl / (MI355X_FP64_TFLOPS * 1e12):
            ach = by / dur_s / 1e9
            roofline = {"bound": "hbm", "achieved": round(ach, 2), "peak": MI355X_HBM_GBS, "unit": "GB/s",
                        "frac": round(ach / MI355X_HBM_GBS, 4)}
        else:
            ach = fl / dur_s / 1e12
            roofline = {"bound": "fp64", "achieved": round(ach, 3), "peak": MI355X_FP64_TFLOPS, "unit": "TFLOP/s",
                        "frac": round(ach / MI355X_FP64_TFLOPS, 4)}
        ratio = _measured_traffic_ratio(dom)
        # the driver-clock figure: the family's stated work in the timed (overlapped)
        # learned steps over their wall time
        _, _, by_t, fl_t = kernel_counts["learned"].get(dom, (0, 0.0, 0.0, 0.0))
        roofline.update({
            "traffic": round(ratio * by / n_l) if ratio else None,
            "kernel": dom, "launches": n_l, "kernel_avg_ms": round(ms_l / n_l, 4),
            "kernel_ms_serial_pass": round(ms_l, 3), "serial_pass_ms": round(pass_s * 1e3, 3),
            "algorithmic_bytes_per_launch": round(by / n_l), "algorithmic_flops_per_launch": round(fl / n_l),
            "timed_steps_rate": ({"TFLOP/s": round(fl_t / args.steps / t_l / 1e12, 3)} if roofline["unit"] == "TFLOP/s"
                                 else {"GB/s": round(by_t / args.steps / t_l / 1e9, 2)}),
            "note": "serial pass over the learned step's groups (one stream): summed stated work over summed "
                    "durations; launches differ in size (one per resolution group); traffic = HBM bytes measured "
                    "per stated byte on the same kernel (profiles/*_traffic.json) times the stated bytes per launch; "
                    "timed_steps_rate = the family's stated work per timed step over the step's wall time"})
    hit = lambda o: float((C.rel_err(o, refs_ev) <= tol).float().mean())  # noqa: E731
    cpu = None if args.no_cpu else _cpu_baseline_controller(c_ev, cot, args.cpu_budget)
    value = n_ev / t_l
    line = {
        "metric": "Magnus fwd+VJP solves/s at rel err <=1e-8, learned vs plain, 1 GPU",
        "value": round(value, 3),
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(t_l * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded Fourier-form A(t), d=8, 3-8 modes, ||A||_2 T ~ U[0.1, 4]; samples formed on the "
                "device from the coefficients)",
        "config": {"workload": "controller", "train": n_tr, "eval": n_ev, "dim": 8, "tol": tol,
                   "samples": list(C.S_GRID), "orders": list(C.N_GRID), "reference": [C.REF_S, C.REF_N],
                   "plain": list(C.PLAIN), "mlp": "24-128-128-63, bf16 weights, fp32 accumulation, Adam"},
        "learned_vs_plain": round(t_p / t_l, 3),
        "learned_vs_no_network": round(t_0 / t_l, 3),
        "plain": {"s": round(t_p, 3), "solves_per_s": round(n_ev / t_p, 3), "hit_rate": hit(out_p)},
        "learned": {"s": round(t_l, 3), "hit_rate": hit(out_l), "judge_work": int(st_l["work"]),
                    "solve_work": int(w_l), "mean_rounds": float(st_l["rounds"].mean())},
        "no_network": {"s": round(t_0, 3), "solves_per_s": round(n_ev / t_0, 3), "hit_rate": hit(out_0),
                       "judge_work": int(st_0["work"]), "solve_work": int(w_0),
                       "mean_rounds": float(st_0["rounds"].mean())},
        "prediction": acc,
        "kernel_ms": kernel_ms,
        "stages": stages,
        "label_s": round(label_s, 2),
        "train_s": round(train_s, 2),
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None, help="timed steps (default 3 for controller, else 20)")
    ap.add_argument("--warmup", type=int, default=None, help="warm-up steps (default 1 for controller, else 5)")
    ap.add_argument("--workload", default="controller",
                    choices=["spacecurve-batch", "dense-d32", "plain-d4", "uneven", "controller", "wbasis-d8"])
    ap.add_argument("--ctrl-train", type=int, default=200000)
    ap.add_argument("--ctrl-eval", type=int, default=10000)
    ap.add_argument("--ctrl-epochs", type=int, default=20)
    ap.add_argument("--uneven-branches", type=int, default=4)
    ap.add_argument("--uneven-workspace-gb", type=float, default=128.0)
    ap.add_argument("--batch", type=int, default=None, help="override the batch size (testing only)")
    ap.add_argument("--dim", type=int, default=32, help="dense workload: matrix dimension (profiling only)")
    ap.add_argument("--sc-residual", default="checkpoint", choices=["checkpoint", "recompute"],
                    help="spacecurve-batch: what the VJP starts from")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=20.0)
    ap.add_argument("--copy-mib", type=int, default=0,
                    help="only run a plain device copy (y = 1.0 * x) of this many MiB, three times: the "
                         "reference reading for scaling the FETCH_SIZE / WRITE_SIZE counters")
    args = ap.parse_args()
    if args.steps is None:
        args.steps = 3 if args.workload == "controller" else 20
    if args.warmup is None:
        args.warmup = 1 if args.workload == "controller" else 5
    if args.gpus != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path (run with --gpus 1)",
              file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    from magnus import capi, torch_ops

    dev = torch.device("cuda:0")
    if args.copy_mib:
        x = torch.ones(args.copy_mib * 131072, dtype=torch.float64, device=dev)
        y = torch.empty_like(x)
        for _ in range(3):
            torch.mul(x, 1.0, out=y)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": x.numel() * 8, "copy_bytes_written": y.numel() * 8}))
        return
    wl = args.workload
    if wl == "uneven":
        _bench_uneven(args, torch, dev)
        return
    if wl == "controller":
        _bench_controller(args, torch, dev)
        return
    if wl == "spacecurve-batch":
        B, S, n, modes = args.batch or 4096, 16385, 6, 4
        a, cot = _sc_inputs(torch, B, S, modes, dev)

        if args.sc_residual == "checkpoint":
            def step():
                # compute_sc (op='many') keeping the per-chunk checkpoint, then
                # compute_sc_vjp from it
                out, ckpt = torch_ops.sc_forward_ckpt(n, a, 0.0, 1.0, op="many", integrator="Simpson")
                dA = torch_ops.sc_vjp_ckpt(n, a, cot, ckpt, 0.0, 1.0, op="many", integrator="Simpson")
                return out, dA
        else:
            def step():
                # compute_sc (op='many') then compute_sc_vjp without a stored carry: the
                # VJP re-derives its forward quantities (the reference's VJP::none path)
                out = torch_ops.sc_forward(n, a, 0.0, 1.0, op="many", integrator="Simpson")
                dA = torch_ops.sc_vjp(n, a, cot, None, 0.0, 1.0, op="many", integrator="Simpson")
                return out, dA

        Q, K, m, m_in, P, n_out = (n + 3) // 2, n - 1, 4, 3, 20, n
        config = {"workload": "spacecurve-batch", "instances": B, "samples": S, "n": n, "integrator": "Simpson",
                  "op": "many", "vector": "su(2) 3-vector", "fourier_modes": modes, "cotangent": "N(0,1) (6,3)",
                  "carry": ("none; compact per-chunk checkpoint kept by the forward" if args.sc_residual == "checkpoint"
                            else "none (VJP recomputes its forward quantities)")}
        do_vjp = True
    elif wl == "dense-d32":
        B, S, d, n, modes = args.batch or 256, 4097, args.dim, 4, 5
        a, cot = _dense_inputs(torch, B, S, d, modes, dev, seed=3)

        def step():
            out, carry = torch_ops.forward(n, a, 0.0, 1.0, op="sum", integrator="Boole", record_vjp=True)
            dA = torch_ops.vjp(n, a, cot, carry, 0.0, 1.0, op="sum", integrator="Boole")
            return out, dA

        Q, K, m, m_in, P, n_out = (n + 3) // 2, n - 1, d * d, d * d, 2 * d ** 3, 1
        config = {"workload": "dense-d32" if d == 32 else f"dense-d{d}", "instances": B, "samples": S, "dim": d,
                  "n": n, "integrator": "Boole",
                  "op": "sum", "carry": "recorded"}
        do_vjp = True
    elif wl == "wbasis-d8":
        # one resolution group of the controller's solve: forward (op 'sum')
        # keeping the basis state + VJP from it on the d = 8 single-pass sweeps,
        # 1000 d = 8 instances at S = 2049 given by their Fourier coefficients
        # (the sweeps form the samples, as the controller runs them)
        import numpy as np

        B, S, d, n, modes = args.batch or 1000, 2049, 8, 8, 5
        coeffs = torch.from_numpy(np.stack([np.random.default_rng([5, b]).standard_normal((2 * modes + 1, d, d))
                                            / np.sqrt(d) for b in range(B)])).to(dev)
        cot = torch.from_numpy(np.stack([np.random.default_rng([5, b, 1]).standard_normal((d, d))
                                         for b in range(B)])).to(dev)

        def step():
            out, state = torch_ops.fourier_forward_basis(n, coeffs, S, 0.0, 1.0, op="sum", integrator="Boole")
            dA = torch_ops.fourier_vjp_basis(n, coeffs, S, cot, state, 0.0, 1.0, op="sum", integrator="Boole")
            return out, dA

        # one job per instance on this path (the W-basis holds all GL nodes)
        Q, K, m, m_in, P, n_out = 1, n - 1, d * d, d * d, 2 * d ** 3, 1
        config = {"workload": "wbasis-d8", "instances": B, "samples": S, "dim": d, "n": n, "integrator": "Boole",
                  "op": "sum", "carry": "none (basis state kept by the forward)", "input": "Fourier coefficients"}
        do_vjp = True
    else:
        B, S, d, n, modes = 1, 1025, 4, 4, 3
        a, cot = _dense_inputs(torch, B, S, d, modes, dev, seed=1)

        def step():
            return torch_ops.forward(n, a, 0.0, 1.0, op="sum", matrix_backend="Manual", integrator="Boole"), None

        Q, K, m, m_in, P, n_out = (n + 3) // 2, n - 1, d * d, d * d, 2 * d ** 3, 1
        config = {"workload": "plain-d4", "instances": 1, "samples": S, "dim": d, "n": n, "integrator": "Boole",
                  "op": "sum", "backend": "Manual"}
        do_vjp = False

    print(f"[bench] workload={wl} B={B} S={S} n={n}: warmup {args.warmup}", file=sys.stderr, flush=True)
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    capi.timing_reset()
    capi.timing_enable(True)
    start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    start.record()
    for i in range(args.steps):
        step()
    stop.record()
    torch.cuda.synchronize()
    capi.timing_enable(False)
    ms_total = start.elapsed_time(stop)
    ms_step = ms_total / args.steps
    per_kernel = {}
    for name in KERNELS:
        cnt, tot = capi.timing_read(name)
        if cnt:
            per_kernel[name] = {"launches": cnt, "avg_ms": tot / cnt}
    print(f"[bench] timed {args.steps} steps: {ms_step:.3f} ms/step; kernels {per_kernel}", file=sys.stderr,
          flush=True)

    F_fwd, F_vjp, B_fwd, B_carry, B_vjp = _counts(Q, K, S, m, m_in, P, n_out)
    kernel_bytes, kernel_flops = _lm_counts(B, Q, K, S, m, m_in, P)
    wbf, wbb = _wb_counts(K, S)
    for k in wbf:
        kernel_flops[k] = B * wbf[k]
        kernel_bytes[k] = B * wbb[k]
    for k in per_kernel:
        if k not in kernel_bytes:  # families whose launch sites state their own work
            by, fl = capi.timing_counts(k)
            kernel_bytes[k] = by / per_kernel[k]["launches"]
            kernel_flops[k] = fl / per_kernel[k]["launches"]
    dom = max(per_kernel, key=lambda k: per_kernel[k]["avg_ms"]) if per_kernel else None
    roofline = None
    if dom in kernel_bytes:
        dur_s = per_kernel[dom]["avg_ms"] / 1e3
        bytes_ach = kernel_bytes[dom] / dur_s / 1e9
        flops_ach = kernel_flops[dom] / dur_s / 1e12
        t_bytes = kernel_bytes[dom] / (MI355X_HBM_GBS * 1e9)
        t_flops = kernel_flops[dom] / (MI355X_FP64_TFLOPS * 1e12)
        if t_bytes >= t_flops:
            roofline = {"bound": "hbm", "achieved": round(bytes_ach, 2), "peak": MI355X_HBM_GBS, "unit": "GB/s",
                        "frac": round(bytes_ach / MI355X_HBM_GBS, 4), "traffic": None}
        else:
            roofline = {"bound": "fp64", "achieved": round(flops_ach, 3), "peak": MI355X_FP64_TFLOPS,
                        "unit": "TFLOP/s", "frac": round(flops_ach / MI355X_FP64_TFLOPS, 4), "traffic": None}
        roofline["traffic"] = _measured_traffic(dom, config)
        if roofline["traffic"] is None:
            ratio = _measured_traffic_ratio(dom)
            roofline["traffic"] = round(ratio * kernel_bytes[dom]) if ratio else None
        roofline["kernel"] = dom
        roofline["kernel_avg_ms"] = round(per_kernel[dom]["avg_ms"], 4)
        roofline["algorithmic_bytes_per_launch"] = kernel_bytes[dom]
        roofline["algorithmic_flops_per_launch"] = kernel_flops[dom]

    cpu = None
    if not args.no_cpu and wl == "spacecurve-batch":
        k = min(B, 1024)
        cpu = _cpu_baseline_sc(a[:k].cpu().numpy(), cot[:k].cpu().numpy(), n, S, args.cpu_budget)
    elif not args.no_cpu and wl == "dense-d32":
        k = min(B, 32)
        cpu = _cpu_baseline_dense(a[:k].cpu().numpy(), cot[:k].cpu().numpy(), n, S, args.cpu_budget)

    value = B / (ms_step / 1e3)
    line = {
        "metric": "Magnus fwd+VJP solves/s" if do_vjp else "Magnus fwd solves/s",
        "value": round(value, 3),
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded Fourier-form A(t), sampled on the device by the library)",
        "config": config,
        "roofline": roofline,
        "cpu_baseline": cpu,
        "kernels": per_kernel,
        "kernel_family": capi.last_kernel(),
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
