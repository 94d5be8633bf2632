"""Learned resolution controller: pick (samples, order) per instance, verify, escalate.

Problem.  For an instance A(t) (dim d, Fourier form) and a tolerance tol, a
resolution whose Magnus sum is within tol (relative Frobenius) of the
S = 32769, n = 8 sum, as cheaply as possible.  A class (S, n) -- S in
{65, ..., 16385} (2^k + 1), n in {2..8} -- is delivered at (S, n) and judged
by the doubling pair (S, 2S - 1).

Pieces, all on the device:
  * labels: one all-orders sweep (magnus_forward_orders_f64) per S gives the
    sums of every order at once, so an instance is labelled with 10 sweeps;
  * features: magnus_ctrl_features_f64 on a 65-point coarse sampling;
  * network: magnus_mlp_* (3 layers, width 128, bfloat16 weights, FP32
    accumulation, Adam), classifying the 63 (S, n) classes;
  * acceptance (the judge, also the fallback): class (S, n) is accepted when
    the sums at (S, n) and (2S - 1, n) agree within JUDGE_FINE * tol and, for
    n < 8, the sums at (2S - 1, n) and (2S - 1, n + 1) within JUDGE_MARGIN *
    tol.  The pair's difference estimates the delivered (S, n) sum's
    discretisation error (the finer member's error is a small fraction of the
    coarser's: at least fourfold smaller per doubling, 64-fold for Boole in
    its asymptotic range), and the next order's term estimates its truncation
    error, so the two bounds add up to under tol.  A failed check doubles S
    and/or raises n, and repeats.  The learned path starts the judge at the
    network's class; a round keeps the basis state of its (S, n) sweep
    (torch_ops.forward_basis) and runs the VJP from it (torch_ops.vjp_basis),
    and checks with a sweep at 2S - 1 that keeps no levels (the d = 8 sweeps
    form their samples from the Fourier coefficients, torch_ops.fourier_*, so
    no sample array is written or read), so an accepted
    instance costs one sweep keeping its levels and one reverse sweep at the
    delivered resolution plus one level-free sweep at twice the samples; the
    same judge started at (65, 2) is the method without the network (judge
    rounds, then the VJP at the accepted resolution); the plain method runs
    every instance at (16385, 8), forward with its kept basis state and the
    VJP from it.
"""
from __future__ import annotations

import ctypes
import math
import time

import numpy as np
import torch

from . import capi, torch_ops

S_GRID = (65, 129, 257, 513, 1025, 2049, 4097, 8193, 16385)
N_GRID = (2, 3, 4, 5, 6, 7, 8)
REF_S, REF_N = 32769, 8
PLAIN = (16385, 8)
CLASSES = tuple((S, n) for S in S_GRID for n in N_GRID)
FEATURE_GRID = 65
FOURIER_MAX_MODES = 8  # the d = 8 sweeps form samples from at most this many modes
MAX_MODES = 8
JUDGE_MARGIN = 0.25  # next-order check: |sum(2S - 1, n + 1) - sum(2S - 1, n)| <= JUDGE_MARGIN tol
JUDGE_FINE = 0.75    # doubling check: |sum(2S - 1, n) - sum(S, n)| <= JUDGE_FINE tol (sum(S, n) is delivered)


def cost(S: int, n: int) -> int:
    return (S - 1) * (n - 1)


def class_index(S: int, n: int) -> int:
    return S_GRID.index(S) * len(N_GRID) + N_GRID.index(n)


def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


# --------------------------------------------------------------------------- data
def make_coefficients(count: int, seed: int, dim: int = 8, modes=(3, 8), norm=(0.1, 4.0), device="cuda"):
    """Fourier-form coefficients [count][2*MAX_MODES+1][dim][dim] (zero beyond
    each instance's mode count M ~ U{modes}), entries N(0,1)/sqrt(dim), scaled
    so that max over the 65-point grid of ||A(t)||_2 * T is U[norm]."""
    coeffs = np.zeros((count, 2 * MAX_MODES + 1, dim, dim))
    target = np.empty(count)
    for i in range(count):
        rng = np.random.default_rng([seed, i])
        M = int(rng.integers(modes[0], modes[1] + 1))
        coeffs[i, : 2 * M + 1] = rng.standard_normal((2 * M + 1, dim, dim)) / math.sqrt(dim)
        target[i] = rng.uniform(*norm)
    c = torch.from_numpy(coeffs).to(device)
    f = features(torch_ops.sample_fourier(c, FEATURE_GRID), 0.0, 1.0)
    scale = torch.from_numpy(target).to(device) / torch.exp(f[:, 0].double())
    return c * scale[:, None, None, None]


def features(A_coarse: torch.Tensor, t0: float, tf: float) -> torch.Tensor:
    """[B][F] float32 features of A sampled on a 2^k + 1 grid ([B][G][d][d])."""
    A_coarse = A_coarse.contiguous()
    B, G, d, _ = A_coarse.shape
    out = torch.empty((B, int(capi.lib().magnus_ctrl_feature_count())), dtype=torch.float32, device=A_coarse.device)
    capi.check(capi.lib().magnus_ctrl_features_f64(B, G, d, t0, tf, ctypes.c_void_p(A_coarse.data_ptr()),
                                                   ctypes.c_void_p(out.data_ptr()), _stream()))
    return out


def _fourier(coeffs: torch.Tensor, n: int, integrator: str = "Boole") -> bool:
    """The d = 8 single-pass sweeps form the samples from the coefficients
    themselves (Boole, n <= 8, at most 8 modes): the same results as sampling
    first, bit for bit, without a sample array."""
    return (coeffs.shape[-1] == 8 and integrator == "Boole" and 2 <= n <= 8
            and (coeffs.shape[1] - 1) // 2 <= FOURIER_MAX_MODES)


def sums(coeffs: torch.Tensor, S: int, n: int, integrator: str = "Boole", first: int = 1) -> torch.Tensor:
    """Magnus sums of orders first..n at S samples: [B][n - first + 1][d][d]."""
    if _fourier(coeffs, n, integrator):
        return torch_ops.fourier_forward_orders(n, coeffs, S, 0.0, 1.0, integrator=integrator, first=first)
    return torch_ops.forward_orders(n, torch_ops.sample_fourier(coeffs, S), 0.0, 1.0, integrator=integrator,
                                    first=first)


def _forward_basis(c: torch.Tensor, S: int, n: int):
    if _fourier(c, n):
        return torch_ops.fourier_forward_basis(n, c, S, 0.0, 1.0, op="sum", integrator="Boole")
    return torch_ops.forward_basis(n, torch_ops.sample_fourier(c, S), 0.0, 1.0, op="sum", integrator="Boole")


def _vjp_basis(c: torch.Tensor, S: int, n: int, cot: torch.Tensor, state: torch.Tensor):
    if _fourier(c, n):
        return torch_ops.fourier_vjp_basis(n, c, S, cot, state, 0.0, 1.0, op="sum", integrator="Boole")
    return torch_ops.vjp_basis(n, torch_ops.sample_fourier(c, S), cot, state, 0.0, 1.0, op="sum", integrator="Boole")


def rel_err(x: torch.Tensor, ref: torch.Tensor) -> torch.Tensor:
    return torch.linalg.vector_norm((x - ref).flatten(1), dim=1) / torch.linalg.vector_norm(ref.flatten(1), dim=1)


def label(coeffs: torch.Tensor, tol: float, batch: int = 2048, progress=None):
    """Per instance: (labels, minimal, reachable, errors [B][9][7], reference
    sums [B][d][d]); progress(done, total) is called after every batch.
    errors[s][k] is the error of class (S_GRID[s], N_GRID[k])'s delivered sum
    (S samples, order n); `minimal` is the cheapest class whose delivered sum is
    within tol; `labels` -- what the network learns -- is the cheapest class
    the judge accepts in one round with its delivered sum within tol, so a
    correct prediction costs one round."""
    B = coeffs.shape[0]
    nmax = max(N_GRID)
    labels = np.empty(B, np.int32)
    minimal = np.empty(B, np.int32)
    reach = np.empty(B, bool)
    errs = np.empty((B, len(S_GRID), len(N_GRID)))
    refs = []
    costs = np.array([cost(S, n) for S, n in CLASSES], dtype=np.float64)
    order = np.argsort(costs, kind="stable")
    for b0 in range(0, B, batch):
        c = coeffs[b0:b0 + batch]
        nb = c.shape[0]
        allS = [sums(c, S, nmax) for S in S_GRID]
        allS.append(sums(c, REF_S, nmax))
        ref = allS[-1][:, REF_N - 1]
        refs.append(ref)
        e = torch.empty((nb, len(S_GRID), len(N_GRID)), dtype=torch.float64, device=c.device)
        ok = torch.empty((nb, len(S_GRID), len(N_GRID)), dtype=torch.bool, device=c.device)
        for si, S in enumerate(S_GRID):
            o, up = allS[si], allS[si + 1]  # o: the delivered member, up: 2S - 1 samples
            for ni, n in enumerate(N_GRID):
                cur = o[:, n - 1]
                e[:, si, ni] = rel_err(cur, ref)
                s_ok = rel_err(up[:, n - 1], cur) <= JUDGE_FINE * tol
                if S >= max(S_GRID):
                    s_ok = torch.ones_like(s_ok)
                n_ok = (rel_err(up[:, n], up[:, n - 1]) <= JUDGE_MARGIN * tol if n < nmax
                        else torch.ones_like(s_ok))
                ok[:, si, ni] = s_ok & n_ok & (e[:, si, ni] <= tol)
        e = e.cpu().numpy()
        ok = ok.cpu().numpy().reshape(nb, -1)
        errs[b0:b0 + nb] = e
        suff = e.reshape(nb, -1)[:, order] <= tol  # classes in order of cost
        acc = ok[:, order]
        reach[b0:b0 + nb] = suff.any(axis=1)
        minimal[b0:b0 + nb] = np.where(suff.any(axis=1), order[suff.argmax(axis=1)], class_index(*PLAIN))
        labels[b0:b0 + nb] = np.where(acc.any(axis=1), order[acc.argmax(axis=1)], class_index(*PLAIN))
        if progress is not None:
            progress(b0 + nb, B)
    return labels, minimal, reach, errs, torch.cat(refs)


# ------------------------------------------------------------------------ network
class MLP:
    """Device MLP of the library (magnus_mlp_*): in -> hidden -> hidden -> out."""

    def __init__(self, in_dim: int, hidden: int = 128, out_dim: int = len(CLASSES), max_batch: int = 1024,
                 seed: int = 0):
        h = ctypes.c_void_p()
        capi.check(capi.lib().magnus_mlp_create(in_dim, hidden, out_dim, max_batch, seed, ctypes.byref(h)))
        self._h = h
        self.in_dim, self.hidden, self.out_dim, self.max_batch = in_dim, hidden, out_dim, max_batch

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            capi.lib().magnus_mlp_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def param_count(self) -> int:
        return int(capi.lib().magnus_mlp_param_count(self._h))

    def _host(self, fn) -> np.ndarray:
        buf = np.empty(self.param_count, np.float32)
        capi.check(fn(self._h, buf.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), buf.size))
        return buf

    def params(self) -> np.ndarray:
        return self._host(capi.lib().magnus_mlp_get_params)

    def grads(self) -> np.ndarray:
        return self._host(capi.lib().magnus_mlp_get_grads)

    def set_params(self, p: np.ndarray) -> None:
        p = np.ascontiguousarray(p, np.float32)
        capi.check(capi.lib().magnus_mlp_set_params(self._h, p.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                                                    p.size))

    def logits(self, x: torch.Tensor) -> torch.Tensor:
        x = x.contiguous()
        out = torch.empty((x.shape[0], self.out_dim), dtype=torch.float32, device=x.device)
        for b0 in range(0, x.shape[0], self.max_batch):
            xb = x[b0:b0 + self.max_batch]
            capi.check(capi.lib().magnus_mlp_forward(self._h, xb.shape[0], ctypes.c_void_p(xb.data_ptr()),
                                                     ctypes.c_void_p(out[b0:].data_ptr()), _stream()))
        return out

    def train_step(self, x: torch.Tensor, y: torch.Tensor, lr: float) -> torch.Tensor:
        x = x.contiguous()
        y = y.to(torch.int32).contiguous()
        loss = torch.empty(x.shape[0], dtype=torch.float32, device=x.device)
        capi.check(capi.lib().magnus_mlp_train_step(self._h, x.shape[0], ctypes.c_void_p(x.data_ptr()),
                                                    ctypes.c_void_p(y.data_ptr()), lr,
                                                    ctypes.c_void_p(loss.data_ptr()), _stream()))
        return loss


class Controller:
    """Feature standardisation + MLP + the verified escalation."""

    def __init__(self, in_dim: int, hidden: int = 128, seed: int = 0, max_batch: int = 16384):
        self.net = MLP(in_dim, hidden, len(CLASSES), max_batch, seed)
        self.mu = None
        self.sd = None

    def fit(self, feats: torch.Tensor, labels: torch.Tensor, epochs: int = 30, batch: int = 1024, lr: float = 2e-3,
            seed: int = 0):
        self.mu = feats.mean(0)
        self.sd = feats.std(0) + 1e-6
        x = (feats - self.mu) / self.sd
        y = labels.to(torch.int32)
        g = torch.Generator(device="cpu").manual_seed(seed)
        history = []
        steps = 0
        for ep in range(epochs):
            perm = torch.randperm(x.shape[0], generator=g).to(x.device)
            tot = 0.0
            for b0 in range(0, x.shape[0] - batch + 1, batch):
                idx = perm[b0:b0 + batch]
                step_lr = lr * 0.5 * (1 + math.cos(math.pi * steps / max(1, epochs * (x.shape[0] // batch))))
                tot += float(self.net.train_step(x[idx], y[idx], step_lr).sum())
                steps += 1
            history.append(tot / max(1, (x.shape[0] // batch) * batch))
        return history

    def predict(self, feats: torch.Tensor) -> torch.Tensor:
        return self.net.logits((feats - self.mu) / self.sd).argmax(1)


# --------------------------------------------------------------------- acceptance
def _budget() -> float:
    """Device bytes one call of a round may use: a share of the free memory per
    stream of the pool, at most 32 GB (large calls fill the GPU: a sweep launch
    has a workgroup per instance pair)."""
    free, _ = torch.cuda.mem_get_info()
    return max(2e9, min(32e9, 0.5 * free / _POOL[0]))


def _chunks(ids, bytes_per_instance, budget):
    """Split an id array so that a call's device buffers stay within ~budget bytes."""
    per = max(1, int(budget // max(1, bytes_per_instance)))
    return [ids[k:k + per] for k in range(0, len(ids), per)]


_STREAMS: list = []
_POOL = [4]


def set_streams(k: int) -> None:
    """Number of side streams the groups of a round are spread over (1: serial)."""
    _POOL[0] = max(1, int(k))


def _streams():
    """A small pool of side streams (default priority) for independent groups."""
    k = _POOL[0]
    while len(_STREAMS) < k:
        _STREAMS.append(torch.cuda.Stream())
    return _STREAMS[:k]


def _index(ids, device):
    return torch.from_numpy(np.ascontiguousarray(ids, dtype=np.int64)).pin_memory().to(device, non_blocking=True)


def _groups(S_arr, n_arr, ids):
    """[(S, n, ids of that resolution)] for the given instance ids, in a fixed order."""
    key = S_arr[ids] * 64 + n_arr[ids]
    order = np.argsort(key, kind="stable")
    key_sorted = key[order]
    cuts = np.flatnonzero(np.diff(key_sorted)) + 1
    return [(int(k // 64), int(k % 64), ids[part])
            for k, part in zip(key_sorted[np.r_[0, cuts]], np.split(order, cuts))]


def _resolutions(start):
    """(S, n) arrays from a list of (S, n) or an array of class indices."""
    if isinstance(start, np.ndarray) and start.ndim == 1 and np.issubdtype(start.dtype, np.integer):
        k = start.astype(np.int64)
        return (np.asarray(S_GRID, np.int64)[k // len(N_GRID)], np.asarray(N_GRID, np.int64)[k % len(N_GRID)])
    return (np.array([r[0] for r in start], dtype=np.int64), np.array([r[1] for r in start], dtype=np.int64))


def delivered(res):
    """The resolution a judged class (S, n) delivers: (S, n) itself, the coarser member of its doubling pair."""
    return [(S, n) for S, n in res]


def accept(coeffs: torch.Tensor, start, tol: float, max_rounds: int = 24, cot: torch.Tensor | None = None):
    """The judge: from per-instance starting classes (list of (S, n)), verify and
    escalate until every instance passes.  Returns (accepted classes, delivered
    sums [B][d][d] at (S, n), stats) with stats = {work: sum of (S-1)(n-1) over
    the sweeps run (a reverse sweep counts twice), rounds: per-instance rounds,
    calls, dA: per-instance dA through dA_of(stats, i), or None}.  With cot,
    every round also runs the VJP of its (S, n) sweep from that sweep's kept
    basis state, and an instance's dA is the one of the round that accepted
    it.  The groups of a round run on a pool of streams; the host waits once
    per round and updates every instance with array operations."""
    B, d = coeffs.shape[0], coeffs.shape[-1]
    S_arr, n_arr = _resolutions(start)
    done = np.zeros(B, bool)
    rounds = np.zeros(B, np.int32)
    out = torch.empty_like(coeffs[:, 0])
    # dA of each instance: chunk tensor and position of the round that accepted it
    dA = {"chunks": [], "chunk": np.full(B, -1, np.int64), "pos": np.zeros(B, np.int64)} if cot is not None else None
    work = 0
    calls = 0
    main = torch.cuda.current_stream()
    pool = _streams()
    S_top = max(S_GRID)
    budget = _budget()
    for _ in range(max_rounds):
        active = np.flatnonzero(~done)
        if active.size == 0:
            break
        ids_l, s_l, n_l, g_l = [], [], [], []
        k = 0
        for S, n, all_ids in _groups(S_arr, n_arr, active):
            S2 = 2 * S - 1
            # per instance: both sweeps' samples; with the VJP the delivered
            # sweep's kept levels, its dA and the reverse sweep's buffers
            per = (S * (2 if cot is None else n + 4) + S2 * 2) * d * d * 8
            for ids in _chunks(all_ids, per, budget):
                st = pool[k % len(pool)]
                k += 1
                st.wait_stream(main)
                with torch.cuda.stream(st):
                    idx = _index(ids, coeffs.device)
                    c = coeffs[idx]
                    if cot is not None:
                        cur, state = _forward_basis(c, S, n)
                        g = _vjp_basis(c, S, n, cot[idx], state)
                        del state
                        work += len(ids) * 3 * cost(S, n)  # sweep + reverse sweep
                    else:
                        cur = sums(c, S, n, first=n)[:, 0]
                        g = None
                        work += len(ids) * cost(S, n)
                    nm = min(n + 1, REF_N)
                    hi = sums(c, S2, nm, first=n)  # orders n..nm at 2S - 1 samples, no levels kept
                    work += len(ids) * cost(S2, nm)
                    calls += 2
                    if S < S_top:
                        s_ok = rel_err(hi[:, 0], cur) <= JUDGE_FINE * tol
                    else:  # the finer member is the reference resolution
                        s_ok = torch.ones(len(ids), dtype=torch.bool, device=coeffs.device)
                    n_ok = (rel_err(hi[:, nm - n], hi[:, 0]) <= JUDGE_MARGIN * tol if nm > n
                            else torch.ones(len(ids), dtype=torch.bool, device=coeffs.device))
                    out[idx] = cur
                    ids_l.append(ids)
                    s_l.append(s_ok)
                    n_l.append(n_ok)
                    g_l.append(g)
        for st in pool:
            main.wait_stream(st)
        s_ok = torch.cat(s_l).cpu().numpy()
        n_ok = torch.cat(n_l).cpu().numpy()
        if dA is not None:  # this round's dA (a later round overwrites the rejected instances')
            for ids, g in zip(ids_l, g_l):
                dA["chunk"][ids] = len(dA["chunks"])
                dA["pos"][ids] = np.arange(len(ids))
                dA["chunks"].append(g)
        ids = np.concatenate(ids_l)
        S_cur, n_cur = S_arr[ids], n_arr[ids]
        rounds[ids] += 1
        ok = (s_ok & n_ok) | ((S_cur >= S_top) & (n_cur >= REF_N))
        done[ids[ok]] = True
        rej = ~ok
        S_arr[ids[rej]] = np.where(s_ok[rej] | (S_cur[rej] >= S_top), S_cur[rej], 2 * S_cur[rej] - 1)
        n_arr[ids[rej]] = np.where(n_ok[rej] | (n_cur[rej] >= REF_N), n_cur[rej], n_cur[rej] + 1)
    res = list(zip(S_arr.tolist(), n_arr.tolist()))
    return res, out, {"work": work, "rounds": rounds, "calls": calls, "dA": dA}


def dA_of(stats, i: int) -> torch.Tensor:
    """Instance i's dA from accept(..., cot=...)'s stats."""
    d = stats["dA"]
    return d["chunks"][d["chunk"][i]][d["pos"][i]]


def solve(coeffs: torch.Tensor, res, cot: torch.Tensor, keep_dA: bool = False, forward_out=None):
    """Forward (op sum) + VJP of every instance at the resolution it is given
    (res: (S, n) per instance), grouped by resolution (groups spread over a
    pool of streams); returns (outputs [B][d][d], dA list or None, work in sweep
    units, a reverse sweep counting twice).  Without forward_out the forward
    keeps its basis state and the VJP starts from it; forward_out: outputs
    computed already at these resolutions (the VJP then reruns the sweep)."""
    d = coeffs.shape[-1]
    S_arr = np.array([r[0] for r in res], dtype=np.int64)
    n_arr = np.array([r[1] for r in res], dtype=np.int64)
    out = torch.empty_like(coeffs[:, 0]) if forward_out is None else forward_out
    dA = [None] * len(res) if keep_dA else None
    work = 0
    main = torch.cuda.current_stream()
    pool = _streams()
    budget = _budget()
    k = 0
    for S, n, all_ids in _groups(S_arr, n_arr, np.arange(len(res))):
        # per instance: samples, dA, the n - 1 kept levels and the reverse sweep's buffers
        for ids in _chunks(all_ids, S * d * d * 8 * (n + 4), budget):
            st = pool[k % len(pool)]
            k += 1
            st.wait_stream(main)
            with torch.cuda.stream(st):
                idx = _index(ids, coeffs.device)
                c = coeffs[idx]
                o, state = _forward_basis(c, S, n)
                if forward_out is None:
                    out[idx] = o
                g = _vjp_basis(c, S, n, cot[idx], state)  # with forward_out: the sweep rerun, then the reverse
                del state
                work += len(ids) * 3 * cost(S, n)  # sweep + reverse sweep
                if keep_dA:
                    for kk, i in enumerate(ids.tolist()):
                        dA[i] = g[kk]
    for st in pool:
        main.wait_stream(st)
    return out, dA, work


def timed(fn):
    torch.cuda.synchronize()
    t = time.perf_counter()
    r = fn()
    torch.cuda.synchronize()
    return r, time.perf_counter() - t
