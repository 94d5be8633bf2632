"""Learned resolution controller: pick (samples, order) per instance, verify, escalate.

Problem.  For an instance A(t) (dim d, Fourier form) and a tolerance tol, the
cheapest resolution (S, n) -- S in {65, ..., 16385} (2^k + 1), n in {2..8} --
whose Magnus sum is within tol (relative Frobenius) of the S = 32769, n = 8
sum.  The cost of a resolution is (S - 1)(n - 1), the number of sample
products of the basis sweep.

Pieces, all on the device:
  * labels: one all-orders sweep (magnus_forward_orders_f64) per S gives the
    sums of every order at once, so an instance is labelled with 10 sweeps;
  * features: magnus_ctrl_features_f64 on a 65-point coarse sampling;
  * network: magnus_mlp_* (3 layers, width 128, bfloat16 weights, FP32
    accumulation, Adam), classifying the 63 (S, n) classes;
  * acceptance (the judge, also the fallback): a resolution (S, n) is accepted
    when the sum at (S, n) agrees within JUDGE_MARGIN * tol with the sum at
    doubled samples (2S - 1, n) and, for n < 8, with the sum at (S, n + 1)
    (the margin covers the error left in the finer of each pair); a failed check
    doubles S and/or raises n, and repeats.  The accepted (S, n) is delivered
    (forward output and VJP at that resolution).  The learned path
    starts the judge at the network's class; the same judge started at
    (65, 2) is the method without the network; the plain method runs every
    instance at (16385, 8).
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
MAX_MODES = 8
JUDGE_MARGIN = 0.5


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


def sums(coeffs: torch.Tensor, S: int, n: int, integrator: str = "Boole") -> torch.Tensor:
    """Magnus sums of orders 1..n at S samples: [B][n][d][d]."""
    return torch_ops.forward_orders(n, torch_ops.sample_fourier(coeffs, S), 0.0, 1.0, integrator=integrator)


def rel_err(x: torch.Tensor, ref: torch.Tensor) -> torch.Tensor:
    return torch.linalg.vector_norm((x - ref).flatten(1), dim=1) / torch.linalg.vector_norm(ref.flatten(1), dim=1)


def label(coeffs: torch.Tensor, tol: float, batch: int = 512, progress=None):
    """Per instance: (labels, minimal, reachable, errors [B][9][7], reference
    sums [B][d][d]); progress(done, total) is called after every batch.  `minimal` is the cheapest class within tol of the
    reference; `labels` -- what the network learns -- is the cheapest class
    the judge accepts in one round with a result within tol (its doubled-samples
    and next-order checks pass), so a correct prediction costs one round."""
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
            o, up = allS[si], allS[si + 1]  # up: 2S - 1 samples
            for ni, n in enumerate(N_GRID):
                cur = o[:, n - 1]
                e[:, si, ni] = rel_err(cur, ref)
                s_ok = rel_err(up[:, n - 1], cur) <= JUDGE_MARGIN * tol
                n_ok = rel_err(o[:, n], cur) <= JUDGE_MARGIN * tol if n < nmax else torch.ones_like(s_ok)
                ok[:, si, ni] = s_ok & n_ok & (e[:, si, ni] <= tol)
        e = e.cpu().numpy()
        ok = ok.cpu().numpy().reshape(nb, -1)
        errs[b0:b0 + nb] = e
        flat = e.reshape(nb, -1)
        for r in range(nb):
            suff = np.nonzero(flat[r][order] <= tol)[0]
            acc = np.nonzero(ok[r][order])[0]
            reach[b0 + r] = suff.size > 0
            minimal[b0 + r] = order[suff[0]] if suff.size else class_index(*PLAIN)
            labels[b0 + r] = order[acc[0]] if acc.size else class_index(*PLAIN)
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

    def __init__(self, in_dim: int, hidden: int = 128, seed: int = 0, max_batch: int = 1024):
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
def _chunks(ids, S, d, budget=4e9):
    """Split an id list so that a call's samples stay within ~budget bytes."""
    per = max(1, int(budget // (S * d * d * 8 * 4)))
    return [ids[k:k + per] for k in range(0, len(ids), per)]


_STREAMS: list = []


def _streams(k: int = 4):
    """A small pool of side streams (default priority) for independent groups."""
    while len(_STREAMS) < k:
        _STREAMS.append(torch.cuda.Stream())
    return _STREAMS[:k]


def _index(ids, device):
    return torch.tensor(ids, dtype=torch.long).pin_memory().to(device, non_blocking=True)


def accept(coeffs: torch.Tensor, start, tol: float, max_rounds: int = 24):
    """The judge: from per-instance starting resolutions (list of (S, n)),
    verify and escalate until every instance passes.  Returns (accepted
    resolutions, accepted sums [B][d][d], stats) with stats = {work: sum of
    (S-1)(n-1) over the sweeps run, rounds: per-instance rounds, calls}.
    The groups of a round run on a pool of streams; the host waits once per
    round."""
    B, d = coeffs.shape[0], coeffs.shape[-1]
    res = list(start)
    done = [False] * B
    rounds = np.zeros(B, np.int32)
    out = torch.empty_like(coeffs[:, 0])
    work = 0
    calls = 0
    main = torch.cuda.current_stream()
    pool = _streams()
    for _ in range(max_rounds):
        groups = {}
        for i in range(B):
            if not done[i]:
                groups.setdefault(res[i], []).append(i)
        if not groups:
            break
        pending = []
        k = 0
        for (S, n), all_ids in groups.items():
            for ids in _chunks(all_ids, S, d):
                st = pool[k % len(pool)]
                k += 1
                st.wait_stream(main)
                with torch.cuda.stream(st):
                    idx = _index(ids, coeffs.device)
                    c = coeffs[idx]
                    nm = min(n + 1, REF_N)
                    hi = sums(c, S, nm)
                    cur = hi[:, n - 1]
                    if S < REF_S:
                        up = sums(c, 2 * S - 1, n)[:, n - 1]
                        s_ok = rel_err(up, cur) <= JUDGE_MARGIN * tol
                        calls += 1
                        work += len(ids) * cost(2 * S - 1, n)
                    else:
                        s_ok = torch.ones(len(ids), dtype=torch.bool, device=coeffs.device)
                    n_ok = (rel_err(hi[:, nm - 1], cur) <= JUDGE_MARGIN * tol if nm > n
                            else torch.ones(len(ids), dtype=torch.bool, device=coeffs.device))
                    out[idx] = cur
                    calls += 1
                    work += len(ids) * cost(S, nm)
                    pending.append(((S, n), ids, s_ok, n_ok))
        for st in pool:
            main.wait_stream(st)
        torch.cuda.synchronize()
        for (S, n), ids, s_ok, n_ok in pending:
            s_ok = s_ok.cpu().numpy()
            n_ok = n_ok.cpu().numpy()
            for kk, i in enumerate(ids):
                rounds[i] += 1
                if (s_ok[kk] and n_ok[kk]) or (S >= REF_S and n >= REF_N):
                    done[i] = True
                    res[i] = (S, n)
                else:
                    S2 = S if s_ok[kk] or S >= REF_S else 2 * S - 1
                    n2 = n if n_ok[kk] or n >= REF_N else n + 1
                    res[i] = (S2, n2)
    return res, out, {"work": work, "rounds": rounds, "calls": calls}


def solve(coeffs: torch.Tensor, res, cot: torch.Tensor, keep_dA: bool = False, forward_out=None):
    """Forward (op sum) + VJP of every instance at its resolution, grouped by
    resolution (groups spread over a pool of streams); returns (outputs
    [B][d][d], dA list or None, work in sweep units).  forward_out: outputs
    the judge already computed at these resolutions (the forward is then not
    repeated)."""
    d = coeffs.shape[-1]
    groups = {}
    for i, r in enumerate(res):
        groups.setdefault(tuple(r), []).append(i)
    out = torch.empty_like(coeffs[:, 0]) if forward_out is None else forward_out
    dA = [None] * len(res) if keep_dA else None
    work = 0
    main = torch.cuda.current_stream()
    pool = _streams()
    k = 0
    for (S, n), all_ids in groups.items():
        for ids in _chunks(all_ids, S, d):
            st = pool[k % len(pool)]
            k += 1
            st.wait_stream(main)
            with torch.cuda.stream(st):
                idx = _index(ids, coeffs.device)
                A = torch_ops.sample_fourier(coeffs[idx], S)
                if forward_out is None:
                    out[idx] = torch_ops.forward(n, A, 0.0, 1.0, op="sum", integrator="Boole")
                    work += len(ids) * cost(S, n)
                g = torch_ops.vjp(n, A, cot[idx], None, 0.0, 1.0, op="sum", integrator="Boole")
                work += len(ids) * 3 * cost(S, n)  # forward recompute + reverse sweep
                if keep_dA:
                    for kk, i in enumerate(ids):
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
