"""Uneven batches: many matrix problems of different dim, samples and order in
one enqueue or one hipGraph launch (C ABI magnus_uneven_*).

    plan = UnevenBatch([dict(n=4, dim=8, samples=1025), dict(n=2, dim=64, samples=257), ...])
    plan.autotune(A_list, cot_list)          # optional: product path per bucket
    outs = plan.forward(A_list)              # list of (dim, dim) / (n, dim, dim) tensors
    dAs = plan.vjp(A_list, cot_list)         # list of (samples, dim, dim) tensors

Each instance's results are those of a per-instance magnus_forward_f64 /
magnus_vjp_f64 call (VJP without a carry) with the plan's product path.
"""
from __future__ import annotations

import ctypes

import torch

from . import capi


def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


class UnevenBatch:
    def __init__(self, problems, workspace_bytes: int = 0, branches: int = 4):
        self.problems = [dict(p) for p in problems]
        arr = (capi.Problem * max(len(self.problems), 1))()
        for i, p in enumerate(self.problems):
            arr[i] = capi.problem(p["n"], p["dim"], p["samples"], p.get("t0", 0.0), p.get("tf", 1.0),
                                  p.get("op", "sum"), p.get("matrix_backend", "Auto"), p.get("integrator", "Auto"))
        self._arr = arr
        h = ctypes.c_void_p()
        capi.check(capi.lib().magnus_uneven_create(arr, len(self.problems), workspace_bytes, branches,
                                                   ctypes.byref(h)))
        self._h = h

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            capi.lib().magnus_uneven_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __len__(self):
        return len(self.problems)

    def _ptrs(self, tensors, shape_of, what):
        if len(tensors) != len(self.problems):
            raise ValueError(f"{what}: expected {len(self.problems)} tensors")
        arr = (ctypes.c_void_p * max(len(tensors), 1))()
        keep = []
        for i, t in enumerate(tensors):
            if not t.is_cuda or t.dtype != torch.float64:
                raise TypeError(f"{what}[{i}] must be a float64 GPU tensor")
            if tuple(t.shape) != shape_of(self.problems[i]):
                raise ValueError(f"{what}[{i}] has shape {tuple(t.shape)}, expected {shape_of(self.problems[i])}")
            t = t.contiguous()
            keep.append(t)
            arr[i] = t.data_ptr()
        return arr, keep

    @staticmethod
    def _in_shape(p):
        return (p["samples"], p["dim"], p["dim"])

    @staticmethod
    def _out_shape(p):
        return (p["n"], p["dim"], p["dim"]) if p.get("op", "sum") == "many" else (p["dim"], p["dim"])

    def forward(self, A, use_graph: bool = True):
        ain, keep = self._ptrs(A, self._in_shape, "A")
        outs = [torch.empty(self._out_shape(p), dtype=torch.float64, device=A[i].device)
                for i, p in enumerate(self.problems)]
        aout, keep2 = self._ptrs(outs, self._out_shape, "out")
        capi.check(capi.lib().magnus_uneven_forward_f64(self._h, ain, aout, int(use_graph), _stream()))
        return outs

    def vjp(self, A, cot, use_graph: bool = True):
        ain, keep = self._ptrs(A, self._in_shape, "A")
        acot, keep2 = self._ptrs(cot, self._out_shape, "cot")
        dA = [torch.empty(self._in_shape(p), dtype=torch.float64, device=A[i].device)
              for i, p in enumerate(self.problems)]
        adA, keep3 = self._ptrs(dA, self._in_shape, "dA")
        capi.check(capi.lib().magnus_uneven_vjp_f64(self._h, ain, acot, adA, int(use_graph), _stream()))
        return dA

    def autotune(self, A, cot):
        ain, keep = self._ptrs(A, self._in_shape, "A")
        acot, keep2 = self._ptrs(cot, self._out_shape, "cot")
        dA = [torch.empty(self._in_shape(p), dtype=torch.float64, device=A[i].device)
              for i, p in enumerate(self.problems)]
        adA, keep3 = self._ptrs(dA, self._in_shape, "dA")
        capi.check(capi.lib().magnus_uneven_autotune(self._h, ain, acot, adA, _stream()))

    @property
    def tuning(self) -> str:
        L = capi.lib()
        n = L.magnus_uneven_tuning(self._h, None, 0)
        buf = ctypes.create_string_buffer(n + 1)
        L.magnus_uneven_tuning(self._h, buf, n + 1)
        return buf.value.decode()

    def set_tuning(self, text: str) -> None:
        capi.check(capi.lib().magnus_uneven_set_tuning(self._h, text.encode()))

    def in_pipeline(self, i: int) -> bool:
        """Whether instance i runs in the grouped walker pipeline (its results then
        agree with a per-instance call to rounding) rather than in a bucket (bit
        for bit)."""
        return bool(capi.lib().magnus_uneven_in_pipeline(self._h, i))

    def info(self) -> dict:
        L = capi.lib()
        return {k: int(L.magnus_uneven_info(self._h, i))
                for i, k in enumerate(("buckets", "sub_batches", "branches", "workspace_bytes", "pipeline_instances",
                                       "pipeline_waves"))}
