"""magnus -- Magnus expansion of Y' = A(t) Y from uniformly sampled A(t),
computed on an AMD Instinct MI355X (gfx950) by hand-written HIP kernels.

Public API (same names, arguments and defaults as the reference package,
src/api/magnus/__init__.py:23-229): sample a callable on
np.linspace(t0, tf, samples), then run the compiled core.  The compiled core
(`_core`, over libmagnus_hip.so) is required: there is no CPU fallback.
"""
from __future__ import annotations

import operator
from typing import Callable

import numpy as np

from ._runtime import bind_runtime

bind_runtime()

try:
    from ._core import _replace_gl_table
    from ._core import compute as _compute_sampled
    from ._core import compute_sc as _compute_sc_sampled
    from ._core import compute_sc_vjp as _compute_sc_vjp_sampled
    from ._core import compute_vjp as _compute_vjp_sampled
    from ._core import gl_backends
    from ._core import integrators
    from ._core import matrix_backends
    from ._core import max_order
    from ._core import numeric_backends
    from ._core import ops
except ImportError as err:  # pragma: no cover - exercised only when unbuilt
    raise ImportError(
        "magnus: the HIP extension is not built (run `python magnus-dp_amd/build.py` "
        "or __graft_entry__.build()); there is no CPU fallback"
    ) from err

__all__ = [
    "max_order",
    "numeric_backends",
    "matrix_backends",
    "integrators",
    "gl_backends",
    "ops",
    "replace_gl_table",
    "compute",
    "compute_sc",
    "compute_vjp",
    "compute_sc_vjp",
]


def replace_gl_table(max_order: int) -> None:
    """Install a Gauss-Legendre table of orders 1..max_order (numpy leggauss on [0, 1])."""
    table_order = operator.index(max_order)
    if table_order < 1:
        raise ValueError("max_order must be at least 1")
    total = table_order * (table_order + 1) // 2
    weights = np.empty(total, dtype=np.float64)
    nodes = np.empty(total, dtype=np.float64)
    offset = 0
    for order in range(1, table_order + 1):
        x, w = np.polynomial.legendre.leggauss(order)
        nodes[offset:offset + order] = (x + 1) / 2
        weights[offset:offset + order] = w / 2
        offset += order
    _replace_gl_table(table_order, weights, nodes)


def _sample(f: Callable, t0: float, tf: float, samples: int, *, dtype=None, vectorized: bool = True) -> np.ndarray:
    """Evaluate f on np.linspace(t0, tf, samples), vectorised or point by point."""
    count = operator.index(samples)
    t = np.linspace(t0, tf, count)
    values = f(t) if vectorized else [f(float(ti)) for ti in t]
    return np.ascontiguousarray(np.asarray(values, dtype=dtype))


def compute(n: int, f: Callable, t0: float, tf: float, samples: int, *, op="sum", dtype=None, vectorized=True,
            matrix_backend="Auto", integrator="Auto", gl_backend="Auto", record_vjp=False):
    """Return the Magnus output and, when requested, its saved forward carry."""
    data = _sample(f, t0, tf, samples, dtype=dtype, vectorized=vectorized)
    return _compute_sampled(n, data, t0, tf, op=op, matrix_backend=matrix_backend, integrator=integrator,
                            gl_backend=gl_backend, record_vjp=record_vjp)


def compute_sc(n: int, f: Callable, t0: float, tf: float, samples: int, *, op="sum", dtype=None, vectorized=True,
               integrator="Auto", gl_backend="Auto", record_vjp=False):
    """Return the SpaceCurve Magnus output and, when requested, its saved forward carry."""
    data = _sample(f, t0, tf, samples, dtype=dtype, vectorized=vectorized)
    return _compute_sc_sampled(n, data, t0, tf, op=op, integrator=integrator, gl_backend=gl_backend,
                               record_vjp=record_vjp)


def compute_vjp(n: int, f: Callable, cotangent, t0: float, tf: float, samples: int, *, op="sum", dtype=None,
                vectorized=True, matrix_backend="Auto", integrator="Auto", vjp_data=None):
    """Vector-Jacobian product of `compute` with respect to the sampled matrices."""
    data = _sample(f, t0, tf, samples, dtype=dtype, vectorized=vectorized)
    cotangent = np.ascontiguousarray(np.asarray(cotangent, dtype=data.dtype))
    return _compute_vjp_sampled(n, data, cotangent, vjp_data=vjp_data, t0=t0, tf=tf, op=op,
                                matrix_backend=matrix_backend, integrator=integrator)


def compute_sc_vjp(n: int, f: Callable, cotangent, t0: float, tf: float, samples: int, *, op="sum", dtype=None,
                   vectorized=True, integrator="Auto", vjp_data=None):
    """Vector-Jacobian product of `compute_sc` with respect to the sampled 3-vectors."""
    data = _sample(f, t0, tf, samples, dtype=dtype, vectorized=vectorized)
    cotangent = np.ascontiguousarray(np.asarray(cotangent, dtype=data.dtype))
    return _compute_sc_vjp_sampled(n, data, cotangent, vjp_data=vjp_data, t0=t0, tf=tf, op=op,
                                   integrator=integrator)
