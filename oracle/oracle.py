"""ctypes front end of the CPU oracle (oracle/magnus_oracle.c).

TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg import this module; the product never does.

Names and "Auto" rules follow the reference registries:
  integrators  integration/backends.hpp:53-58   Auto -> Boole
  matrix       linalg/backends.hpp:63-68         Auto -> Fixed2 (d==2), Manual (d<64), Blas
  ops          composer/dispatch.hpp:285-300
"""
from __future__ import annotations

import ctypes
import os
import threading
from pathlib import Path

import numpy as np

_HERE = Path(__file__).resolve().parent
_LIB_PATH = _HERE / "libmagnus_oracle.so"

INTEGRATORS = {"Boole": 0, "Simpson": 1, "Trapezoid": 2, "Riemann": 3}
MATRIX = {"Fixed2": 0, "Manual": 1, "Blas": 2, "SpaceCurve": 3}
OPS = {"one": 0, "ONE": 0, "many": 1, "MANY": 1, "sum": 2, "SUM": 2}

_lib = None
_lib_mu = threading.Lock()


def build() -> None:
    """Compile the oracle with its Makefile (gcc, no GPU needed)."""
    import subprocess

    subprocess.run(["make", "-s", "-C", str(_HERE)], check=True)


def lib() -> ctypes.CDLL:
    """The loaded library (built on first use if absent); safe to call from several threads."""
    global _lib
    if _lib is not None:
        return _lib
    with _lib_mu:
        if _lib is not None:
            return _lib
        if not _LIB_PATH.exists():
            build()
        L = ctypes.CDLL(str(_LIB_PATH))
        D = ctypes.POINTER(ctypes.c_double)
        sz = ctypes.c_size_t
        i = ctypes.c_int
        d = ctypes.c_double
        L.oracle_gl_init.argtypes = [i]
        L.oracle_gl_get.argtypes = [i, D, D]
        L.oracle_gl_replace.argtypes = [i, D, D]
        L.oracle_gl_max_order.argtypes = []
        L.oracle_prefix.argtypes = [i, sz, sz, d, D]
        L.oracle_sum.argtypes = [i, sz, sz, d, D, D]
        L.oracle_prefix_vjp.argtypes = [i, sz, sz, d, D]
        L.oracle_sum_vjp.argtypes = [i, sz, sz, d, D, D]
        L.oracle_sum_vjp_add.argtypes = [i, sz, sz, d, D, D]
        L.oracle_sample_update.argtypes = [i, sz, sz, D, D, D, d]
        L.oracle_sample_update_vjp.argtypes = [i, sz, sz, D, D, D, D, d]
        L.oracle_forward.argtypes = [i, i, i, i, i, sz, sz, d, d, D, D, D]
        L.oracle_vjp.argtypes = [i, i, i, i, sz, sz, d, d, D, D, D, D]
        L.oracle_sc_forward.argtypes = [i, i, i, i, sz, d, d, D, D, D]
        L.oracle_sc_vjp.argtypes = [i, i, i, sz, d, d, D, D, D, D]
        L.oracle_forward_batch.argtypes = [i, i, i, i, i, i, sz, sz, sz, d, d, D, D, D]
        L.oracle_vjp_batch.argtypes = [i, i, i, i, i, sz, sz, sz, d, d, D, D, D, D]
        L.oracle_sc_forward_batch.argtypes = [i, i, i, i, i, sz, sz, d, d, D, D, D]
        L.oracle_sc_vjp_batch.argtypes = [i, i, i, i, sz, sz, d, d, D, D, D, D]
        L.oracle_last_error.restype = ctypes.c_char_p
        _lib = L
    return _lib


def _p(a):
    if a is None:
        return ctypes.POINTER(ctypes.c_double)()
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _check(rc):
    if rc != 0:
        raise ValueError(lib().oracle_last_error().decode())


def resolve_integrator(name: str) -> int:
    if name == "Auto":
        return INTEGRATORS["Boole"]
    return INTEGRATORS[name]


def resolve_matrix(name: str, dim: int) -> int:
    if name == "Auto":
        if dim == 2:
            return MATRIX["Fixed2"]
        return MATRIX["Manual"] if dim < 64 else MATRIX["Blas"]
    return MATRIX[name]


def gl(order: int):
    w = np.empty(order)
    x = np.empty(order)
    _check(lib().oracle_gl_get(order, _p(w), _p(x)))
    return w, x


def gl_replace(max_order: int, weights, nodes) -> None:
    w = np.ascontiguousarray(weights, dtype=np.float64)
    x = np.ascontiguousarray(nodes, dtype=np.float64)
    _check(lib().oracle_gl_replace(max_order, _p(w), _p(x)))


def gl_reset() -> None:
    _check(lib().oracle_gl_init(64))


def carry_shape(n: int, samples: int, dim: int):
    return ((n + 3) // 2, n - 1, samples, dim, dim)


# ---- primitives (operate on (len, m) float64 arrays, return new arrays) ----

def prefix(integrator: str, A, dt):
    A = np.array(A, dtype=np.float64, order="C")
    _check(lib().oracle_prefix(resolve_integrator(integrator), A.shape[0], A[0].size, dt, _p(A)))
    return A


def isum(integrator: str, A, dt, out=None):
    A = np.ascontiguousarray(A, dtype=np.float64)
    o = np.zeros(A[0].shape) if out is None else np.array(out, dtype=np.float64, order="C")
    _check(lib().oracle_sum(resolve_integrator(integrator), A.shape[0], A[0].size, dt, _p(A), _p(o)))
    return o


def prefix_vjp(integrator: str, A, dt):
    A = np.array(A, dtype=np.float64, order="C")
    _check(lib().oracle_prefix_vjp(resolve_integrator(integrator), A.shape[0], A[0].size, dt, _p(A)))
    return A


def sum_vjp(integrator: str, out, length, dt):
    out = np.ascontiguousarray(out, dtype=np.float64)
    A = np.zeros((length,) + out.shape)
    _check(lib().oracle_sum_vjp(resolve_integrator(integrator), length, out.size, dt, _p(A), _p(out)))
    return A


def sum_vjp_add(integrator: str, A, out, dt):
    A = np.array(A, dtype=np.float64, order="C")
    out = np.ascontiguousarray(out, dtype=np.float64)
    _check(lib().oracle_sum_vjp_add(resolve_integrator(integrator), A.shape[0], out.size, dt, _p(A), _p(out)))
    return A


def sample_update(matrix: str, A, Y, total, shift):
    A = np.ascontiguousarray(A, dtype=np.float64)
    Y = np.array(Y, dtype=np.float64, order="C")
    total = np.ascontiguousarray(total, dtype=np.float64)
    dim = 2 if matrix == "SpaceCurve" else A.shape[1]
    _check(lib().oracle_sample_update(MATRIX[matrix], dim, A.shape[0], _p(A), _p(Y), _p(total), shift))
    return Y


def sample_update_vjp(matrix: str, dA, A, barY, prefix_, shift):
    dA = np.array(dA, dtype=np.float64, order="C")
    A = np.ascontiguousarray(A, dtype=np.float64)
    barY = np.array(barY, dtype=np.float64, order="C")
    prefix_ = np.ascontiguousarray(prefix_, dtype=np.float64)
    dim = 2 if matrix == "SpaceCurve" else A.shape[1]
    _check(lib().oracle_sample_update_vjp(MATRIX[matrix], dim, A.shape[0], _p(dA), _p(A), _p(barY), _p(prefix_), shift))
    return dA, barY


# ---- drivers ----

def forward(n, data, t0, tf, *, op="sum", matrix_backend="Auto", integrator="Auto", lanes=1,
            record_vjp=False):
    data = np.ascontiguousarray(data, dtype=np.float64)
    S, d, _ = data.shape
    out = np.zeros((n, d, d) if op == "many" else (d, d))
    carry = np.zeros(carry_shape(n, S, d)) if record_vjp else None
    _check(lib().oracle_forward(n, OPS[op], resolve_matrix(matrix_backend, d), resolve_integrator(integrator),
                                lanes, S, d, t0, tf, _p(data), _p(out), _p(carry)))
    return (out, carry) if record_vjp else out


def vjp(n, data, cotangent, t0=0.0, tf=1.0, *, op="sum", matrix_backend="Auto", integrator="Auto",
        vjp_data=None):
    data = np.ascontiguousarray(data, dtype=np.float64)
    cot = np.ascontiguousarray(cotangent, dtype=np.float64)
    S, d, _ = data.shape
    dA = np.zeros_like(data)
    carry = None if vjp_data is None else np.ascontiguousarray(vjp_data, dtype=np.float64)
    _check(lib().oracle_vjp(n, OPS[op], resolve_matrix(matrix_backend, d), resolve_integrator(integrator),
                            S, d, t0, tf, _p(data), _p(cot), _p(carry), _p(dA)))
    return dA


def sc_forward(n, data, t0, tf, *, op="sum", integrator="Auto", lanes=1, record_vjp=False):
    data = np.ascontiguousarray(data, dtype=np.float64)
    S = data.shape[0]
    out = np.zeros((n, 3) if op == "many" else (3,))
    carry = np.zeros(carry_shape(n, S, 2)) if record_vjp else None
    _check(lib().oracle_sc_forward(n, OPS[op], resolve_integrator(integrator), lanes, S, t0, tf,
                                   _p(data), _p(out), _p(carry)))
    return (out, carry) if record_vjp else out


def sc_vjp(n, data, cotangent, t0=0.0, tf=1.0, *, op="sum", integrator="Auto", vjp_data=None):
    data = np.ascontiguousarray(data, dtype=np.float64)
    cot = np.ascontiguousarray(cotangent, dtype=np.float64)
    S = data.shape[0]
    dA = np.zeros_like(data)
    carry = None if vjp_data is None else np.ascontiguousarray(vjp_data, dtype=np.float64)
    _check(lib().oracle_sc_vjp(n, OPS[op], resolve_integrator(integrator), S, t0, tf, _p(data), _p(cot),
                               _p(carry), _p(dA)))
    return dA


def sc_forward_batch(n, data, t0, tf, *, op="many", integrator="Auto", lanes=1, threads=1, carry=None):
    """data: (B, S, 3).  carry (optional, preallocated): (B, Q, n-1, S, 2, 2)."""
    data = np.ascontiguousarray(data, dtype=np.float64)
    B, S, _ = data.shape
    out = np.zeros((B, n, 3) if op == "many" else (B, 3))
    _check(lib().oracle_sc_forward_batch(n, OPS[op], resolve_integrator(integrator), lanes, threads, B, S, t0, tf,
                                         _p(data), _p(out), _p(carry)))
    return out


def sc_vjp_batch(n, data, cot, t0, tf, *, op="many", integrator="Auto", threads=1, carry=None):
    data = np.ascontiguousarray(data, dtype=np.float64)
    cot = np.ascontiguousarray(cot, dtype=np.float64)
    B, S, _ = data.shape
    dA = np.zeros_like(data)
    _check(lib().oracle_sc_vjp_batch(n, OPS[op], resolve_integrator(integrator), threads, B, S, t0, tf,
                                     _p(data), _p(cot), _p(carry), _p(dA)))
    return dA


def forward_batch(n, data, t0, tf, *, op="sum", matrix_backend="Auto", integrator="Auto", lanes=1,
                  threads=1, carry=None):
    data = np.ascontiguousarray(data, dtype=np.float64)
    B, S, d, _ = data.shape
    out = np.zeros((B, n, d, d) if op == "many" else (B, d, d))
    _check(lib().oracle_forward_batch(n, OPS[op], resolve_matrix(matrix_backend, d), resolve_integrator(integrator),
                                      lanes, threads, B, S, d, t0, tf, _p(data), _p(out), _p(carry)))
    return out


def vjp_batch(n, data, cot, t0, tf, *, op="sum", matrix_backend="Auto", integrator="Auto", threads=1,
              carry=None):
    data = np.ascontiguousarray(data, dtype=np.float64)
    cot = np.ascontiguousarray(cot, dtype=np.float64)
    B, S, d, _ = data.shape
    dA = np.zeros_like(data)
    _check(lib().oracle_vjp_batch(n, OPS[op], resolve_matrix(matrix_backend, d), resolve_integrator(integrator),
                                  threads, B, S, d, t0, tf, _p(data), _p(cot), _p(carry), _p(dA)))
    return dA


def auto_lanes(n: int, samples: int, threads: int) -> int:
    """Reference Auto GL rule (util/gausslegendre_backends.hpp:11-19, 55-59):
    OpenMP lanes (<= Q) only when S >= ceil(30000 / ((n-1) Q))."""
    if n <= 1:
        return 1
    Q = (n + 3) // 2
    work = (n - 1) * Q
    min_samples = (30000 + work - 1) // work
    return min(Q, threads) if samples >= min_samples else 1


if os.environ.get("MAGNUS_ORACLE_EAGER"):
    lib()
