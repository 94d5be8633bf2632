"""ctypes binding of libmagnus_hip.so (include/magnus_capi.h).

Used by the device-tensor front end (magnus.torch_ops) and by bench.py to call
the C ABI with device pointers and a HIP stream, without staging through host
memory.  The library must be built; loading fails loudly otherwise.
"""
from __future__ import annotations

import ctypes
import re
from pathlib import Path

_HERE = Path(__file__).resolve().parent
LIB_PATH = _HERE / "libmagnus_hip.so"
HEADER_PATH = _HERE.parents[1] / "include" / "magnus_capi.h"

OK, EINVAL, ERANGE, ENOMEM, EDEVICE = 0, 1, 2, 3, 4
OP = {"one": 0, "many": 1, "sum": 2}
AUTO = -1


class Problem(ctypes.Structure):
    _fields_ = [
        ("n", ctypes.c_uint64),
        ("dim", ctypes.c_uint64),
        ("samples", ctypes.c_uint64),
        ("t0", ctypes.c_double),
        ("tf", ctypes.c_double),
        ("op", ctypes.c_int32),
        ("matrix_backend", ctypes.c_int32),
        ("integrator", ctypes.c_int32),
        ("gl_backend", ctypes.c_int32),
    ]


_lib = None


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise ImportError(f"{LIB_PATH} is not built; run `python magnus-dp_amd/build.py`")
        from ._runtime import bind_runtime

        bind_runtime()
        L = ctypes.CDLL(str(LIB_PATH))
        P = ctypes.POINTER(Problem)
        vp = ctypes.c_void_p
        u64 = ctypes.c_uint64
        i32p = ctypes.POINTER(ctypes.c_int32)
        for name in ("magnus_forward_f64", "magnus_sc_forward_f64"):
            getattr(L, name).argtypes = [P, u64, vp, vp, vp, vp]
        for name in ("magnus_vjp_f64", "magnus_sc_vjp_f64"):
            getattr(L, name).argtypes = [P, u64, vp, vp, vp, vp, vp]
        L.magnus_forward_orders_f64.argtypes = [P, u64, vp, vp, vp]
        L.magnus_forward_orders_from_f64.argtypes = [P, u64, u64, vp, vp, vp]
        L.magnus_sc_forward_ckpt_f64.argtypes = [P, u64, vp, vp, vp, vp]
        L.magnus_sc_vjp_ckpt_f64.argtypes = [P, u64, vp, vp, vp, vp, vp]
        L.magnus_sc_vjp_noalloc_f64.argtypes = [P, u64, vp, vp, vp, vp]
        L.magnus_basis_state_elems.argtypes = [P, u64]
        L.magnus_basis_state_elems.restype = u64
        L.magnus_forward_basis_f64.argtypes = [P, u64, vp, vp, vp, vp]
        L.magnus_vjp_basis_f64.argtypes = [P, u64, vp, vp, vp, vp, vp]
        L.magnus_fourier_forward_basis_f64.argtypes = [P, u64, vp, u64, vp, vp, vp]
        L.magnus_fourier_vjp_basis_f64.argtypes = [P, u64, vp, u64, vp, vp, vp, vp]
        L.magnus_fourier_forward_orders_from_f64.argtypes = [P, u64, u64, vp, u64, vp, vp]
        L.magnus_vjp_segmented_f64.argtypes = [P, u64, vp, vp, vp, u64, vp]
        L.magnus_vjp_segmented_scratch_bytes.argtypes = [P, u64, u64]
        L.magnus_vjp_segmented_scratch_bytes.restype = u64
        L.magnus_vjp_segment_length.argtypes = [P, u64]
        L.magnus_vjp_segment_length.restype = u64
        L.magnus_sc_checkpoint_elems.argtypes = [P, u64]
        L.magnus_sc_checkpoint_elems.restype = u64
        for name in ("magnus_carry_elems", "magnus_output_elems", "magnus_sc_carry_elems", "magnus_sc_output_elems"):
            getattr(L, name).argtypes = [P]
            getattr(L, name).restype = u64
        L.magnus_validate.argtypes = [P, ctypes.c_int, ctypes.c_int]
        for name in ("magnus_resolve_op", "magnus_resolve_matrix_backend", "magnus_resolve_integrator",
                     "magnus_resolve_gl_backend"):
            getattr(L, name).argtypes = [ctypes.c_char_p, i32p]
        for name in ("magnus_matrix_backend_name", "magnus_integrator_name", "magnus_gl_backend_name",
                     "magnus_op_name", "magnus_numeric_backend_name"):
            getattr(L, name).argtypes = [ctypes.c_int32]
            getattr(L, name).restype = ctypes.c_char_p
        L.magnus_max_order.restype = u64
        L.magnus_gl_order.argtypes = [u64, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]
        L.magnus_replace_gl_table.argtypes = [u64, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]
        d = ctypes.c_double
        L.magnus_sample_fourier_f64.argtypes = [u64, u64, d, d, u64, u64, vp, vp, vp]
        L.magnus_sample_sc_fourier_f64.argtypes = [u64, u64, d, d, u64, vp, vp, vp, vp]
        L.magnus_expm_f64.argtypes = [u64, u64, vp, vp, vp]
        L.magnus_sc_expm_f64.argtypes = [u64, vp, vp, vp]
        pp = ctypes.POINTER(ctypes.c_void_p)
        L.magnus_uneven_create.argtypes = [P, u64, u64, ctypes.c_int, ctypes.POINTER(ctypes.c_void_p)]
        L.magnus_uneven_destroy.argtypes = [vp]
        L.magnus_uneven_destroy.restype = None
        L.magnus_uneven_forward_f64.argtypes = [vp, pp, pp, ctypes.c_int, vp]
        L.magnus_uneven_vjp_f64.argtypes = [vp, pp, pp, pp, ctypes.c_int, vp]
        L.magnus_uneven_autotune.argtypes = [vp, pp, pp, pp, vp]
        L.magnus_uneven_tuning.argtypes = [vp, ctypes.c_char_p, u64]
        L.magnus_uneven_tuning.restype = u64
        L.magnus_uneven_set_tuning.argtypes = [vp, ctypes.c_char_p]
        L.magnus_uneven_info.argtypes = [vp, ctypes.c_int]
        L.magnus_uneven_info.restype = u64
        L.magnus_uneven_in_pipeline.argtypes = [vp, u64]
        f32p = ctypes.POINTER(ctypes.c_float)
        u32 = ctypes.c_uint32
        L.magnus_ctrl_feature_count.restype = u64
        L.magnus_ctrl_features_f64.argtypes = [u64, u64, u64, d, d, vp, vp, vp]
        L.magnus_mlp_create.argtypes = [u32, u32, u32, u64, u64, ctypes.POINTER(ctypes.c_void_p)]
        L.magnus_mlp_destroy.argtypes = [vp]
        L.magnus_mlp_destroy.restype = None
        L.magnus_mlp_param_count.argtypes = [vp]
        L.magnus_mlp_param_count.restype = u64
        for name in ("magnus_mlp_get_params", "magnus_mlp_set_params", "magnus_mlp_get_grads"):
            getattr(L, name).argtypes = [vp, f32p, u64]
        L.magnus_mlp_forward.argtypes = [vp, u64, vp, vp, vp]
        L.magnus_mlp_train_step.argtypes = [vp, u64, vp, vp, ctypes.c_float, vp, vp]
        L.magnus_release_scratch.argtypes = []
        L.magnus_set_scratch_limit.argtypes = [u64]
        L.magnus_scratch_limit.argtypes = []
        L.magnus_scratch_limit.restype = u64
        L.magnus_timing_enable.argtypes = [ctypes.c_int]
        L.magnus_timing_read.argtypes = [ctypes.c_char_p, ctypes.POINTER(u64), ctypes.POINTER(ctypes.c_double)]
        L.magnus_timing_counts.argtypes = [ctypes.c_char_p, ctypes.POINTER(ctypes.c_double),
                                           ctypes.POINTER(ctypes.c_double)]
        L.magnus_timing_reset.restype = None
        L.magnus_last_error.restype = ctypes.c_char_p
        L.magnus_last_kernel.restype = ctypes.c_char_p
        _lib = L
    return _lib


def header_functions() -> list[str]:
    """Names of every function declared in include/magnus_capi.h."""
    text = HEADER_PATH.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(magnus_[a-z0-9_]+)\s*\(", text)))


def check(rc: int) -> None:
    if rc != OK:
        msg = lib().magnus_last_error().decode()
        if rc == EINVAL:
            raise ValueError(msg)
        if rc == ERANGE:
            raise IndexError(msg)
        if rc == ENOMEM:
            raise MemoryError(msg)
        raise RuntimeError(msg)


def _resolve(fn, name: str) -> int:
    v = ctypes.c_int32()
    check(fn(name.encode(), ctypes.byref(v)))
    return v.value


def problem(n: int, dim: int, samples: int, t0: float, tf: float, op: str = "sum", matrix_backend: str = "Auto",
            integrator: str = "Auto", gl_backend: str = "Auto") -> Problem:
    L = lib()
    return Problem(n, dim, samples, t0, tf, _resolve(L.magnus_resolve_op, op),
                   _resolve(L.magnus_resolve_matrix_backend, matrix_backend),
                   _resolve(L.magnus_resolve_integrator, integrator),
                   _resolve(L.magnus_resolve_gl_backend, gl_backend))


def timing_enable(on: bool = True) -> None:
    lib().magnus_timing_enable(1 if on else 0)


def timing_reset() -> None:
    lib().magnus_timing_reset()


def timing_read(name: str) -> tuple[int, float]:
    """(launches, total milliseconds) of kernel family `name` since the last reset."""
    n = ctypes.c_uint64()
    ms = ctypes.c_double()
    check(lib().magnus_timing_read(name.encode(), ctypes.byref(n), ctypes.byref(ms)))
    return n.value, ms.value


def timing_counts(name: str) -> tuple[float, float]:
    """(algorithmic HBM bytes, flops) of the timed launches of kernel family `name`
    since the last reset, as their launch sites state them."""
    b = ctypes.c_double()
    f = ctypes.c_double()
    check(lib().magnus_timing_counts(name.encode(), ctypes.byref(b), ctypes.byref(f)))
    return b.value, f.value


def last_kernel() -> str:
    return lib().magnus_last_kernel().decode()


def set_scratch_limit(nbytes: int) -> None:
    """Bound the bytes a call may take from its stream's scratch cache (0: none)."""
    check(lib().magnus_set_scratch_limit(int(nbytes)))


def scratch_limit() -> int:
    return int(lib().magnus_scratch_limit())
