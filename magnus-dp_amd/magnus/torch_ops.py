"""Device-tensor front end: batched Magnus on torch tensors resident on the GPU.

The tensors' device memory is handed to the C ABI directly (no host staging)
and work is enqueued on torch's current HIP stream.  `MagnusFunction` /
`SpaceCurveFunction` are torch.autograd.Functions whose backward is the HIP
reverse sweep, the counterpart of the reference's JAX custom_vjp
(src/api/magnus/jax.py:209-342): the forward residual is the input plus,
when recorded, the carry; without it, SpaceCurve problems on the single-pass
path keep a compact per-chunk checkpoint instead of recomputing.

Shapes (B = batch of independent instances):
  matrix:     A (B, S, d, d) -> out (B, d, d) or (B, n, d, d)
  SpaceCurve: a (B, S, 3)    -> out (B, 3)    or (B, n, 3)
"""
from __future__ import annotations

import ctypes

import torch

from . import capi


def _stream() -> ctypes.c_void_p:
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def _check_input(A: torch.Tensor, rank: int, what: str) -> None:
    if not A.is_cuda:
        raise ValueError(f"{what} must be a GPU tensor (the MI355X build has no CPU path)")
    if A.dtype != torch.float64:
        raise TypeError(f"{what} must be float64")
    if A.dim() != rank:
        raise ValueError(f"{what} must have {rank} dimensions")


def carry_shape(n: int, samples: int, dim: int):
    return ((n + 3) // 2, n - 1, samples, dim, dim)


def _check_operand(t: torch.Tensor, A: torch.Tensor, shape, what: str) -> torch.Tensor:
    """A cotangent or carry handed to the C ABI next to A: a float64 tensor on
    A's device of exactly `shape` (the kernels read it by A's problem shape), as
    the reference checks validate_matrix_output_shape / validate_vjp_data_shape
    (src/composer/backends.hpp:124-180).  Returns it contiguous."""
    if not isinstance(t, torch.Tensor):
        raise TypeError(f"{what} must be a tensor")
    if not t.is_cuda or t.device != A.device:
        raise ValueError(f"{what} must be a GPU tensor on the input's device ({A.device})")
    if t.dtype != torch.float64:
        raise TypeError(f"{what} must be float64")
    if tuple(t.shape) != tuple(shape):
        raise ValueError(f"{what} has shape {tuple(t.shape)}; expected {tuple(shape)}")
    return t.contiguous()


def forward(n: int, A: torch.Tensor, t0: float = 0.0, tf: float = 1.0, *, op: str = "sum",
            matrix_backend: str = "Auto", integrator: str = "Auto", record_vjp: bool = False):
    """Batched matrix Magnus: A (B, S, d, d) float64 on the GPU."""
    _check_input(A, 4, "A")
    A = A.contiguous()
    B, S, d, d2 = A.shape
    if d != d2:
        raise ValueError("data must contain square matrices")
    p = capi.problem(n, d, S, t0, tf, op, matrix_backend, integrator)
    out = torch.empty((B, n, d, d) if op == "many" else (B, d, d), dtype=A.dtype, device=A.device)
    carry = torch.empty((B,) + carry_shape(n, S, d), dtype=A.dtype, device=A.device) if (record_vjp and n > 1) else None
    capi.check(capi.lib().magnus_forward_f64(ctypes.byref(p), B, _ptr(A), _ptr(out), _ptr(carry), _stream()))
    if record_vjp:
        if carry is None:
            carry = torch.empty((B,) + carry_shape(n, S, d), dtype=A.dtype, device=A.device)
        return out, carry
    return out


def forward_orders(n: int, A: torch.Tensor, t0: float = 0.0, tf: float = 1.0, *, matrix_backend: str = "Auto",
                   integrator: str = "Auto", first: int = 1):
    """Magnus sums of orders first..n from one basis sweep: (B, n - first + 1, d, d),
    row k - first = forward(k, A, op="sum")."""
    _check_input(A, 4, "A")
    A = A.contiguous()
    B, S, d, d2 = A.shape
    if d != d2:
        raise ValueError("data must contain square matrices")
    p = capi.problem(n, d, S, t0, tf, "sum", matrix_backend, integrator)
    out = torch.empty((B, n - first + 1, d, d), dtype=A.dtype, device=A.device)
    if first == 1:
        capi.check(capi.lib().magnus_forward_orders_f64(ctypes.byref(p), B, _ptr(A), _ptr(out), _stream()))
    else:
        capi.check(capi.lib().magnus_forward_orders_from_f64(ctypes.byref(p), B, first, _ptr(A), _ptr(out),
                                                             _stream()))
    return out


def vjp(n: int, A: torch.Tensor, cot: torch.Tensor, carry: torch.Tensor | None = None, t0: float = 0.0,
        tf: float = 1.0, *, op: str = "sum", matrix_backend: str = "Auto", integrator: str = "Auto"):
    _check_input(A, 4, "A")
    A = A.contiguous()
    B, S, d, d2 = A.shape
    if d != d2:
        raise ValueError("data must contain square matrices")
    p = capi.problem(n, d, S, t0, tf, op, matrix_backend, integrator)
    cot = _check_operand(cot, A, (B, n, d, d) if op == "many" else (B, d, d), "cotangent")
    c = _check_operand(carry, A, (B,) + carry_shape(n, S, d), "carry") if (carry is not None and n > 1) else None
    dA = torch.empty_like(A)
    capi.check(capi.lib().magnus_vjp_f64(ctypes.byref(p), B, _ptr(A), _ptr(cot), _ptr(c), _ptr(dA), _stream()))
    return dA


def basis_state_elems(n: int, A: torch.Tensor, t0: float = 0.0, tf: float = 1.0, *, op: str = "sum",
                      matrix_backend: str = "Auto", integrator: str = "Auto") -> int:
    """Doubles of the basis state forward_basis keeps for input A (B, S, d, d); 0
    when the problem does not take the dense W-basis path (op 'one', n = 1)."""
    B, S, d, _ = A.shape
    p = capi.problem(n, d, S, t0, tf, op, matrix_backend, integrator)
    return int(capi.lib().magnus_basis_state_elems(ctypes.byref(p), B))


def forward_basis(n: int, A: torch.Tensor, t0: float = 0.0, tf: float = 1.0, *, op: str = "sum",
                  matrix_backend: str = "Auto", integrator: str = "Auto"):
    """forward() that also returns its basis state (a flat float64 tensor: the
    sweep's levels, sums and totals), from which vjp_basis starts without
    rerunning the sweep."""
    _check_input(A, 4, "A")
    A = A.contiguous()
    B, S, d, d2 = A.shape
    if d != d2:
        raise ValueError("data must contain square matrices")
    p = capi.problem(n, d, S, t0, tf, op, matrix_backend, integrator)
    elems = int(capi.lib().magnus_basis_state_elems(ctypes.byref(p), B))
    if elems == 0:
        raise ValueError("this problem has no basis state (it does not take the dense W-basis path)")
    out = torch.empty((B, n, d, d) if op == "many" else (B, d, d), dtype=A.dtype, device=A.device)
    state = torch.empty((elems,), dtype=A.dtype, device=A.device)
    capi.check(capi.lib().magnus_forward_basis_f64(ctypes.byref(p), B, _ptr(A), _ptr(out), _ptr(state), _stream()))
    return out, state


def vjp_basis(n: int, A: torch.Tensor, cot: torch.Tensor, state: torch.Tensor, t0: float = 0.0, tf: float = 1.0, *,
              op: str = "sum", matrix_backend: str = "Auto", integrator: str = "Auto"):
    """vjp() from the basis state forward_basis returned for the same A."""
    _check_input(A, 4, "A")
    A = A.contiguous()
    B, S, d, d2 = A.shape
    if d != d2:
        raise ValueError("data must contain square matrices")
    p = capi.problem(n, d, S, t0, tf, op, matrix_backend, integrator)
    cot = _check_operand(cot, A, (B, n, d, d) if op == "many" else (B, d, d), "cotangent")
    elems = int(capi.lib().magnus_basis_state_elems(ctypes.byref(p), B))
    if elems == 0:
        raise ValueError("this problem has no basis state (it does not take the dense W-basis path)")
    state = _check_operand(state, A, (elems,), "basis state")
    dA = torch.empty_like(A)
    capi.check(capi.lib().magnus_vjp_basis_f64(ctypes.byref(p), B, _ptr(A), _ptr(cot), _ptr(state), _ptr(dA),
                                               _stream()))
    return dA


def _fourier_shape(coeffs: torch.Tensor):
    _check_input(coeffs, 4, "coeffs")
    B, terms, d, d2 = coeffs.shape
    if d != d2 or terms % 2 != 1:
        raise ValueError("coeffs must have shape (B, 2*modes+1, d, d)")
    return B, (terms - 1) // 2, d


def fourier_forward_basis(n: int, coeffs: torch.Tensor, samples: int, t0: float = 0.0, tf: float = 1.0, *,
                          op: str = "sum", matrix_backend: str = "Auto", integrator: str = "Auto"):
    """forward_basis(n, sample_fourier(coeffs, samples)) with the samples formed
    inside the sweep (dim 8, Boole, n <= 8, at most 8 modes): the same results
    bit for bit, and no sample array in HBM."""
    coeffs = coeffs.contiguous()
    B, modes, d = _fourier_shape(coeffs)
    p = capi.problem(n, d, samples, t0, tf, op, matrix_backend, integrator)
    elems = int(capi.lib().magnus_basis_state_elems(ctypes.byref(p), B))
    if elems == 0:
        raise ValueError("this problem has no basis state (it does not take the dense W-basis path)")
    out = torch.empty((B, n, d, d) if op == "many" else (B, d, d), dtype=coeffs.dtype, device=coeffs.device)
    state = torch.empty((elems,), dtype=coeffs.dtype, device=coeffs.device)
    capi.check(capi.lib().magnus_fourier_forward_basis_f64(ctypes.byref(p), B, _ptr(coeffs), modes, _ptr(out),
                                                           _ptr(state), _stream()))
    return out, state


def fourier_vjp_basis(n: int, coeffs: torch.Tensor, samples: int, cot: torch.Tensor, state: torch.Tensor,
                      t0: float = 0.0, tf: float = 1.0, *, op: str = "sum", matrix_backend: str = "Auto",
                      integrator: str = "Auto"):
    """vjp_basis(n, sample_fourier(coeffs, samples), cot, state) with the samples
    formed inside the reverse sweep; dA has shape (B, samples, d, d)."""
    coeffs = coeffs.contiguous()
    B, modes, d = _fourier_shape(coeffs)
    p = capi.problem(n, d, samples, t0, tf, op, matrix_backend, integrator)
    cot = _check_operand(cot, coeffs, (B, n, d, d) if op == "many" else (B, d, d), "cotangent")
    elems = int(capi.lib().magnus_basis_state_elems(ctypes.byref(p), B))
    if elems == 0:
        raise ValueError("this problem has no basis state (it does not take the dense W-basis path)")
    state = _check_operand(state, coeffs, (elems,), "basis state")
    dA = torch.empty((B, samples, d, d), dtype=coeffs.dtype, device=coeffs.device)
    capi.check(capi.lib().magnus_fourier_vjp_basis_f64(ctypes.byref(p), B, _ptr(coeffs), modes, _ptr(cot),
                                                       _ptr(state), _ptr(dA), _stream()))
    return dA


def fourier_forward_orders(n: int, coeffs: torch.Tensor, samples: int, t0: float = 0.0, tf: float = 1.0, *,
                           matrix_backend: str = "Auto", integrator: str = "Auto", first: int = 1):
    """forward_orders(n, sample_fourier(coeffs, samples), first=first) with the
    samples formed inside the sweep."""
    coeffs = coeffs.contiguous()
    B, modes, d = _fourier_shape(coeffs)
    p = capi.problem(n, d, samples, t0, tf, "sum", matrix_backend, integrator)
    out = torch.empty((B, n - first + 1, d, d), dtype=coeffs.dtype, device=coeffs.device)
    capi.check(capi.lib().magnus_fourier_forward_orders_from_f64(ctypes.byref(p), B, first, _ptr(coeffs), modes,
                                                                 _ptr(out), _stream()))
    return out


def vjp_segmented(n: int, A: torch.Tensor, cot: torch.Tensor, t0: float = 0.0, tf: float = 1.0, *, op: str = "sum",
                  matrix_backend: str = "Auto", integrator: str = "Auto", segments: int = 0):
    """vjp() without a carry that keeps the basis levels for one time segment at
    a time (dense dim 16, 32 or 64, Boole, op many or sum, n >= 2); segments = 0
    picks the segment length.  The result does not depend on the segments."""
    _check_input(A, 4, "A")
    A = A.contiguous()
    B, S, d, d2 = A.shape
    if d != d2:
        raise ValueError("data must contain square matrices")
    p = capi.problem(n, d, S, t0, tf, op, matrix_backend, integrator)
    cot = _check_operand(cot, A, (B, n, d, d) if op == "many" else (B, d, d), "cotangent")
    dA = torch.empty_like(A)
    capi.check(capi.lib().magnus_vjp_segmented_f64(ctypes.byref(p), B, _ptr(A), _ptr(cot), _ptr(dA), int(segments),
                                                   _stream()))
    return dA


def vjp_segmented_scratch_bytes(n: int, B: int, S: int, d: int, t0: float = 0.0, tf: float = 1.0, *,
                                op: str = "sum", matrix_backend: str = "Auto", integrator: str = "Auto",
                                segments: int = 0) -> int:
    """Device bytes vjp_segmented allocates for B instances (0 where it does not apply)."""
    p = capi.problem(n, d, S, t0, tf, op, matrix_backend, integrator)
    return int(capi.lib().magnus_vjp_segmented_scratch_bytes(ctypes.byref(p), B, int(segments)))


def sc_forward(n: int, a: torch.Tensor, t0: float = 0.0, tf: float = 1.0, *, op: str = "sum",
               integrator: str = "Auto", record_vjp: bool = False):
    """Batched SpaceCurve Magnus: a (B, S, 3) float64 on the GPU."""
    _check_input(a, 3, "a")
    a = a.contiguous()
    B, S, w = a.shape
    if w != 3:
        raise ValueError("data must have shape (samples, 3)")
    p = capi.problem(n, 2, S, t0, tf, op, "SpaceCurve", integrator)
    out = torch.empty((B, n, 3) if op == "many" else (B, 3), dtype=a.dtype, device=a.device)
    carry = torch.empty((B,) + carry_shape(n, S, 2), dtype=a.dtype, device=a.device) if (record_vjp and n > 1) else None
    capi.check(capi.lib().magnus_sc_forward_f64(ctypes.byref(p), B, _ptr(a), _ptr(out), _ptr(carry), _stream()))
    if record_vjp:
        if carry is None:
            carry = torch.empty((B,) + carry_shape(n, S, 2), dtype=a.dtype, device=a.device)
        return out, carry
    return out


def sc_vjp(n: int, a: torch.Tensor, cot: torch.Tensor, carry: torch.Tensor | None = None, t0: float = 0.0,
           tf: float = 1.0, *, op: str = "sum", integrator: str = "Auto"):
    _check_input(a, 3, "a")
    a = a.contiguous()
    B, S, w = a.shape
    if w != 3:
        raise ValueError("data must have shape (samples, 3)")
    p = capi.problem(n, 2, S, t0, tf, op, "SpaceCurve", integrator)
    cot = _check_operand(cot, a, (B, n, 3) if op == "many" else (B, 3), "cotangent")
    c = _check_operand(carry, a, (B,) + carry_shape(n, S, 2), "carry") if (carry is not None and n > 1) else None
    da = torch.empty_like(a)
    capi.check(capi.lib().magnus_sc_vjp_f64(ctypes.byref(p), B, _ptr(a), _ptr(cot), _ptr(c), _ptr(da), _stream()))
    return da


def sc_checkpoint_elems(n: int, a: torch.Tensor, t0: float = 0.0, tf: float = 1.0, *, op: str = "sum",
                        integrator: str = "Auto") -> int:
    """Doubles in the compact checkpoint of sc_forward_ckpt for input a (B, S, 3);
    0 when the problem does not take the single-pass path."""
    B, S, _ = a.shape
    p = capi.problem(n, 2, S, t0, tf, op, "SpaceCurve", integrator)
    return int(capi.lib().magnus_sc_checkpoint_elems(ctypes.byref(p), B))


def sc_forward_ckpt(n: int, a: torch.Tensor, t0: float = 0.0, tf: float = 1.0, *, op: str = "sum",
                    integrator: str = "Auto"):
    """sc_forward that also returns the per-chunk summaries of the single-pass
    path (a flat float64 tensor), from which sc_vjp_ckpt starts its reverse sweep."""
    _check_input(a, 3, "a")
    a = a.contiguous()
    B, S, w = a.shape
    if w != 3:
        raise ValueError("data must have shape (samples, 3)")
    p = capi.problem(n, 2, S, t0, tf, op, "SpaceCurve", integrator)
    elems = int(capi.lib().magnus_sc_checkpoint_elems(ctypes.byref(p), B))
    out = torch.empty((B, n, 3) if op == "many" else (B, 3), dtype=a.dtype, device=a.device)
    if elems == 0:
        raise ValueError("this problem has no compact checkpoint (it does not take the single-pass path)")
    ckpt = torch.empty((elems,), dtype=a.dtype, device=a.device)
    capi.check(capi.lib().magnus_sc_forward_ckpt_f64(ctypes.byref(p), B, _ptr(a), _ptr(out), _ptr(ckpt), _stream()))
    return out, ckpt


def sc_vjp_ckpt(n: int, a: torch.Tensor, cot: torch.Tensor, ckpt: torch.Tensor, t0: float = 0.0, tf: float = 1.0, *,
                op: str = "sum", integrator: str = "Auto"):
    """sc_vjp from the summaries sc_forward_ckpt returned for the same a."""
    _check_input(a, 3, "a")
    a = a.contiguous()
    B, S, w = a.shape
    if w != 3:
        raise ValueError("data must have shape (samples, 3)")
    p = capi.problem(n, 2, S, t0, tf, op, "SpaceCurve", integrator)
    cot = _check_operand(cot, a, (B, n, 3) if op == "many" else (B, 3), "cotangent")
    elems = int(capi.lib().magnus_sc_checkpoint_elems(ctypes.byref(p), B))
    # the chunk layout depends on the batch size, so a checkpoint of another
    # batch (even a larger one) would be read with the wrong layout
    if elems == 0:
        raise ValueError("this problem has no compact checkpoint (it does not take the single-pass path)")
    if (not ckpt.is_cuda or ckpt.device != a.device or ckpt.dtype != torch.float64 or ckpt.numel() != elems
            or not ckpt.is_contiguous()):
        raise ValueError(f"ckpt must be the contiguous float64 GPU tensor of exactly {elems} elements that "
                         f"sc_forward_ckpt returned for this batch")
    da = torch.empty_like(a)
    capi.check(capi.lib().magnus_sc_vjp_ckpt_f64(ctypes.byref(p), B, _ptr(a), _ptr(cot), _ptr(ckpt), _ptr(da),
                                                 _stream()))
    return da


def _residual_fits(nbytes: int, device) -> bool:
    """A residual of nbytes leaves at least as much device memory free as it
    takes (free device memory plus torch's cached, unallocated blocks)."""
    free, _ = torch.cuda.mem_get_info(device)
    cached = torch.cuda.memory_reserved(device) - torch.cuda.memory_allocated(device)
    return 2 * nbytes <= free + cached


class MagnusFunction(torch.autograd.Function):
    """out = Magnus(A); backward = HIP reverse sweep.  record_vjp keeps the
    reference's carry as the residual; otherwise, where the dense W-basis path
    applies (op many or sum, n >= 2) and its state fits in half the free device
    memory, the residual is the basis state of the forward's sweep (n - 1
    levels instead of the carry's Q (n - 1)) and the backward skips its sweep;
    elsewhere the backward recomputes (VJP::none), keeping the basis levels for
    one time segment at a time where the full levels do not fit."""

    @staticmethod
    def forward(ctx, A, n, t0, tf, op, matrix_backend, integrator, record_vjp):
        kw = dict(op=op, matrix_backend=matrix_backend, integrator=integrator)
        mode = "none"
        if record_vjp:
            out, res = forward(n, A, t0, tf, record_vjp=True, **kw)
            mode = "carry" if n > 1 else "none"
        elif (A.is_cuda and A.dim() == 4 and (elems := basis_state_elems(n, A, t0, tf, **kw)) > 0
              and _residual_fits(8 * elems, A.device)):
            out, res = forward_basis(n, A, t0, tf, **kw)
            mode = "basis"
        else:
            out, res = forward(n, A, t0, tf, **kw), None
        ctx.save_for_backward(A, res if mode != "none" else torch.empty(0, device=A.device))
        ctx.cfg = (n, t0, tf, op, matrix_backend, integrator, mode)
        return out

    @staticmethod
    def backward(ctx, g):
        A, res = ctx.saved_tensors
        n, t0, tf, op, mb, integ, mode = ctx.cfg
        kw = dict(op=op, matrix_backend=mb, integrator=integ)
        if mode == "basis":
            dA = vjp_basis(n, A, g, res, t0, tf, **kw)
        else:
            dA = vjp(n, A, g, res if mode == "carry" else None, t0, tf, **kw)
        return dA, None, None, None, None, None, None, None


class SpaceCurveFunction(torch.autograd.Function):
    """out = Magnus(a) for su(2) inputs.  record_vjp keeps the reference's carry;
    otherwise, where the single-pass path applies, the residual is its compact
    per-chunk checkpoint (2n-1 4-vectors per chunk of samples) and the backward
    skips the forward sweep; elsewhere the backward recomputes (VJP::none)."""

    @staticmethod
    def forward(ctx, a, n, t0, tf, op, integrator, record_vjp):
        mode = "none"
        if record_vjp:
            out, res = sc_forward(n, a, t0, tf, op=op, integrator=integrator, record_vjp=True)
            mode = "carry" if n > 1 else "none"
        elif a.is_cuda and sc_checkpoint_elems(n, a, t0, tf, op=op, integrator=integrator) > 0:
            out, res = sc_forward_ckpt(n, a, t0, tf, op=op, integrator=integrator)
            mode = "ckpt"
        else:
            out, res = sc_forward(n, a, t0, tf, op=op, integrator=integrator), None
        ctx.save_for_backward(a, res if mode != "none" else torch.empty(0, device=a.device))
        ctx.cfg = (n, t0, tf, op, integrator, mode)
        return out

    @staticmethod
    def backward(ctx, g):
        a, res = ctx.saved_tensors
        n, t0, tf, op, integ, mode = ctx.cfg
        if mode == "ckpt":
            da = sc_vjp_ckpt(n, a, g, res, t0, tf, op=op, integrator=integ)
        else:
            da = sc_vjp(n, a, g, res if mode == "carry" else None, t0, tf, op=op, integrator=integ)
        return da, None, None, None, None, None, None


def sample_fourier(coeffs: torch.Tensor, samples: int, t0: float = 0.0, tf: float = 1.0) -> torch.Tensor:
    """A(t_i) = C_0 + sum_k sin(2 pi k t_i) C_{2k-1} + cos(2 pi k t_i) C_{2k} on
    np.linspace(t0, tf, samples), for coeffs (B, 2M+1, d, d) float64 on the GPU."""
    _check_input(coeffs, 4, "coeffs")
    coeffs = coeffs.contiguous()
    B, terms, d, d2 = coeffs.shape
    if d != d2 or terms % 2 != 1:
        raise ValueError("coeffs must have shape (B, 2*modes+1, d, d)")
    out = torch.empty((B, samples, d, d), dtype=torch.float64, device=coeffs.device)
    capi.check(capi.lib().magnus_sample_fourier_f64(B, samples, t0, tf, d * d, (terms - 1) // 2, _ptr(coeffs),
                                                    _ptr(out), _stream()))
    return out


def sample_sc_fourier(amp: torch.Tensor, phase: torch.Tensor, samples: int, t0: float = 0.0,
                      tf: float = 1.0) -> torch.Tensor:
    """a_c(t_i) = sum_k amp[k, c] sin(2 pi k t_i + phase[k, c]) on np.linspace(t0, tf, samples),
    for amp, phase (B, M, 3) float64 on the GPU; returns (B, samples, 3)."""
    _check_input(amp, 3, "amp")
    _check_input(phase, 3, "phase")
    amp, phase = amp.contiguous(), phase.contiguous()
    if amp.shape != phase.shape or amp.shape[2] != 3:
        raise ValueError("amp and phase must have the same shape (B, modes, 3)")
    B, modes, _ = amp.shape
    out = torch.empty((B, samples, 3), dtype=torch.float64, device=amp.device)
    capi.check(capi.lib().magnus_sample_sc_fourier_f64(B, samples, t0, tf, modes, _ptr(amp), _ptr(phase), _ptr(out),
                                                       _stream()))
    return out


def expm(omega: torch.Tensor) -> torch.Tensor:
    """Batched matrix exponential of real (B, d, d) float64 GPU tensors (the
    propagator exp(Omega) of a Magnus sum)."""
    _check_input(omega, 3, "omega")
    omega = omega.contiguous()
    B, d, d2 = omega.shape
    if d != d2:
        raise ValueError("omega must have shape (B, d, d)")
    out = torch.empty_like(omega)
    capi.check(capi.lib().magnus_expm_f64(B, d, _ptr(omega), _ptr(out), _stream()))
    return out


def sc_expm(omega: torch.Tensor) -> torch.Tensor:
    """exp(-i omega.sigma) for SpaceCurve Magnus sums omega (B, 3) float64 on the
    GPU; returns the complex128 (B, 2, 2) propagators."""
    _check_input(omega, 2, "omega")
    omega = omega.contiguous()
    if omega.shape[1] != 3:
        raise ValueError("omega must have shape (B, 3)")
    out = torch.empty((omega.shape[0], 2, 2, 2), dtype=torch.float64, device=omega.device)
    capi.check(capi.lib().magnus_sc_expm_f64(omega.shape[0], _ptr(omega), _ptr(out), _stream()))
    return torch.view_as_complex(out)


def magnus(A, n, t0=0.0, tf=1.0, *, op="sum", matrix_backend="Auto", integrator="Auto", record_vjp=True):
    """Differentiable batched Magnus on GPU tensors (A: (B, S, d, d))."""
    return MagnusFunction.apply(A, n, t0, tf, op, matrix_backend, integrator, record_vjp)


def magnus_sc(a, n, t0=0.0, tf=1.0, *, op="sum", integrator="Auto", record_vjp=True):
    """Differentiable batched SpaceCurve Magnus on GPU tensors (a: (B, S, 3))."""
    return SpaceCurveFunction.apply(a, n, t0, tf, op, integrator, record_vjp)
