"""GPU: magnus_sc_vjp_noalloc_f64, the SpaceCurve VJP on the library's fixed
device scratch array, against magnus_sc_vjp_f64 (carry NULL) through the C
ABI: bit-identical dA on two calls in a row, the second with a smaller batch,
fewer samples and a lower order (the array keeps the first call's contents),
and the oracle at 1e-12; batches beyond the array and problems off the
single-pass path are refused before any launch."""
from __future__ import annotations

import ctypes

import numpy as np
import pytest

import oracle as O
from helpers import fourier_spacecurve, rel_fro

pytestmark = pytest.mark.gpu


def _problem(n, S, integrator, op="many"):
    from magnus import capi

    return capi.problem(n, 2, S, 0.0, 1.0, op, "SpaceCurve", integrator)


def _vjp(fn, n, a, cot, integrator, op="many"):
    import torch

    from magnus import capi

    p = _problem(n, a.shape[1], integrator, op)
    dA = torch.full_like(a, float("nan"))
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    if fn == "noalloc":
        rc = capi.lib().magnus_sc_vjp_noalloc_f64(ctypes.byref(p), a.shape[0], ctypes.c_void_p(a.data_ptr()),
                                                  ctypes.c_void_p(cot.data_ptr()), ctypes.c_void_p(dA.data_ptr()),
                                                  stream)
    else:
        rc = capi.lib().magnus_sc_vjp_f64(ctypes.byref(p), a.shape[0], ctypes.c_void_p(a.data_ptr()),
                                          ctypes.c_void_p(cot.data_ptr()), None, ctypes.c_void_p(dA.data_ptr()),
                                          stream)
    assert rc == 0, capi.lib().magnus_last_error().decode()
    torch.cuda.synchronize()
    return dA


@pytest.mark.parametrize("integrator", ["Simpson", "Boole"])
def test_noalloc_vjp_matches_the_allocating_entry_point(integrator):
    import torch

    cases = [(64, 4097, 6), (5, 513, 3)]  # second call: smaller batch, samples and order
    for seed, (B, S, n) in enumerate(cases):
        print(f"[noalloc] {integrator} B={B} S={S} n={n}", flush=True)
        a = torch.from_numpy(np.stack([fourier_spacecurve(np.random.default_rng([51, seed, b]), S)
                                       for b in range(B)])).cuda()
        cot = torch.from_numpy(np.random.default_rng([51, seed]).standard_normal((B, n, 3))).cuda()
        ref = _vjp("alloc", n, a, cot, integrator)
        got = _vjp("noalloc", n, a, cot, integrator)
        assert torch.equal(got, ref)
        g0 = O.sc_vjp(n, a[0].cpu().numpy(), cot[0].cpu().numpy(), 0.0, 1.0, op="many", integrator=integrator)
        assert rel_fro(got[0].cpu().numpy(), g0) <= 1e-12


def test_noalloc_vjp_refuses_oversize_and_off_path_problems():
    import torch

    from magnus import capi

    a = torch.zeros((1, 16385, 3), dtype=torch.float64, device="cuda")
    dA = torch.empty_like(a)
    cot = torch.zeros((1, 6, 3), dtype=torch.float64, device="cuda")
    L = capi.lib()
    ptr = lambda t: ctypes.c_void_p(t.data_ptr())  # noqa: E731
    # 16384 trajectories of 16385 samples need more than the array holds: refused
    # before any launch, so the count beyond the buffers' own is never read
    p = _problem(6, 16385, "Simpson")
    assert L.magnus_sc_vjp_noalloc_f64(ctypes.byref(p), 16384, ptr(a), ptr(cot), ptr(dA), None) == capi.ERANGE
    assert "scratch bytes" in L.magnus_last_error().decode()
    p = _problem(6, 16385, "Trapezoid")
    assert L.magnus_sc_vjp_noalloc_f64(ctypes.byref(p), 1, ptr(a), ptr(cot), ptr(dA), None) == capi.EINVAL
    assert "single-pass" in L.magnus_last_error().decode()
    p = _problem(6, 16385, "Simpson")
    assert L.magnus_sc_vjp_noalloc_f64(ctypes.byref(p), 1, ptr(a), None, ptr(dA), None) == capi.EINVAL
    torch.cuda.synchronize()
