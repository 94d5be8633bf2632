"""CPU-side checks of the product library: it loads, exports every function
include/magnus_capi.h declares, and applies the reference's validation rules
and message texts before any device work (no GPU needed)."""
from __future__ import annotations

import ctypes

import numpy as np
import pytest

from magnus import capi
from helpers import matrix_values, spacecurve_values


def test_library_exports_every_header_function():
    L = capi.lib()
    names = capi.header_functions()
    assert len(names) >= 20
    for name in names:
        assert hasattr(L, name), name


def test_core_module_loads_native_library():
    import magnus
    from magnus import _core

    assert "libmagnus_hip" in open("/proc/self/maps").read()
    assert magnus.max_order() == 64
    assert magnus.ops() == ("one", "many", "sum")
    assert magnus.matrix_backends() == ("Fixed2", "Manual", "Blas", "SpaceCurve", "Auto")
    assert magnus.integrators() == ("Boole", "Simpson", "Trapezoid", "Riemann", "Auto")
    assert magnus.gl_backends() == ("threadpool", "openmp", "serial", "Auto")
    assert "double" in magnus.numeric_backends() and "Auto" not in magnus.numeric_backends()
    assert callable(_core.compute)


def test_public_api_matches_reference_surface():
    import magnus

    assert set(magnus.__all__) == {
        "max_order", "numeric_backends", "matrix_backends", "integrators", "gl_backends", "ops",
        "replace_gl_table", "compute", "compute_sc", "compute_vjp", "compute_sc_vjp",
    }
    for removed in ["one", "many", "sum", "one_sc", "many_sc", "sum_sc"]:
        assert not hasattr(magnus, removed)


def _validate(**kw):
    base = dict(n=4, dim=2, samples=9, t0=0.0, tf=1.0, op="sum", matrix_backend="Auto", integrator="Auto",
                gl_backend="Auto")
    base.update(kw)
    sc = base.pop("sc", False)
    fwd = base.pop("forward", True)
    p = capi.problem(**base)
    rc = capi.lib().magnus_validate(ctypes.byref(p), int(sc), int(fwd))
    return rc, capi.lib().magnus_last_error().decode()


def test_validation_rules_and_messages():
    assert _validate()[0] == capi.OK
    assert _validate(n=0) == (capi.EINVAL, "n must be at least 1")
    assert _validate(samples=1) == (capi.EINVAL, "data must contain at least two samples")
    assert _validate(dim=3, matrix_backend="Fixed2")[1] == "invalid dispatch option (based on params)"
    assert _validate(dim=3, matrix_backend="SpaceCurve")[1] == "invalid dispatch option (based on params)"
    assert "divisibility" in _validate(samples=7, integrator="Boole")[1]
    assert _validate(samples=7, integrator="Simpson")[0] == capi.OK
    assert "divisibility" in _validate(samples=8, integrator="Simpson")[1]
    assert _validate(samples=8, integrator="Riemann")[0] == capi.OK
    assert _validate(gl_backend="threadpool")[1] == "invalid dispatch option (based on params)"
    assert _validate(gl_backend="threadpool", forward=False)[0] == capi.OK
    rc, msg = _validate(n=127)  # Q = (n + 3) // 2 = 65 > max_order 64
    assert rc == capi.EINVAL and msg.startswith("requested magnus order exceeds GL table")
    assert _validate(n=126)[0] == capi.OK
    with pytest.raises(ValueError, match="Could not resolve dispatch name option"):
        capi.problem(4, 2, 9, 0.0, 1.0, matrix_backend="Missing")
    with pytest.raises(ValueError, match="deduce kernel op"):
        capi.problem(4, 2, 9, 0.0, 1.0, op="missing")


def test_gl_table_replacement_moves_the_order_bound():
    """magnus_replace_gl_table (pybind.cpp:102-116, gausslegendre.cpp:74-111):
    max_order follows the installed table and the order check uses it; numpy's
    64-order table (the default, bit for bit) is restored afterwards."""
    import magnus

    try:
        magnus.replace_gl_table(5)
        assert magnus.max_order() == 5
        assert _validate(n=8)[0] == capi.OK  # Q = 5
        rc, msg = _validate(n=9)  # Q = 6 > 5
        assert rc == capi.EINVAL and msg.startswith("requested magnus order exceeds GL table")
        with pytest.raises(ValueError):
            magnus.replace_gl_table(0)
        w = np.empty(5)
        x = np.empty(5)
        dp = ctypes.POINTER(ctypes.c_double)
        assert capi.lib().magnus_gl_order(5, w.ctypes.data_as(dp), x.ctypes.data_as(dp)) == 0
        xr, wr = np.polynomial.legendre.leggauss(5)
        assert np.array_equal(w, wr / 2) and np.array_equal(x, (xr + 1) / 2)
        assert capi.lib().magnus_gl_order(6, None, None) == capi.ERANGE
    finally:
        magnus.replace_gl_table(64)
    assert magnus.max_order() == 64
    assert _validate(n=126)[0] == capi.OK and _validate(n=127)[0] == capi.EINVAL


def test_sizes():
    p = capi.problem(4, 3, 9, 0.0, 1.0, op="many")
    assert capi.lib().magnus_carry_elems(ctypes.byref(p)) == 3 * 3 * 9 * 9
    assert capi.lib().magnus_output_elems(ctypes.byref(p)) == 4 * 9
    assert capi.lib().magnus_sc_carry_elems(ctypes.byref(p)) == 3 * 3 * 9 * 4
    assert capi.lib().magnus_sc_output_elems(ctypes.byref(p)) == 4 * 3
    p1 = capi.problem(1, 3, 9, 0.0, 1.0)
    assert capi.lib().magnus_carry_elems(ctypes.byref(p1)) == 0


def test_segmented_vjp_sizes_and_scratch_limit():
    """Segment lengths are block multiples of about sqrt(2 K (S - 1) / (K + 3))
    samples (or (S - 1) / segments), the segmented VJP's scratch is 0 where it
    does not apply, and the scratch bound reads back as set."""
    L = capi.lib()
    p = capi.problem(4, 16, 65, 0.0, 1.0, op="sum", integrator="Boole")
    assert L.magnus_vjp_segment_length(ctypes.byref(p), 0) == 8
    assert L.magnus_vjp_segment_length(ctypes.byref(p), 2) == 32
    assert L.magnus_vjp_segment_length(ctypes.byref(p), 3) == 24
    assert L.magnus_vjp_segment_length(ctypes.byref(p), 1000) == 4
    assert L.magnus_vjp_segmented_scratch_bytes(ctypes.byref(p), 2, 0) > 0
    for kw in (dict(d=8, op="sum", integ="Boole"), dict(d=16, op="one", integ="Boole"),
               dict(d=16, op="sum", integ="Simpson"), dict(d=48, op="sum", integ="Boole")):
        q = capi.problem(4, kw["d"], 65, 0.0, 1.0, op=kw["op"], integrator=kw["integ"])
        assert L.magnus_vjp_segmented_scratch_bytes(ctypes.byref(q), 2, 0) == 0, kw
    assert capi.scratch_limit() == 0
    try:
        capi.set_scratch_limit(1 << 20)
        assert capi.scratch_limit() == 1 << 20
    finally:
        capi.set_scratch_limit(0)
    assert capi.scratch_limit() == 0


def test_sampling_argument_checks_before_device_work():
    L = capi.lib()
    dummy = ctypes.c_void_p(16)  # never dereferenced: every call below fails validation
    assert L.magnus_sample_fourier_f64(1, 1, 0.0, 1.0, 4, 2, dummy, dummy, None) == capi.EINVAL
    assert L.magnus_sample_fourier_f64(1, 9, 0.0, 1.0, 0, 2, dummy, dummy, None) == capi.EINVAL
    assert L.magnus_sample_fourier_f64(1, 9, 0.0, 1.0, 4, 33, dummy, dummy, None) == capi.ERANGE
    assert L.magnus_sample_fourier_f64(1, 9, 0.0, 1.0, 4, 2, None, dummy, None) == capi.EINVAL
    assert L.magnus_sample_sc_fourier_f64(1, 9, 0.0, 1.0, 40, dummy, dummy, dummy, None) == capi.ERANGE
    assert L.magnus_sample_sc_fourier_f64(1, 9, 0.0, 1.0, 4, dummy, None, dummy, None) == capi.EINVAL
    L.magnus_sample_sc_fourier_f64(1, 9, 0.0, 1.0, 40, dummy, dummy, dummy, None)
    assert "modes" in L.magnus_last_error().decode()
    assert L.magnus_sample_fourier_f64(0, 9, 0.0, 1.0, 4, 2, dummy, dummy, None) == capi.OK  # empty batch


def test_fourier_input_argument_checks_before_device_work():
    """The Fourier-form entry points refuse, before any device call, problems
    off the d = 8 single-pass sweeps (other dims, Simpson, orders above 8, op
    one), more than 8 modes, an order range outside 1..n and null pointers;
    an empty batch is a no-op."""
    L = capi.lib()
    dummy = ctypes.c_void_p(16)  # never dereferenced: every call below fails validation or is empty
    ok = capi.problem(4, 8, 65, 0.0, 1.0, op="sum", integrator="Boole")
    fwd, vjp, orders = (L.magnus_fourier_forward_basis_f64, L.magnus_fourier_vjp_basis_f64,
                        L.magnus_fourier_forward_orders_from_f64)
    for kw in (dict(n=4, d=16, integ="Boole", op="sum"), dict(n=4, d=8, integ="Simpson", op="sum"),
               dict(n=9, d=8, integ="Boole", op="sum")):
        q = capi.problem(kw["n"], kw["d"], 65, 0.0, 1.0, op=kw["op"], integrator=kw["integ"])
        assert fwd(ctypes.byref(q), 1, dummy, 3, dummy, dummy, None) == capi.EINVAL, kw
        assert vjp(ctypes.byref(q), 1, dummy, 3, dummy, dummy, dummy, None) == capi.EINVAL, kw
        assert orders(ctypes.byref(q), 1, 1, dummy, 3, dummy, None) == capi.EINVAL, kw
    assert fwd(ctypes.byref(ok), 1, dummy, 9, dummy, dummy, None) == capi.ERANGE
    assert "modes" in L.magnus_last_error().decode()
    assert vjp(ctypes.byref(ok), 1, dummy, 9, dummy, dummy, dummy, None) == capi.ERANGE
    assert orders(ctypes.byref(ok), 1, 1, dummy, 9, dummy, None) == capi.ERANGE
    one = capi.problem(4, 8, 65, 0.0, 1.0, op="one", integrator="Boole")
    assert fwd(ctypes.byref(one), 1, dummy, 3, dummy, dummy, None) == capi.EINVAL
    assert vjp(ctypes.byref(one), 1, dummy, 3, dummy, dummy, dummy, None) == capi.EINVAL
    assert orders(ctypes.byref(ok), 1, 0, dummy, 3, dummy, None) == capi.EINVAL
    assert orders(ctypes.byref(ok), 1, 5, dummy, 3, dummy, None) == capi.EINVAL
    assert fwd(ctypes.byref(ok), 1, None, 3, dummy, dummy, None) == capi.EINVAL
    assert vjp(ctypes.byref(ok), 1, dummy, 3, dummy, None, dummy, None) == capi.EINVAL
    assert orders(ctypes.byref(ok), 1, 1, dummy, 3, None, None) == capi.EINVAL
    assert fwd(ctypes.byref(ok), 0, dummy, 3, dummy, dummy, None) == capi.OK  # empty batch
    assert orders(ctypes.byref(ok), 0, 2, dummy, 3, dummy, None) == capi.OK


def test_spacecurve_checkpoint_size_and_argument_checks():
    """Compact checkpoint: 2(n-1)+1 4-vectors per chunk of samples on the
    single-pass path; none (size 0, EINVAL) elsewhere.  Host-only calls."""
    L = capi.lib()
    dummy = ctypes.c_void_p(16)  # never dereferenced: every call below fails validation or is empty
    p = capi.problem(3, 2, 1025, 0.0, 1.0, op="sum", matrix_backend="SpaceCurve", integrator="Simpson")
    # one instance: 16-sample chunks -> 64 chunks of 5 4-vectors
    assert L.magnus_sc_checkpoint_elems(ctypes.byref(p), 1) == 64 * 5 * 4
    # headline batch shape (4096 x 16385 samples, n = 6): 512-sample chunks, 46 MB
    ph = capi.problem(6, 2, 16385, 0.0, 1.0, op="sum", matrix_backend="SpaceCurve", integrator="Boole")
    assert L.magnus_sc_checkpoint_elems(ctypes.byref(ph), 4096) * 8 == 4096 * 32 * 11 * 4 * 8
    for n, integ in [(3, "Trapezoid"), (1, "Simpson"), (9, "Boole")]:
        q = capi.problem(n, 2, 1025, 0.0, 1.0, matrix_backend="SpaceCurve", integrator=integ)
        assert L.magnus_sc_checkpoint_elems(ctypes.byref(q), 2) == 0
        assert L.magnus_sc_forward_ckpt_f64(ctypes.byref(q), 2, dummy, dummy, dummy, None) == capi.EINVAL
        assert L.magnus_sc_vjp_ckpt_f64(ctypes.byref(q), 2, dummy, dummy, dummy, dummy, None) == capi.EINVAL
    assert L.magnus_sc_forward_ckpt_f64(ctypes.byref(p), 2, dummy, dummy, None, None) == capi.EINVAL
    assert L.magnus_sc_vjp_ckpt_f64(ctypes.byref(p), 2, dummy, None, dummy, dummy, None) == capi.EINVAL
    assert L.magnus_sc_forward_ckpt_f64(ctypes.byref(p), 0, None, None, None, None) == capi.OK  # empty batch


def test_controller_and_uneven_argument_checks_before_device_work():
    """New entry points reject bad arguments before any device call."""
    L = capi.lib()
    dummy = ctypes.c_void_p(16)  # never dereferenced
    # features: dim 1..16, grid 2^k + 1 in 9..129
    assert L.magnus_ctrl_features_f64(1, 65, 0, 0.0, 1.0, dummy, dummy, None) == capi.EINVAL
    assert L.magnus_ctrl_features_f64(1, 65, 17, 0.0, 1.0, dummy, dummy, None) == capi.EINVAL
    assert L.magnus_ctrl_features_f64(1, 66, 8, 0.0, 1.0, dummy, dummy, None) == capi.EINVAL
    assert L.magnus_ctrl_features_f64(1, 257, 8, 0.0, 1.0, dummy, dummy, None) == capi.EINVAL
    assert L.magnus_ctrl_features_f64(1, 65, 8, 0.0, 1.0, None, dummy, None) == capi.EINVAL
    assert L.magnus_ctrl_features_f64(0, 65, 8, 0.0, 1.0, dummy, dummy, None) == capi.OK
    assert L.magnus_ctrl_feature_count() == 24
    # MLP creation: positive widths and batch
    h = ctypes.c_void_p()
    assert L.magnus_mlp_create(0, 128, 63, 1024, 0, ctypes.byref(h)) == capi.EINVAL
    assert L.magnus_mlp_create(24, 128, 63, 0, 0, ctypes.byref(h)) == capi.EINVAL
    assert L.magnus_mlp_create(24, 5000, 63, 1024, 0, ctypes.byref(h)) == capi.EINVAL
    assert L.magnus_mlp_param_count(None) == 0
    # all-orders forward: matrix problems only
    sc = capi.problem(3, 2, 33, 0.0, 1.0, matrix_backend="SpaceCurve")
    assert L.magnus_forward_orders_f64(ctypes.byref(sc), 1, dummy, dummy, None) == capi.EINVAL
    p = capi.problem(3, 4, 33, 0.0, 1.0)
    assert L.magnus_forward_orders_f64(ctypes.byref(p), 0, None, None, None) == capi.OK
    assert L.magnus_forward_orders_f64(ctypes.byref(p), 1, None, dummy, None) == capi.EINVAL
    # uneven plans: branch count, then per-instance validation naming the instance
    probs = (capi.Problem * 2)(capi.problem(2, 4, 33, 0.0, 1.0),
                               capi.problem(2, 4, 32, 0.0, 1.0, integrator="Boole"))
    plan = ctypes.c_void_p()
    assert L.magnus_uneven_create(probs, 2, 0, 0, ctypes.byref(plan)) == capi.EINVAL
    assert L.magnus_uneven_create(probs, 2, 0, 9, ctypes.byref(plan)) == capi.EINVAL
    assert L.magnus_uneven_create(probs, 2, 0, 2, ctypes.byref(plan)) == capi.EINVAL
    assert L.magnus_last_error().decode().startswith("instance 1: ")
    assert L.magnus_uneven_info(None, 0) == 0


def test_expm_argument_checks_before_device_work():
    L = capi.lib()
    dummy = ctypes.c_void_p(16)  # never dereferenced: every call below fails validation or is empty
    assert L.magnus_expm_f64(2, 3, None, dummy, None) == capi.EINVAL
    assert L.magnus_expm_f64(2, 0, dummy, dummy, None) == capi.EINVAL
    assert L.magnus_sc_expm_f64(2, dummy, None, None) == capi.EINVAL
    assert L.magnus_expm_f64(0, 3, dummy, dummy, None) == capi.OK


def test_default_gl_table_is_numpy_leggauss_bitwise():
    L = capi.lib()
    for order in (1, 2, 3, 5, 17, 64):
        w = (ctypes.c_double * order)()
        x = (ctypes.c_double * order)()
        capi.check(L.magnus_gl_order(order, w, x))
        xr, wr = np.polynomial.legendre.leggauss(order)
        np.testing.assert_array_equal(np.array(x[:]), 0.5 * (xr + 1.0))
        np.testing.assert_array_equal(np.array(w[:]), 0.5 * wr)


def test_error_contract_through_public_api_without_device():
    """These all fail during validation, before any device allocation."""
    import magnus

    def zeros3(t):
        return np.zeros((np.asarray(t).size, 3, 3))

    with pytest.raises(ValueError, match="invalid dispatch option"):
        magnus.compute(1, zeros3, 0.0, 1.0, 9, op="one", matrix_backend="Fixed2")
    with pytest.raises(ValueError, match="dispatch name"):
        magnus.compute(1, zeros3, 0.0, 1.0, 9, op="one", matrix_backend="Missing")
    with pytest.raises(ValueError, match="dispatch name"):
        magnus.compute(1, zeros3, 0.0, 1.0, 9, op="one", integrator="Missing")
    with pytest.raises(ValueError, match="dispatch name"):
        magnus.compute(1, zeros3, 0.0, 1.0, 9, op="one", gl_backend="Missing")
    with pytest.raises(ValueError, match="deduce kernel op"):
        magnus.compute(1, zeros3, 0.0, 1.0, 9, op="missing")
    with pytest.raises(ValueError, match="at least two samples"):
        magnus.compute(1, matrix_values, 0.0, 1.0, 1)
    with pytest.raises(ValueError, match="at least two samples"):
        magnus.compute_sc(1, spacecurve_values, 0.0, 1.0, 1)
    with pytest.raises(ValueError, match="n must be at least 1"):
        magnus.compute(0, matrix_values, 0.0, 1.0, 9)
    with pytest.raises(ValueError, match="n must be at least 1"):
        magnus.compute_sc(0, spacecurve_values, 0.0, 1.0, 9)
    with pytest.raises(TypeError, match="supports dtypes"):
        magnus.compute(1, lambda t: matrix_values(t).astype(np.int64), 0.0, 1.0, 9)
    with pytest.raises(TypeError, match="supports dtypes"):
        magnus.compute_sc(1, lambda t: spacecurve_values(t).astype(np.int64), 0.0, 1.0, 9)
    with pytest.raises(ValueError, match="shape"):
        magnus.compute_sc(1, lambda t: np.zeros((np.asarray(t).size, 4)), 0.0, 1.0, 9)
    for bad in (np.zeros((9, 2)), np.zeros((8, 2, 2)), np.zeros((9, 2, 3))):
        with pytest.raises(ValueError):
            magnus.compute(1, lambda t, b=bad: b, 0.0, 1.0, 9)
    for bad in (np.zeros((9, 2)), np.zeros((8, 3)), np.zeros((9, 3, 1))):
        with pytest.raises(ValueError):
            magnus.compute_sc(1, lambda t, b=bad: b, 0.0, 1.0, 9)
    with pytest.raises(ValueError, match="divisibility"):
        magnus.compute(2, matrix_values, 0.0, 1.0, 7, integrator="Boole")
    with pytest.raises(ValueError, match="one/sum output must have shape"):
        magnus.compute_vjp(2, matrix_values, np.zeros((3, 3)), 0.0, 1.0, 9)
    with pytest.raises(ValueError, match="VJP data has an invalid shape"):
        magnus.compute_vjp(2, matrix_values, np.zeros((2, 2)), 0.0, 1.0, 9, vjp_data=np.zeros((3, 1, 9, 2, 2)))


def test_one_hip_runtime_per_process():
    """Importing the package maps exactly one HIP and one HSA runtime, so device
    pointers and streams from torch and from the library are interchangeable."""
    import subprocess
    import sys

    code = (
        "import sys; sys.path.insert(0, 'magnus-dp_amd'); import magnus; from magnus import capi; capi.lib();"
        "m = open('/proc/self/maps').read().splitlines();"
        "print(len({l.split()[-1] for l in m if 'libamdhip64' in l}), len({l.split()[-1] for l in m if 'libhsa-runtime64' in l}))"
    )
    from pathlib import Path

    out = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, check=True,
                         cwd=Path(__file__).resolve().parents[1]).stdout.split()
    assert out == ["1", "1"], out
