"""Pins for the CPU oracle (oracle/magnus_oracle.c).

The reference cannot be built or imported in this image and its own tests hold
no golden vectors, so the oracle is pinned by what the mathematics fixes:
closed forms, exactness of the quadrature rules, adjoint identities, a known
Omega_2, the su(2) <-> complex 2x2 correspondence, finite differences and the
ODE propagator.  Each check is chosen so that a dropped term, a wrong
coefficient or index, or a transposed operand in the restatement fails it.
"""
from __future__ import annotations

import numpy as np
import pytest
import scipy.integrate
import scipy.linalg

import oracle as O
from helpers import (
    dense_matrix_values,
    fourier_matrices,
    fourier_spacecurve,
    gradient_values,
    matrix_values,
    rel_fro,
    spacecurve_values,
)

INTEGRATORS = ["Boole", "Simpson", "Trapezoid", "Riemann"]


# --------------------------------------------------------------------------
# Gauss-Legendre table (util/gen_nodes.py:11-47)
# --------------------------------------------------------------------------
def _gl_truth(N):
    """40-digit Gauss-Legendre nodes/weights on [0,1] (mpmath Newton on P_N)."""
    import mpmath

    mpmath.mp.dps = 40
    xs, ws = [], []
    for i in range(1, N + 1):
        x = mpmath.cos(mpmath.pi * (i - mpmath.mpf(0.25)) / (N + mpmath.mpf(0.5)))
        for _ in range(100):
            p = mpmath.legendre(N, x)
            dp = N * (x * p - mpmath.legendre(N - 1, x)) / (x * x - 1)
            dx = p / dp
            x -= dx
            if abs(dx) < mpmath.mpf(10) ** -35:
                break
        dp = N * (x * mpmath.legendre(N, x) - mpmath.legendre(N - 1, x)) / (x * x - 1)
        xs.append((x + 1) / 2)
        ws.append(1 / ((1 - x * x) * dp * dp))
    order = np.argsort([float(v) for v in xs])
    return np.array([float(ws[j]) for j in order]), np.array([float(xs[j]) for j in order])


@pytest.mark.parametrize("order", list(range(1, 17)) + [22, 32, 40, 64])
def test_gl_matches_high_precision_and_numpy(order):
    w, x = O.gl(order)
    wt, xt = _gl_truth(order)
    np.testing.assert_allclose(x, xt, rtol=0, atol=2.3e-16)
    np.testing.assert_allclose(w, wt, rtol=5e-14, atol=0)
    # numpy leggauss (what the reference embeds, util/gen_nodes.py:21-24) agrees to
    # the accuracy of numpy's weights, whose error grows like order^2 (~1e-12 at 64).
    xr, wr = np.polynomial.legendre.leggauss(order)
    np.testing.assert_allclose(x, (xr + 1) / 2, rtol=0, atol=2.3e-16)
    np.testing.assert_allclose(w, wr / 2, rtol=max(5e-15, 5e-16 * order**2), atol=0)
    assert abs(w.sum() - 1.0) < 1e-14
    # exactness on monomials t^j, j <= 2*order-1, over [0,1]
    for j in range(0, 2 * order):
        assert abs(np.dot(w, x**j) - 1.0 / (j + 1)) < 1e-13


# --------------------------------------------------------------------------
# Integrator exactness on polynomials
# --------------------------------------------------------------------------
def _poly(t, deg, rng):
    c = rng.standard_normal(deg + 1)
    val = np.polynomial.polynomial.polyval(t, c)
    anti = np.polynomial.polynomial.polyint(c)
    return val, lambda a: np.polynomial.polynomial.polyval(a, anti) - np.polynomial.polynomial.polyval(0.0, anti)


@pytest.mark.parametrize(
    "integrator,deg,samples",
    [("Boole", 5, 17), ("Boole", 5, 33), ("Simpson", 3, 9), ("Simpson", 3, 33), ("Trapezoid", 1, 9)],
)
def test_sum_is_exact_on_polynomials(integrator, deg, samples):
    rng = np.random.default_rng(deg * 100 + samples)
    t = np.linspace(0.0, 1.3, samples)
    dt = t[1] - t[0]
    f, F = _poly(t, deg, rng)
    A = np.stack([f, 2 * f, -f, 0.5 * f], axis=1)
    got = O.isum(integrator, A, dt)
    np.testing.assert_allclose(got, F(1.3) * np.array([1, 2, -1, 0.5]), rtol=1e-13, atol=1e-13)
    # one degree more is NOT exact (guards against a rule that is "too good" by accident)
    f2, F2 = _poly(t, deg + 1 if integrator != "Boole" else deg + 2, rng)
    got2 = O.isum(integrator, f2[:, None], dt)
    assert abs(got2[0] - F2(1.3)) > 1e-12


def test_riemann_sum_closed_form():
    rng = np.random.default_rng(5)
    A = rng.standard_normal((11, 3))
    dt = 0.37
    np.testing.assert_allclose(O.isum("Riemann", A, dt), dt * A.sum(axis=0), rtol=1e-14)


def test_boole_prefix_local_rules():
    # prefix at +1 is trapezoid (exact for linear), +2 Simpson, +3 3/8 (exact cubic), +4 Boole (exact quintic)
    rng = np.random.default_rng(11)
    t = np.linspace(0.0, 2.0, 21)
    dt = t[1] - t[0]
    f1, F1 = _poly(t, 1, rng)
    P = O.prefix("Boole", f1[:, None], dt)[:, 0]
    np.testing.assert_allclose(P, F1(t), rtol=1e-13, atol=1e-13)
    f3, F3 = _poly(t, 3, rng)
    P = O.prefix("Boole", f3[:, None], dt)[:, 0]
    idx = np.arange(21)
    exact = idx % 4 != 1
    np.testing.assert_allclose(P[exact], F3(t[exact]), rtol=1e-12, atol=1e-12)
    assert np.max(np.abs(P[~exact] - F3(t[~exact]))) > 1e-10
    f5, F5 = _poly(t, 5, rng)
    P = O.prefix("Boole", f5[:, None], dt)[:, 0]
    np.testing.assert_allclose(P[idx % 4 == 0], F5(t[idx % 4 == 0]), rtol=1e-12, atol=1e-12)


def test_simpson_prefix_local_rules():
    rng = np.random.default_rng(12)
    t = np.linspace(0.0, 2.0, 21)
    dt = t[1] - t[0]
    f3, F3 = _poly(t, 3, rng)
    P = O.prefix("Simpson", f3[:, None], dt)[:, 0]
    even = np.arange(21) % 2 == 0
    np.testing.assert_allclose(P[even], F3(t[even]), rtol=1e-12, atol=1e-12)
    f1, F1 = _poly(t, 1, rng)
    np.testing.assert_allclose(O.prefix("Simpson", f1[:, None], dt)[:, 0], F1(t), rtol=1e-13, atol=1e-13)


def test_trapezoid_and_riemann_prefix():
    rng = np.random.default_rng(13)
    t = np.linspace(0.0, 1.0, 17)
    dt = t[1] - t[0]
    f1, F1 = _poly(t, 1, rng)
    np.testing.assert_allclose(O.prefix("Trapezoid", f1[:, None], dt)[:, 0], F1(t), rtol=1e-12, atol=1e-12)
    f = rng.standard_normal(17)
    np.testing.assert_allclose(O.prefix("Riemann", f[:, None], dt)[:, 0], dt * np.cumsum(f), rtol=1e-14)


# --------------------------------------------------------------------------
# Adjoint identities: <L x, y> == <x, L^T y>
# --------------------------------------------------------------------------
@pytest.mark.parametrize("integrator", INTEGRATORS)
@pytest.mark.parametrize("samples", [5, 9, 13, 33])
def test_prefix_vjp_is_adjoint(integrator, samples):
    if integrator == "Boole" and (samples - 1) % 4:
        pytest.skip("Boole needs (S-1) % 4 == 0")
    rng = np.random.default_rng(samples)
    m = 3
    x = rng.standard_normal((samples, m))
    y = rng.standard_normal((samples, m))
    dt = 0.113
    lhs = np.sum(O.prefix(integrator, x, dt) * y)
    rhs = np.sum(x * O.prefix_vjp(integrator, y, dt))
    assert abs(lhs - rhs) <= 1e-13 * max(1.0, abs(lhs))


@pytest.mark.parametrize("integrator", INTEGRATORS)
def test_sum_vjp_is_adjoint_and_add_is_consistent(integrator):
    rng = np.random.default_rng(7)
    samples, m, dt = 17, 4, 0.21
    x = rng.standard_normal((samples, m))
    y = rng.standard_normal(m)
    lhs = np.dot(O.isum(integrator, x, dt), y)
    sv = O.sum_vjp(integrator, y, samples, dt)
    assert abs(lhs - np.sum(x * sv)) <= 1e-13 * max(1.0, abs(lhs))
    base = rng.standard_normal((samples, m))
    np.testing.assert_allclose(O.sum_vjp_add(integrator, base, y, dt), base + sv, rtol=1e-14, atol=1e-15)


@pytest.mark.parametrize("matrix,dim", [("Fixed2", 2), ("Manual", 2), ("Manual", 3), ("Blas", 5), ("SpaceCurve", 2)])
def test_sample_update_vjp_is_adjoint(matrix, dim):
    rng = np.random.default_rng(dim + len(matrix))
    S = 7
    m = 4 if matrix == "SpaceCurve" else dim * dim
    shape = (S, m) if matrix == "SpaceCurve" else (S, dim, dim)
    A = rng.standard_normal(shape)
    if matrix == "SpaceCurve":
        A[:, 0] = 0.0  # padded input has component 0 == 0
    P = rng.standard_normal(shape)
    G = rng.standard_normal(shape)
    shift = -0.37
    total = P[-1].copy()
    dA0 = rng.standard_normal(shape)
    dA, barP = O.sample_update_vjp(matrix, dA0, A, G, P, shift)
    # linear in P (total = P[last] is part of the map)
    dP = rng.standard_normal(shape)
    Yp = O.sample_update(matrix, A, dP, dP[-1], shift)
    assert abs(np.sum(Yp * G) - np.sum(dP * barP)) < 1e-12 * max(1, abs(np.sum(Yp * G)))
    # linear in A for fixed P
    dAdir = rng.standard_normal(shape)
    if matrix == "SpaceCurve":
        dAdir[:, 0] = 0.0
    Ya = O.sample_update(matrix, dAdir, P, total, shift)
    assert abs(np.sum(Ya * G) - np.sum(dAdir * (dA - dA0))) < 1e-12 * max(1, abs(np.sum(Ya * G)))


# --------------------------------------------------------------------------
# Forward: known answers and consistency
# --------------------------------------------------------------------------
def _comm(a, b):
    return a @ b - b @ a


@pytest.mark.parametrize("integrator", ["Boole", "Simpson"])
@pytest.mark.parametrize("T", [1.0, 0.7])
def test_omega2_closed_form_linear_generator(integrator, T):
    rng = np.random.default_rng(3)
    d = 3
    A0 = rng.standard_normal((d, d))
    A1 = rng.standard_normal((d, d))
    S = 33
    t = np.linspace(0.0, T, S)
    data = A0[None] + t[:, None, None] * A1[None]
    many = O.forward(2, data, 0.0, T, op="many", integrator=integrator)
    np.testing.assert_allclose(many[0], T * A0 + T * T / 2 * A1, rtol=1e-13, atol=1e-13)
    expected = -(T**3 / 12.0) * _comm(A0, A1)
    np.testing.assert_allclose(many[1], expected, rtol=1e-12, atol=1e-13)


def test_commuting_generator_has_no_higher_terms():
    # A(t) = f(t) M: every product is a power of M, so Omega_k = alpha_k M^k with
    # alpha_k -> 0 under refinement (the discrete rules do not integrate by parts exactly).
    rng = np.random.default_rng(4)
    M = rng.standard_normal((3, 3))
    alphas = []
    for S in (65, 257):
        t = np.linspace(0.0, 1.0, S)
        data = (np.sin(3 * t) + 0.5)[:, None, None] * M[None]
        many = O.forward(5, data, 0.0, 1.0, op="many", integrator="Boole")
        row = []
        for k in range(1, 5):
            Mk = np.linalg.matrix_power(M, k + 1)
            a = np.sum(many[k] * Mk) / np.sum(Mk * Mk)
            assert np.linalg.norm(many[k] - a * Mk) < 1e-13 * max(1.0, np.linalg.norm(many[k]))
            row.append(abs(a))
        alphas.append(row)
    for k in range(4):
        assert alphas[1][k] < alphas[0][k] / 10
    const = np.broadcast_to(M, (9, 3, 3)).copy()
    many = O.forward(2, const, 0.0, 1.0, op="many", integrator="Boole")
    assert np.max(np.abs(many[1])) < 1e-15


@pytest.mark.parametrize("integrator", INTEGRATORS)
def test_one_many_sum_consistency(integrator):
    data = matrix_values(np.linspace(0, 1, 9))
    one = O.forward(4, data, 0.0, 1.0, op="one", integrator=integrator)
    many = O.forward(4, data, 0.0, 1.0, op="many", integrator=integrator)
    total = O.forward(4, data, 0.0, 1.0, op="sum", integrator=integrator)
    np.testing.assert_allclose(one, many[3], rtol=1e-11, atol=1e-12)
    np.testing.assert_allclose(total, many.sum(axis=0), rtol=1e-11, atol=1e-12)


def test_gl_lanes_match_serial():
    data = matrix_values(np.linspace(0, 1, 1025))
    serial, c1 = O.forward(8, data, 0.0, 1.0, integrator="Boole", lanes=1, record_vjp=True)
    par, c2 = O.forward(8, data, 0.0, 1.0, integrator="Boole", lanes=5, record_vjp=True)
    np.testing.assert_allclose(par, serial, rtol=1e-13, atol=1e-14)
    np.testing.assert_array_equal(c1, c2)


def test_propagator_converges_to_ode_solution():
    """exp(sum_k Omega_k) approaches the ODE propagator of Y' = A Y as n grows."""
    d, S = 3, 513
    data = 0.4 * fourier_matrices(np.random.default_rng(21), S, d, modes=2)
    coeffs = np.random.default_rng(21).standard_normal((5, d, d)) / np.sqrt(d) * 0.4

    def A_of(t):
        out = coeffs[0].copy()
        for mth in (1, 2):
            out += np.sin(2 * np.pi * mth * t) * coeffs[2 * mth - 1] + np.cos(2 * np.pi * mth * t) * coeffs[2 * mth]
        return out

    np.testing.assert_allclose(A_of(np.linspace(0, 1, S)[7]), data[7], rtol=1e-13, atol=1e-14)
    sol = scipy.integrate.solve_ivp(
        lambda t, y: (A_of(t) @ y.reshape(d, d)).ravel(), (0.0, 1.0), np.eye(d).ravel(),
        rtol=1e-13, atol=1e-14, method="DOP853",
    )
    Y = sol.y[:, -1].reshape(d, d)
    errs = [np.max(np.abs(scipy.linalg.expm(O.forward(n, data, 0.0, 1.0, op="sum", integrator="Boole")) - Y))
            for n in range(1, 9)]
    assert all(b < a for a, b in zip(errs, errs[1:])), errs
    assert errs[1] < 2e-2 and errs[3] < 5e-4 and errs[5] < 1e-5 and errs[7] < 1e-6, errs


# --------------------------------------------------------------------------
# SpaceCurve == dense complex 2x2 product form with A(t) = -i a(t).sigma
# --------------------------------------------------------------------------
_SIG = np.array([[[0, 1], [1, 0]], [[0, -1j], [1j, 0]], [[1, 0], [0, -1]]], dtype=np.complex128)


def _M(c):
    c = np.asarray(c)
    return -c[..., 0, None, None] * np.eye(2) - 1j * np.einsum("...j,jab->...ab", c[..., 1:], _SIG)


def _complex_magnus_many(n, a3, t0, tf, integrator):
    """Product-form Magnus on complex 2x2 matrices (small, python loop over q, k)."""
    S = a3.shape[0]
    dt = (tf - t0) / (S - 1)
    A = _M(np.concatenate([np.zeros((S, 1)), a3], axis=1))

    def pack(Z):
        return np.concatenate([Z.real.reshape(S, 4), Z.imag.reshape(S, 4)], axis=1)

    def unpack(R):
        return (R[:, :4] + 1j * R[:, 4:]).reshape(S, 2, 2)

    out = np.zeros((n, 2, 2), dtype=np.complex128)
    s1 = O.isum(integrator, pack(A), dt)
    out[0] = (s1[:4] + 1j * s1[4:]).reshape(2, 2)
    if n == 1:
        return out
    Q = (n + 3) // 2
    w, x = O.gl(Q)
    for q in range(Q):
        Y = A.copy()
        for k in range(2, n + 1):
            P = unpack(O.prefix(integrator, pack(Y), dt))
            T = P[-1]
            Y = A @ (P + (x[q] - 1) * T)
            sk = O.isum(integrator, pack(Y), dt * w[q])
            out[k - 1] += (sk[:4] + 1j * sk[4:]).reshape(2, 2)
    return out


@pytest.mark.parametrize("integrator", ["Boole", "Simpson", "Riemann"])
def test_spacecurve_equals_complex_2x2(integrator):
    rng = np.random.default_rng(9)
    S, n = 33, 5
    a3 = fourier_spacecurve(rng, S)
    got = O.sc_forward(n, a3, 0.0, 1.0, op="many", integrator=integrator)
    Z = _complex_magnus_many(n, a3, 0.0, 1.0, integrator)
    # c_j = (i/2) tr(Z sigma_j)
    c = 0.5j * np.einsum("kab,jba->kj", Z, _SIG)
    np.testing.assert_allclose(c.imag, 0.0, atol=1e-13)
    np.testing.assert_allclose(got, c.real, rtol=1e-12, atol=1e-13)


def test_spacecurve_fixed_axis_has_no_higher_terms():
    # a(t) = f(t) n: all products commute, so every Omega_k (k >= 2) lies along n
    # (odd k) or in the dropped scalar part (even k), and shrinks under refinement.
    axis = np.array([0.3, -0.5, 0.8])
    mags = []
    for S in (65, 257):
        t = np.linspace(0, 1, S)
        a3 = (np.cos(2 * t) + 0.2)[:, None] * axis[None]
        many = O.sc_forward(6, a3, 0.0, 1.0, op="many", integrator="Boole")
        for k in range(1, 6):
            assert np.linalg.norm(np.cross(many[k], axis)) < 1e-13 * np.linalg.norm(many[k]) + 1e-16
        for k in (1, 3, 5):
            assert np.linalg.norm(many[k]) < 1e-15
        mags.append([np.linalg.norm(many[k]) for k in (2, 4)])
    assert mags[1][0] < mags[0][0] / 10 and mags[1][1] < mags[0][1] / 10


# --------------------------------------------------------------------------
# VJP: finite differences and carry == recompute
# (pattern of the reference tests/test_python_api.py:64-147)
# --------------------------------------------------------------------------
def _fd_check(data, n, op, matrix_backend, integrator, sc=False, eps=1e-6):
    direction = gradient_values(data.shape, (-0.3, 0.7))
    if sc:
        out, carry = O.sc_forward(n, data, 0.0, 1.0, op=op, integrator=integrator, record_vjp=True)
    else:
        out, carry = O.forward(n, data, 0.0, 1.0, op=op, matrix_backend=matrix_backend, integrator=integrator,
                               record_vjp=True)
    cot = gradient_values(out.shape, (0.2, 1.1))
    if sc:
        g = O.sc_vjp(n, data, cot, op=op, integrator=integrator, vjp_data=carry)
        g2 = O.sc_vjp(n, data, cot, op=op, integrator=integrator)
        plus = O.sc_forward(n, data + eps * direction, 0.0, 1.0, op=op, integrator=integrator)
        minus = O.sc_forward(n, data - eps * direction, 0.0, 1.0, op=op, integrator=integrator)
    else:
        kw = dict(op=op, matrix_backend=matrix_backend, integrator=integrator)
        g = O.vjp(n, data, cot, vjp_data=carry, **kw)
        g2 = O.vjp(n, data, cot, **kw)
        plus = O.forward(n, data + eps * direction, 0.0, 1.0, **kw)
        minus = O.forward(n, data - eps * direction, 0.0, 1.0, **kw)
    expected = np.sum((plus - minus) * cot) / (2 * eps)
    np.testing.assert_allclose(np.sum(g * direction), expected, rtol=1e-7, atol=1e-9)
    np.testing.assert_allclose(g2, g, rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("op", ["one", "many", "sum"])
@pytest.mark.parametrize("integrator", INTEGRATORS)
def test_vjp_matches_finite_differences(op, integrator):
    _fd_check(matrix_values(np.linspace(0, 1, 9)), 4, op, "Fixed2", integrator)


@pytest.mark.parametrize("op", ["one", "many", "sum"])
@pytest.mark.parametrize("integrator", INTEGRATORS)
def test_sc_vjp_matches_finite_differences(op, integrator):
    _fd_check(spacecurve_values(np.linspace(0, 1, 9)), 4, op, None, integrator, sc=True)


@pytest.mark.parametrize("matrix_backend,dim", [("Fixed2", 2), ("Manual", 3), ("Blas", 3)])
def test_vjp_matches_fd_per_backend(matrix_backend, dim):
    _fd_check(dense_matrix_values(9, dim), 4, "many", matrix_backend, "Boole")


@pytest.mark.parametrize("n,samples", [(1, 5), (2, 9), (3, 13), (6, 17)])
def test_vjp_across_orders(n, samples):
    _fd_check(matrix_values(np.linspace(0, 1, samples)), n, "sum", "Fixed2", "Boole")
    _fd_check(spacecurve_values(np.linspace(0, 1, samples)), n, "sum", None, "Boole", sc=True)


def test_errors():
    data = np.zeros((9, 3, 3))
    with pytest.raises(ValueError, match="invalid dispatch option"):
        O.forward(1, data, 0.0, 1.0, op="one", matrix_backend="Fixed2")
    with pytest.raises(ValueError, match="divisibility"):
        O.forward(2, np.zeros((7, 2, 2)), 0.0, 1.0, integrator="Boole")
    with pytest.raises(ValueError, match="at least two samples"):
        O.forward(1, np.zeros((1, 2, 2)), 0.0, 1.0)
