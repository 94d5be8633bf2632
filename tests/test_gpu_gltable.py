"""Runtime replacement of the Gauss-Legendre table (reference:
src/api/pybind.cpp:102-116 `_replace_gl_table`, src/util/gausslegendre.cpp:74-111
OwningTable, src/api/magnus/__init__.py:41-55 `replace_gl_table`).

The table is process-global; every test here restores numpy's 64-order table
(the library's default, bit for bit) in both the product and the oracle.
"""
from __future__ import annotations

import numpy as np
import pytest

import oracle as O
from helpers import dense_matrix_values, fourier_spacecurve, gradient_values, rel_fro

pytestmark = pytest.mark.gpu

TOL = 1e-12
ORDER_ERROR = "requested magnus order exceeds GL table. Consider providing a larger table."


def _table(max_order, rule):
    """Flattened (weights, nodes) for orders 1..max_order (order o at o(o-1)/2)."""
    total = max_order * (max_order + 1) // 2
    w, x = np.empty(total), np.empty(total)
    off = 0
    for order in range(1, max_order + 1):
        wo, xo = rule(order)
        w[off:off + order] = wo
        x[off:off + order] = xo
        off += order
    return w, x


def _leggauss(order):
    xr, wr = np.polynomial.legendre.leggauss(order)
    return wr / 2, (xr + 1) / 2


def _midpoint(order):
    """A deliberately non-Gauss rule: midpoints of `order` equal cells, weights
    1/order (any table is accepted, as in the reference's OwningTable)."""
    return np.full(order, 1.0 / order), (np.arange(order) + 0.5) / order


@pytest.fixture(autouse=True)
def restore_tables():
    yield
    import magnus

    magnus.replace_gl_table(64)
    w, x = _table(64, _leggauss)
    O.gl_replace(64, w, x)


def test_replace_changes_max_order_and_bounds_n():
    """replace_gl_table(k) sets max_order() = k; n with (n + 3) // 2 <= k runs and
    n = 2k - 1, the first order needing k + 1 nodes, raises the reference's text."""
    import magnus

    k = 5
    magnus.replace_gl_table(k)
    assert magnus.max_order() == k
    data = dense_matrix_values(17, 3)
    w, x = _table(k, _leggauss)
    O.gl_replace(k, w, x)
    n_ok = 2 * k - 2
    got = magnus._core.compute(n_ok, data, 0.0, 1.0, op="sum", integrator="Boole")
    ref = O.forward(n_ok, data, 0.0, 1.0, op="sum", integrator="Boole")
    assert rel_fro(got, ref) <= TOL
    for call in (lambda: magnus._core.compute(2 * k - 1, data, 0.0, 1.0),
                 lambda: magnus._core.compute_vjp(2 * k - 1, data, np.ones((3, 3))),
                 lambda: magnus._core.compute_sc(2 * k - 1, dense_matrix_values(17, 3)[:, 0, :], 0.0, 1.0)):
        with pytest.raises(ValueError, match="exceeds GL table"):
            call()
    magnus.replace_gl_table(64)
    assert magnus.max_order() == 64


@pytest.mark.parametrize("op", ["one", "many", "sum"])
def test_non_gauss_table_matches_oracle(op):
    """A midpoint-rule table installed in both the product and the oracle: the
    forward (carry and carry-free paths), the VJP and the SpaceCurve path agree
    with the oracle run on the same table."""
    import magnus
    from magnus import _core

    max_order = 6
    w, x = _table(max_order, _midpoint)
    _core._replace_gl_table(max_order, w, x)
    O.gl_replace(max_order, w, x)
    assert magnus.max_order() == max_order
    n = 8  # Q = 5 nodes from the midpoint table
    data = dense_matrix_values(33, 4)
    kw = dict(op=op, integrator="Boole")
    out, carry = _core.compute(n, data, 0.0, 1.0, record_vjp=True, **kw)
    ref, ref_carry = O.forward(n, data, 0.0, 1.0, record_vjp=True, **kw)
    assert rel_fro(out, ref) <= TOL and rel_fro(carry, ref_carry) <= TOL
    assert rel_fro(_core.compute(n, data, 0.0, 1.0, **kw), ref) <= TOL
    cot = gradient_values(out.shape, (0.2, 1.1))
    ref_g = O.vjp(n, data, cot, **kw)
    assert rel_fro(_core.compute_vjp(n, data, cot, vjp_data=carry, **kw), ref_g) <= TOL
    assert rel_fro(_core.compute_vjp(n, data, cot, **kw), ref_g) <= TOL
    # the table really differs from Gauss-Legendre: results move well beyond the bar
    lg_w, lg_x = _table(max_order, _leggauss)
    O.gl_replace(max_order, lg_w, lg_x)
    assert rel_fro(O.forward(n, data, 0.0, 1.0, **kw), ref) > 1e-6
    O.gl_replace(max_order, w, x)
    a = fourier_spacecurve(np.random.default_rng([31]), 257)
    got = _core.compute_sc(n, a, 0.0, 1.0, op=op, integrator="Simpson")
    ref_sc = O.sc_forward(n, a, 0.0, 1.0, op=op, integrator="Simpson")
    assert rel_fro(got, ref_sc) <= TOL


def test_order_126_uses_all_64_nodes():
    """n = 126 is the largest order of the default 64-order table (Q = (n + 3) // 2
    = 64 nodes): the op='sum' forward and the VJP, with and without the carry,
    against the oracle on numpy's table; n = 127 needs 65 nodes and raises."""
    from magnus import _core

    w, x = _table(64, _leggauss)
    O.gl_replace(64, w, x)
    n, S = 126, 9
    data = 0.3 * dense_matrix_values(S, 2)
    kw = dict(op="sum", integrator="Boole")
    out, carry = _core.compute(n, data, 0.0, 1.0, record_vjp=True, **kw)
    ref, ref_carry = O.forward(n, data, 0.0, 1.0, record_vjp=True, **kw)
    assert carry.shape[0] == 64
    assert rel_fro(out, ref) <= TOL and rel_fro(carry, ref_carry) <= TOL
    assert rel_fro(_core.compute(n, data, 0.0, 1.0, **kw), ref) <= TOL
    cot = gradient_values(out.shape, (0.2, 1.1))
    ref_g = O.vjp(n, data, cot, **kw)
    assert rel_fro(_core.compute_vjp(n, data, cot, vjp_data=carry, **kw), ref_g) <= TOL
    assert rel_fro(_core.compute_vjp(n, data, cot, **kw), ref_g) <= TOL
    with pytest.raises(ValueError, match="exceeds GL table"):
        _core.compute(127, data, 0.0, 1.0, **kw)
