"""Shared input generators for the tests.

The small closed-form generators mirror the shapes and value ranges the
reference tests use (tests/test_python_api.py:15-59); the Fourier-form
generators produce the seeded workloads named in the measurement plan
(A(t) = A0 + sum_m B_m sin(2 pi m t) + C_m cos(2 pi m t)).
"""
from __future__ import annotations

import numpy as np


def matrix_values(t, dtype=np.float64):
    t = np.asarray(t, dtype=np.float64)
    out = np.empty((t.size, 2, 2), dtype=dtype)
    out[:, 0, 0] = 1.0 + t
    out[:, 0, 1] = t * t
    out[:, 1, 0] = -0.5 * t
    out[:, 1, 1] = 2.0 - t
    return out


def spacecurve_values(t, dtype=np.float64):
    t = np.asarray(t, dtype=np.float64)
    out = np.empty((t.size, 3), dtype=dtype)
    out[:, 0] = np.sin(1.3 * t) + 0.2 * t
    out[:, 1] = np.cos(0.7 * t) - 0.1
    out[:, 2] = 0.4 + t * t
    return out


def dense_matrix_values(samples, dim):
    t = np.linspace(0.0, 1.0, samples)[:, None, None]
    row = np.arange(1, dim + 1, dtype=np.float64)[None, :, None]
    col = np.arange(1, dim + 1, dtype=np.float64)[None, None, :]
    return np.sin((row + 0.3 * col) * t) + 0.1 * row - 0.07 * col


def gradient_values(shape, bounds):
    return np.linspace(*bounds, int(np.prod(shape))).reshape(shape)


def fourier_matrices(rng, samples, dim, modes, t0=0.0, tf=1.0, scale=None):
    """A(t_i) on np.linspace(t0, tf, samples): entries ~ N(0,1)/sqrt(dim),
    drawn in the order A0, B1, C1, ..., BM, CM."""
    s = 1.0 / np.sqrt(dim) if scale is None else scale
    coeffs = rng.standard_normal((2 * modes + 1, dim, dim)) * s
    t = np.linspace(t0, tf, samples)
    out = np.broadcast_to(coeffs[0], (samples, dim, dim)).copy()
    for mth in range(1, modes + 1):
        out += np.sin(2 * np.pi * mth * t)[:, None, None] * coeffs[2 * mth - 1]
        out += np.cos(2 * np.pi * mth * t)[:, None, None] * coeffs[2 * mth]
    return out


def fourier_spacecurve(rng, samples, modes=4, t0=0.0, tf=1.0):
    """f(t) = sum_m a_m sin(2 pi m t + phi_m), a_m ~ N(0,1)/m, phi_m ~ U[0, 2 pi)."""
    m = np.arange(1, modes + 1, dtype=np.float64)
    a = rng.standard_normal((modes, 3)) / m[:, None]
    phi = rng.uniform(0.0, 2 * np.pi, size=(modes, 3))
    t = np.linspace(t0, tf, samples)
    arg = 2 * np.pi * m[None, :, None] * t[:, None, None] + phi[None, :, :]
    return np.sum(a[None] * np.sin(arg), axis=1)


def rel_fro(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    den = np.linalg.norm(b.ravel())
    return np.linalg.norm((a - b).ravel()) / (den if den > 0 else 1.0)
