"""Pins the oracle's omega equation (omega_eqn_rhs, derivatives.f90:1469-1586;
omega_equation, elliptic.f90:463-567):

* the right-hand side of psi = A(z) sin x + B(z) cos y is the closed form
  2 J(psi_z, lap psi) = 2 (dA Bbar - Abar dB) cos x sin y at each interface
  (dA = (A_{k+1} - A_k)/dz, Abar = (A_{k+1} + A_k)/2) and 0 at the top;
* the solve returns a manufactured w from the operator applied to it, and
  agrees with LAPACK dgtsv (the routine the reference calls) column by column;
  the top level and the kh2 = 0 / L = 0 modes are 0.
"""
import math

import numpy as np
from scipy.linalg import lapack

from oracle import qg_oracle as O

TWOPI = 2 * math.pi


def test_rhs_two_mode_closed_form():
    n, n3 = 16, 6
    g = O.Grid(n, n, n3)
    x = TWOPI * np.arange(n) / n
    X, Y = np.meshgrid(x, x, indexing="xy")  # (y, x)
    A = np.array([0.3, -0.7, 1.1, 0.4, -0.2, 0.9])
    B = np.array([-0.5, 0.2, 0.8, -1.3, 0.6, 0.1])
    psi = A[:, None, None] * np.sin(X)[None] + B[:, None, None] * np.cos(Y)[None]
    rhs = O.fft_c2r(O.omega_eqn_rhs(g, O.fft_r2c(psi)))
    dz = g.dz
    for k in range(n3 - 1):
        dA, dB = (A[k + 1] - A[k]) / dz, (B[k + 1] - B[k]) / dz
        Ab, Bb = 0.5 * (A[k + 1] + A[k]), 0.5 * (B[k + 1] + B[k])
        want = 2.0 * (dA * Bb - Ab * dB) * np.cos(X) * np.sin(Y)
        np.testing.assert_allclose(rhs[k], want, rtol=0, atol=1e-13)
    assert np.max(np.abs(rhs[-1])) < 1e-14


def test_solve_manufactured_and_dgtsv():
    n, n3 = 16, 9
    g = O.Grid(n, n, n3)
    rng = np.random.default_rng(4)
    N2u = 0.5 + rng.random(n3) * 3
    Bu = 0.9
    a = Bu / N2u
    sel = g.Lb & (g.kh2 > 0)
    w = np.zeros((n3, g.ikty, g.iktx), np.complex128)
    w[:-1][:, sel] = rng.standard_normal((n3 - 1, sel.sum())) + 1j * rng.standard_normal((n3 - 1, sel.sum()))
    dz = g.dz
    kh2 = g.kh2.astype(np.float64)
    wp = np.concatenate([np.zeros_like(w[:1]), w])  # w_0 = 0 below the first interface
    rhs = np.zeros_like(w)
    for k in range(n3 - 1):
        rhs[k] = (wp[k] - (2.0 + dz * dz * kh2 / a[k]) * wp[k + 1] + wp[k + 2]) / (dz * dz)
    rhs[-1] = rng.standard_normal(rhs[-1].shape)  # the top level is not solved for
    got = O.omega_equation(g, a, rhs)
    assert np.max(np.abs(got - w)) < 1e-12 * np.max(np.abs(w))
    assert np.all(got[-1] == 0) and np.all(got[:, ~sel] == 0)
    # one column against LAPACK
    iy, ix = np.argwhere(sel)[3]
    m = n3 - 1
    d = -(2.0 + dz * dz * kh2[iy, ix] / a[:m])
    for part in (np.real, np.imag):
        _, _, _, sol, info = lapack.dgtsv(np.ones(m - 1), d.copy(), np.ones(m - 1), (dz * dz) * part(rhs[:m, iy, ix]))
        assert info == 0
        np.testing.assert_allclose(part(got[:m, iy, ix]), sol, rtol=1e-13, atol=1e-14)
