"""Pins the oracle's vertical inversions (elliptic.f90:145-460) and init_q
(init.f90:591-652) against LAPACK dgtsv, analytic Neumann eigenpairs and
the init_q operator as exact inverse."""
import math

import numpy as np
import pytest
from scipy.linalg import lapack

from oracle import qg_oracle as O


def test_thomas_matches_lapack_dgtsv():
    rng = np.random.default_rng(7)
    n, ncol = 40, 25
    a = rng.uniform(0.5, 2.0, n)
    dl, dsum, du = O._coefficients(a)
    kh = rng.uniform(0.01, 3.0, ncol)
    d = -(dsum[:, None] + kh[None, :])
    b = rng.standard_normal((n, ncol))
    x = O.thomas_dgtsv(np.repeat(dl[:, None], ncol, 1), d, np.repeat(du[:, None], ncol, 1), b)
    for j in range(ncol):
        _, _, _, xl, info = lapack.dgtsv(dl.copy(), d[:, j].copy(), du.copy(), b[:, j].copy())
        assert info == 0
        np.testing.assert_allclose(x[:, j], xl, rtol=1e-13, atol=1e-15)


def _neumann_eig(n3, m):
    """Eigenvector / eigenvalue of the constant-coefficient Neumann second
    difference: v_k = cos(m pi (k-1/2)/n3), lambda = -2 (1 - cos(m pi/n3))."""
    k = np.arange(1, n3 + 1)
    v = np.cos(m * math.pi * (k - 0.5) / n3)
    lam = -2.0 * (1.0 - math.cos(m * math.pi / n3))
    return v, lam


@pytest.mark.parametrize("m", [0, 1, 3])
def test_psi_solver_on_neumann_eigenvectors(m):
    n, n3 = 16, 12
    g = O.Grid(n, n, n3)
    Bu = 0.7
    a_ell, _, _ = O.base_state(g, Bu, "constant")
    v, lam = _neumann_eig(n3, m)
    dz = g.dz
    qs = np.where(g.Lb, 1.0 + 0.5j, 0)[None] * v[:, None, None]
    psi = O.psi_solver(g, a_ell, qs)
    # T(kh2) = Bu * D2 - kh2 dz^2 I;  T psi = dz^2 q  ->  psi = dz^2 q / (Bu lam - kh2 dz^2)
    kh2 = g.kh2.astype(float)
    expect = np.where(g.Lb & (g.kh2 != 0), dz * dz * qs / (Bu * lam - kh2 * dz * dz), 0)
    np.testing.assert_allclose(psi, expect, rtol=1e-12, atol=1e-14)


@pytest.mark.parametrize("m", [1, 2])
def test_A_solver_on_neumann_eigenvectors(m):
    n, n3 = 16, 10
    g = O.Grid(n, n, n3)
    Bu = 1.3
    a_ell, _, _ = O.base_state(g, Bu, "constant")
    v, lam = _neumann_eig(n3, m)
    dz = g.dz
    B = np.where(g.Lb, 0.3 - 1.1j, 0)[None] * v[:, None, None]
    A, C = O.A_solver_ybj_plus(g, a_ell, Bu, B)
    kh2 = g.kh2.astype(float)
    expect = np.where(g.Lb & (g.kh2 != 0), dz * dz * Bu * B / (Bu * lam - kh2 * dz * dz / 4.0), 0)
    np.testing.assert_allclose(A, expect, rtol=1e-12, atol=1e-14)
    np.testing.assert_allclose(C[:-1], (A[1:] - A[:-1]) / dz, rtol=1e-14, atol=1e-15)
    assert np.all(C[-1] == 0)


@pytest.mark.parametrize("kind", ["constant", "exponential", "skewed_gaussian"])
def test_psi_solver_inverts_init_q(kind):
    n, n3 = 32, 16
    g = O.Grid(n, n, n3)
    a_ell, _, _ = O.base_state(g, 0.8, kind)
    rng = np.random.default_rng(11)
    psi = rng.standard_normal((n3,) + g.L.shape) + 1j * rng.standard_normal((n3,) + g.L.shape)
    psi = np.where(g.Lb & (g.kh2 != 0), psi, 0)
    q = O.init_q(g, a_ell, psi)
    back = O.psi_solver(g, a_ell, q)
    np.testing.assert_allclose(back, psi, rtol=1e-11, atol=1e-12)


def test_init_q_on_constant_in_z_is_minus_kh2():
    """For z-independent psi the Neumann second difference vanishes, so
    init_q gives -kh2 psi exactly (including kh2 = 0 -> 0)."""
    n, n3 = 16, 6
    g = O.Grid(n, n, n3)
    a_ell, _, _ = O.base_state(g, 1.0, "exponential")
    psi = np.where(g.Lb, 1.0 + 2.0j, 0)[None].repeat(n3, 0)
    q = O.init_q(g, a_ell, psi)
    np.testing.assert_allclose(q, -g.kh2 * psi, rtol=1e-12, atol=1e-12)


def test_operator_is_symmetric():
    n3 = 9
    a = np.exp(np.linspace(0, 2, n3))
    dl, dsum, du = O._coefficients(a)
    np.testing.assert_array_equal(dl, du)
    assert dsum[0] == a[0] and dsum[-1] == a[-2]
