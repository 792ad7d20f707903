"""Pins the oracle's nonlinear operators with closed forms
(derivatives.f90:607-1032, 1295-1338) and its time update with analytic
solutions (main_waqg.f90:237-264, 362-415)."""
import math

import numpy as np
import pytest

from oracle import qg_oracle as O

TWOPI = 2 * math.pi


def _xy(n):
    x = TWOPI * np.arange(n) / n
    return np.meshgrid(x, x, indexing="xy")


def _k(r):
    return O.fft_r2c(np.asarray(r, float))


def test_jacobian_closed_form():
    """psi = sin x sin y, q = cos 2x:  J = u q_x + v q_y = 2 sin x sin 2x cos y."""
    n, nz = 32, 3
    g = O.Grid(n, n, nz)
    X, Y = _xy(n)
    psik = _k(np.broadcast_to(np.sin(X) * np.sin(Y), (nz, n, n)))
    qk = _k(np.broadcast_to(np.cos(2 * X), (nz, n, n)))
    BRk = _k(np.broadcast_to(np.cos(2 * Y), (nz, n, n)))
    BIk = _k(np.broadcast_to(np.sin(X + 3 * Y), (nz, n, n)))
    uk, vk = O.compute_velo(g, psik)
    nq, nBR, nBI = O.convol_waqg(g, uk, vk, qk, BRk, BIk)
    u = -np.sin(X) * np.cos(Y)
    v = np.cos(X) * np.sin(Y)
    jq = u * (-2 * np.sin(2 * X))
    jbr = v * (-2 * np.sin(2 * Y))
    jbi = u * np.cos(X + 3 * Y) + v * 3 * np.cos(X + 3 * Y)
    for got, want in ((nq, jq), (nBR, jbr), (nBI, jbi)):
        ref = np.where(g.Lb, _k(np.broadcast_to(want, (nz, n, n))), 0)
        np.testing.assert_allclose(got, ref, atol=1e-14)


def test_refraction_closed_form():
    """zeta = lap(psi) = -2 sin x sin y; rBR = L*F(zeta BR)."""
    n, nz = 32, 2
    g = O.Grid(n, n, nz)
    X, Y = _xy(n)
    psik = _k(np.broadcast_to(np.sin(X) * np.sin(Y), (nz, n, n)))
    BR = np.cos(X) * np.cos(2 * Y)
    BI = np.sin(3 * X)
    rBR, rBI = O.refraction_waqg(g, _k(np.broadcast_to(BR, (nz, n, n))), _k(np.broadcast_to(BI, (nz, n, n))), psik)
    zeta = -2 * np.sin(X) * np.sin(Y)
    np.testing.assert_allclose(rBR, np.where(g.Lb, _k(np.broadcast_to(zeta * BR, (nz, n, n))), 0), atol=1e-14)
    np.testing.assert_allclose(rBI, np.where(g.Lb, _k(np.broadcast_to(zeta * BI, (nz, n, n))), 0), atol=1e-14)


def test_feedback_closed_form():
    """B = cos x + i sin y: (i/2)J(B*,B) + lap|B|^2/4
    = sin x cos y - cos 2x / 2 + cos 2y / 2."""
    n, nz = 32, 2
    g = O.Grid(n, n, nz)
    X, Y = _xy(n)
    Ro, W2F = 0.3, 0.05
    qw = O.compute_qw(g, _k(np.broadcast_to(np.cos(X), (nz, n, n))), _k(np.broadcast_to(np.sin(Y), (nz, n, n))),
                      Ro, W2F)
    want = np.sin(X) * np.cos(Y) - 0.5 * np.cos(2 * X) + 0.5 * np.cos(2 * Y)
    np.testing.assert_allclose(qw, Ro * W2F * np.where(g.Lb, _k(np.broadcast_to(want, (nz, n, n))), 0), atol=1e-15)


def _params(g, **kw):
    n = g.n1
    z = np.zeros((n, n // 2 + 1))
    d = dict(n1=n, n3=g.n3, dt=1e-2, gamma=0.0, Ro=0.2, Bu=1.0, W2F=0.0, nuz=0.0, flags=0, ifac_q=z, ifac_w=z,
             a_ell_ut=None, N2u=np.ones(g.n3), N2s=np.ones(g.n3))
    d.update(kw)
    if d["a_ell_ut"] is None:
        d["a_ell_ut"] = d["Bu"] / d["N2u"]
    return O.Params(**d)


def test_hyperdiffusion_decay_is_exponential():
    """No flow, no dispersion, gamma = 0: q^n = q^0 exp(-n f) exactly
    (Euler e^{-f}, then leapfrog e^{-2f} on the old value)."""
    n, n3 = 16, 4
    g = O.Grid(n, n, n3)
    f = np.abs(np.random.default_rng(0).standard_normal(g.L.shape)) * 0.1
    P = _params(g, ifac_q=f, ifac_w=f * 0.5, flags=O.LINEAR | O.NO_DISPERSION | O.FIXED_FLOW)
    rng = np.random.default_rng(1)
    B0 = np.where(g.Lb, rng.standard_normal((n3,) + g.L.shape) + 1j * rng.standard_normal((n3,) + g.L.shape), 0)
    z = np.zeros_like(B0)
    S = O.State(psik=z.copy(), qk=z.copy(), qok=z.copy(), BRk=B0.copy(), BIk=B0.copy() * 0.5, BRok=z.copy(),
                BIok=z.copy(), ARk=z.copy(), AIk=z.copy(), CRk=z.copy(), CIk=z.copy(), iter=0)
    S = O.run(g, P, S, 7)
    np.testing.assert_allclose(S.BRk, B0 * np.exp(-7 * 0.5 * f), rtol=1e-13, atol=1e-15)
    np.testing.assert_allclose(S.BIk, 0.5 * B0 * np.exp(-7 * 0.5 * f), rtol=1e-13, atol=1e-15)


def test_robert_filter_fixed_point():
    """A constant mode (f = 0, no tendencies) is a fixed point of the
    leapfrog + Robert-Asselin step for any gamma."""
    n, n3 = 16, 3
    g = O.Grid(n, n, n3)
    P = _params(g, gamma=0.2, flags=O.LINEAR | O.NO_DISPERSION | O.FIXED_FLOW)
    B0 = np.where(g.Lb, 0.7 - 0.2j, 0) * np.ones((n3, 1, 1))
    z = np.zeros_like(B0)
    S = O.State(psik=z.copy(), qk=z.copy(), qok=z.copy(), BRk=B0.copy(), BIk=z.copy(), BRok=z.copy(),
                BIok=z.copy(), ARk=z.copy(), AIk=z.copy(), CRk=z.copy(), CIk=z.copy(), iter=0)
    S = O.run(g, P, S, 5)
    np.testing.assert_allclose(S.BRk, B0, rtol=1e-15)
    np.testing.assert_allclose(S.BRok, B0, rtol=1e-15)


def test_dispersion_rotation_rate():
    """Waves only (no flow): B on the horizontal mode k = (1, 0) times the
    Neumann eigenvector m = 1 rotates as B ~ exp(i w t) with
    w = c kh2 * dz^2 Bu / mu,  mu = Bu*(-2(1-cos(pi/n3))) - kh2 dz^2/4,
    c = 1/(2 Bu Ro) -- the YBJ+ dispersion term i/(2 Bu Ro) lap A with
    L+ A = B.  With A^0 = 0 the Euler step leaves B^1 = B^0; afterwards
    leapfrog gives B^{n+1} = B^{n-1} + 2 i w dt B^n, so
    B^n = a r+^n + b r-^n, r+- = i w dt +- sqrt(1 - (w dt)^2)."""
    n, n3 = 16, 8
    g = O.Grid(n, n, n3)
    Bu, Ro, dt = 1.0, 0.5, 1e-3
    P = _params(g, Bu=Bu, Ro=Ro, dt=dt, flags=O.LINEAR | O.FIXED_FLOW)
    k = np.arange(1, n3 + 1)
    v = np.cos(math.pi * (k - 0.5) / n3)
    dz = g.dz
    mu = Bu * (-2 * (1 - math.cos(math.pi / n3))) - 1 * dz * dz / 4
    w = (0.5 / (Bu * Ro)) * 1 * (dz * dz * Bu / mu)
    B0 = np.zeros((n3,) + g.L.shape, complex)
    B0[:, 0, 1] = v
    z = np.zeros_like(B0)
    S = O.State(psik=z.copy(), qk=z.copy(), qok=z.copy(), BRk=B0.real.astype(complex), BIk=z.copy(),
                BRok=z.copy(), BIok=z.copy(), ARk=z.copy(), AIk=z.copy(), CRk=z.copy(), CIk=z.copy(), iter=0)
    wd = w * dt
    rp = 1j * wd + math.sqrt(1 - wd * wd)
    rm = 1j * wd - math.sqrt(1 - wd * wd)
    # a + b = 1 (n = 0), a rp + b rm = 1 (n = 1)
    a = (1 - rm) / (rp - rm)
    b = 1 - a
    assert w < 0  # the operator is negative definite
    for it in range(1, 61):
        S = O.step(g, P, S)
        got = S.BRk[:, 0, 1] + 1j * S.BIk[:, 0, 1]
        want = v * (a * rp ** it + b * rm ** it)
        np.testing.assert_allclose(got, want, rtol=1e-10, atol=1e-13)
