"""Pins the oracle's regular-YBJ operator (ybj_plus = 0: sumB, compute_sigma,
compute_A, derivatives.f90:1037-1231) with the mathematics it discretises:

* sumB leaves a zero vertical sum at every L = 1 mode with kh2 > 0 and the
  other modes unchanged;
* compute_A inverts the discrete YBJ relation B = (A_z / N^2)_z: with
  D_k = C_k / N^2_u(k) (0 at the top interface and below the bottom),
  (D_k - D_{k-1}) / dz = B_k at every level, and sum_z A = sigma (a shifted
  N^2 index or a reordered running sum fails the first check);
* compute_sigma's sums are the ones that keep the vertical sum of B constant
  under the update equations: sum_z dB/dt = 0 for BR and BI (a sign or
  factor error in sigma fails it).
"""
import numpy as np

from oracle import qg_oracle as O


def _rand(rng, g, n3):
    return rng.standard_normal((n3, g.ikty, g.iktx)) + 1j * rng.standard_normal((n3, g.ikty, g.iktx))


def test_sumB_zero_vertical_sum():
    g = O.Grid(32, 32, 9)
    rng = np.random.default_rng(1)
    BR, BI = _rand(rng, g, 9), _rand(rng, g, 9)
    R, I = O.sumB(g, BR, BI)
    sel = g.Lb & (g.kh2 > 0)
    for new, old in ((R, BR), (I, BI)):
        assert np.max(np.abs(new.sum(axis=0)[sel])) < 1e-13
        assert np.array_equal(new[:, ~sel], old[:, ~sel])
        # a constant shift per column
        d = (old - new)[:, sel]
        assert np.max(np.abs(d - d[0])) < 1e-14


def test_compute_A_inverts_ybj_relation():
    n3 = 11
    g = O.Grid(32, 32, n3)
    rng = np.random.default_rng(2)
    N2u = 0.5 + rng.random(n3) * 3
    BR, BI = O.sumB(g, _rand(rng, g, n3), _rand(rng, g, n3))
    sigma = (rng.standard_normal((g.ikty, g.iktx)) + 1j * rng.standard_normal((g.ikty, g.iktx)),
             rng.standard_normal((g.ikty, g.iktx)) + 1j * rng.standard_normal((g.ikty, g.iktx)))
    AR, AI, CR, CI = O.compute_A(g, N2u, BR, BI, sigma)
    sel = g.Lb & (g.kh2 > 0)
    dz = g.dz
    for A, C, B, sg in ((AR, CR, BR, sigma[0]), (AI, CI, BI, sigma[1])):
        assert np.array_equal(C[:-1], (A[1:] - A[:-1]) / dz) and np.all(C[-1] == 0)
        D = C / N2u[:, None, None]
        D[-1] = 0
        lhs = (D - np.concatenate([np.zeros_like(D[:1]), D[:-1]])) / dz
        scale = np.max(np.abs(B[:, sel]))
        assert np.max(np.abs((lhs - B)[:, sel])) < 1e-12 * scale
        np.testing.assert_allclose(A.sum(axis=0)[sel], sg[sel], rtol=1e-12, atol=1e-12 * np.abs(sg).max())
        assert np.all(A[:, ~sel] == 0)


def test_sigma_keeps_vertical_sum_of_B():
    n3 = 7
    g = O.Grid(32, 32, n3)
    rng = np.random.default_rng(3)
    Bu, Ro = 0.7, 0.15
    nBR, nBI, rBR, rBI = (_rand(rng, g, n3) for _ in range(4))
    BR, BI = O.sumB(g, _rand(rng, g, n3), _rand(rng, g, n3))
    AR, AI, _, _ = O.compute_A(g, np.ones(n3), BR, BI, O.compute_sigma(g, Bu, Ro, nBR, nBI, rBR, rBI))
    c = 0.5 / (Bu * Ro)
    kh2 = g.kh2.astype(np.float64)
    sel = g.Lb & (g.kh2 > 0)
    # the B tendencies of the update (main_waqg.f90:251-253, 376-377)
    dBR = -(nBR + (c * kh2) * AI - 0.5 * rBI)
    dBI = -(nBI - (c * kh2) * AR + 0.5 * rBR)
    for d in (dBR, dBI):
        scale = np.max(np.abs(d[:, sel]))
        assert np.max(np.abs(d.sum(axis=0)[sel])) < 1e-12 * n3 * scale
