"""Pins the oracle's transforms and mask against independent arithmetic:
brute-force DFT sums, analytic single modes, and the mask formula evaluated
mode by mode (fft.f90:55-107, init.f90:32-80)."""
import math

import numpy as np
import pytest

from oracle import qg_oracle as O

TWOPI = 2 * math.pi


def brute_r2c(r):
    """Direct DFT: F(kx,ky) = 1/(n1 n2) sum f(x,y) exp(-i(kx x + ky y)),
    kx = 0..n1/2, ky in storage order."""
    nz, n2, n1 = r.shape
    x = np.arange(n1)
    y = np.arange(n2)
    out = np.zeros((nz, n2, n1 // 2 + 1), complex)
    for iky in range(n2):
        for kx in range(n1 // 2 + 1):
            ph = np.exp(-2j * np.pi * (kx * x[None, :] / n1 + iky * y[:, None] / n2))
            out[:, iky, kx] = np.sum(r * ph[None], axis=(1, 2)) / (n1 * n2)
    return out


def brute_c2r(k):
    """Direct inverse of the Hermitian extension: sum over the full (kx, ky)
    grid with X(-kx,-ky) = conj X(kx,ky) for the kx = 1..n1/2-1 columns, and
    the kx = 0, n1/2 columns taken as given (after realit) with the real-part
    convention of a c2r on the x axis."""
    nz, n2, h = k.shape
    n1 = 2 * (h - 1)
    full = np.zeros((nz, n2, n1), complex)
    full[:, :, :h] = k
    for kx in range(1, n1 // 2):
        for iky in range(n2):
            full[:, (-iky) % n2, n1 - kx] = np.conj(k[:, iky, kx])
    x = np.arange(n1)
    y = np.arange(n2)
    out = np.zeros((nz, n2, n1))
    for iky in range(n2):
        for kx in range(n1):
            ph = np.exp(2j * np.pi * (kx * x[None, :] / n1 + iky * y[:, None] / n2))
            out += np.real(full[:, iky, kx][:, None, None] * ph[None])
    # the x-direction c2r keeps only the real part of the kx = 0 and n1/2
    # columns after the y transform; for a Hermitian column the imaginary
    # part of the full sum is zero anyway
    return out


@pytest.mark.parametrize("n", [8, 16])
def test_r2c_matches_brute_force(n):
    rng = np.random.default_rng(1)
    r = rng.standard_normal((3, n, n))
    np.testing.assert_allclose(O.fft_r2c(r), brute_r2c(r), atol=1e-14, rtol=0)


@pytest.mark.parametrize("n", [8, 16])
def test_c2r_matches_brute_force_on_realit_input(n):
    rng = np.random.default_rng(2)
    k = rng.standard_normal((2, n, n // 2 + 1)) + 1j * rng.standard_normal((2, n, n // 2 + 1))
    # a Hermitian-consistent half spectrum on the kx = 0 and n/2 columns
    for c in (0, n // 2):
        k[:, 0, c] = k[:, 0, c].real
        k[:, n // 2, c] = k[:, n // 2, c].real
        for j in range(1, n // 2):
            k[:, n - j, c] = np.conj(k[:, j, c])
    np.testing.assert_allclose(O.fft_c2r(k), brute_c2r(k), atol=1e-12, rtol=0)


def test_realit_fills_negative_ky_on_kx0_only():
    n = 16
    rng = np.random.default_rng(3)
    k = rng.standard_normal((1, n, n // 2 + 1)) + 1j * rng.standard_normal((1, n, n // 2 + 1))
    k2 = O.realit(k.copy())
    for j in range(1, n // 2):
        assert k2[0, n - j, 0] == np.conj(k[0, j, 0])
    # untouched: kx > 0, ky = 0 row, Nyquist row
    np.testing.assert_array_equal(k2[:, :, 1:], k[:, :, 1:])
    assert k2[0, 0, 0] == k[0, 0, 0] and k2[0, n // 2, 0] == k[0, n // 2, 0]


def test_c2r_uses_realit_not_stored_negative_ky():
    """c2r must ignore what is stored at (0, ky<0): the reference overwrites
    it with conj(0, ky>0) before transforming (fft.f90:77, 88-107)."""
    n = 16
    rng = np.random.default_rng(4)
    k = rng.standard_normal((1, n, n // 2 + 1)) + 1j * rng.standard_normal((1, n, n // 2 + 1))
    k2 = k.copy()
    k2[0, n // 2 + 1:, 0] = 123.0 + 45j
    np.testing.assert_array_equal(O.fft_c2r(k), O.fft_c2r(k2))


def test_known_answer_sinx_siny():
    n = 16
    x = TWOPI * np.arange(n) / n
    X, Y = np.meshgrid(x, x, indexing="xy")
    r = (np.sin(X) * np.sin(Y))[None]
    k = O.fft_r2c(r)
    # sin x sin y = -1/4 (e^{i(x+y)} - e^{i(x-y)} - e^{-i(x-y)} + e^{-i(x+y)})
    assert abs(k[0, 1, 1] - (-0.25)) < 1e-15
    assert abs(k[0, n - 1, 1] - 0.25) < 1e-15
    k[0, 1, 1] = 0
    k[0, n - 1, 1] = 0
    assert np.max(np.abs(k)) < 1e-15


def test_round_trip_masked_spectrum():
    n = 32
    g = O.Grid(n, n, 4)
    rng = np.random.default_rng(5)
    k = (rng.standard_normal((4, n, n // 2 + 1)) + 1j * rng.standard_normal((4, n, n // 2 + 1)))
    k[:, 0, 0] = k[:, 0, 0].real
    k = np.where(g.Lb, k, 0)
    back = O.fft_r2c(O.fft_c2r(k))
    # the round trip restores L = 1 modes and fills (0,-ky) by realit
    np.testing.assert_allclose(np.where(g.Lb, back, 0), k, atol=1e-14)
    np.testing.assert_allclose(back[:, n - 3, 0], np.conj(k[:, 3, 0]), atol=1e-14)


# ---------------------------------------------------------------------------
# mask and wavenumbers
# ---------------------------------------------------------------------------
def brute_mask(n):
    iktx, ikty = n // 2 + 1, n
    L = np.zeros((ikty, iktx), int)
    for iky in range(ikty):
        ky = iky if iky < n // 2 else (0 if iky == n // 2 else iky - n)
        for kx in range(iktx):
            if iky == n // 2:
                continue
            if kx == 0 and ky < 0:
                continue
            kh = float(np.float32(math.sqrt(float(kx * kx + ky * ky))))
            cut = int(n / 3 + 0.5) - 0.5
            if kh > cut:
                continue
            L[iky, kx] = 1
    return L


@pytest.mark.parametrize("n,cut,count", [(64, 20.5, 657), (128, 42.5, 2841), (256, 84.5, 11215),
                                         (512, 170.5, 45681)])
def test_mask_cut_and_counts(n, cut, count):
    g = O.Grid(n, n, 2)
    assert g.cut == cut
    assert int(g.L.sum()) == count
    np.testing.assert_array_equal(g.L, brute_mask(n))


def test_kya_order():
    g = O.Grid(8, 8, 2)
    assert list(g.kya) == [0, 1, 2, 3, 0, -3, -2, -1]
    assert list(g.kxa) == [0, 1, 2, 3, 4]
