"""Pins the oracle's wave spectra (hspec_waves, diagnostics.f90:547-638) with
closed forms: a single wave mode lands in its shell with 1/2 |BR + i BI|^2;
a wave amplitude linear in z gives the potential part from C = A_z with the
interface N^2 of the levels below and above (an off-by-one level fails); the
shell counts of the first shells and the total equal the retained-mode
counts; and the shell index the HIP kernel uses, floor(|k| + 1/2), equals the
reference's REAL(4) expression for every wavenumber up to n1 = 512."""
import math
from types import SimpleNamespace

import numpy as np
import pytest

from oracle import qg_oracle as O

TWOPI = 2 * math.pi


def _state(g, **fields):
    z = np.zeros((g.n3, g.ikty, g.iktx), np.complex128)
    kw = {f: z.copy() for f in ("psik", "qk", "qok", "BRk", "BIk", "BRok", "BIok", "ARk", "AIk", "CRk", "CIk")}
    kw.update(fields)
    return O.State(**kw)


def test_single_mode_kinetic_part():
    g = O.Grid(32, 32, 6)
    P = SimpleNamespace(Bu=1.0, N2u=np.ones(g.n3))
    BR = np.zeros((g.n3, g.ikty, g.iktx), np.complex128)
    BI = np.zeros_like(BR)
    kx, ky, lev = 3, 4, 3  # |k| = 5
    c, d = 0.7 - 0.2j, 0.1 + 0.4j
    BR[lev - 1, ky, kx] = c
    BI[lev - 1, ky, kx] = d
    S = _state(g, BRk=BR, BIk=BI)
    sp = O.hspec_waves(g, P, S, lev)
    assert sp.shape == (16, 3)
    want = np.zeros(16)
    want[5] = 0.5 * abs(c + 1j * d) ** 2
    np.testing.assert_allclose(sp[:, 0], want, rtol=1e-15, atol=0)
    assert np.all(sp[:, 1] == 0)
    assert np.all(O.hspec_waves(g, P, S, lev + 1)[:, 0] == 0)


@pytest.mark.parametrize("n,count", [(64, 657), (128, 2841)])
def test_shell_counts(n, count):
    g = O.Grid(n, n, 2)
    P = SimpleNamespace(Bu=1.0, N2u=np.ones(2))
    sp = O.hspec_waves(g, P, _state(g), 1)
    # shell 0: (0,0); shell 1: (1,0), (0,1), (1,1), (1,-1) (kx = 0, ky < 0 is not retained)
    assert sp[0, 2] == 2 and sp[1, 2] == 8
    assert sp[:, 2].sum() == 2 * count


def test_potential_part_linear_amplitude():
    g = O.Grid(32, 32, 7)
    N2u = 1.0 + np.arange(g.n3)  # distinct interface values pin the level indexing
    Bu = 0.8
    P = SimpleNamespace(Bu=Bu, N2u=N2u)
    kx, ky = 2, 1  # kh2 = 5, shell floor(sqrt(5) + 1/2) = 2
    a = 0.3 + 0.5j
    AR = np.zeros((g.n3, g.ikty, g.iktx), np.complex128)
    AR[:, ky, kx] = a * g.dz * np.arange(1, g.n3 + 1)  # C = a below the top interface, 0 at it
    S = _state(g, ARk=AR)
    kh2 = 5.0
    for lev in range(1, g.n3 + 1):
        c = [abs(a) ** 2 if k < g.n3 - 1 else 0.0 for k in range(g.n3)]
        z = lev - 1
        pe = kh2 * c[z] / (Bu * N2u[z])
        if z > 0:
            pe += kh2 * c[z - 1] / (Bu * N2u[z - 1])
        sp = O.hspec_waves(g, P, S, lev)
        np.testing.assert_allclose(sp[2, 1], 0.5 * 0.25 * pe, rtol=1e-14)
        assert np.count_nonzero(sp[:, 1]) == (1 if pe else 0)


def test_kernel_shell_index_matches_reference_expression():
    k = np.arange(0, 513)
    kh2 = (k[:, None] ** 2 + k[None, :] ** 2).astype(np.float64)
    kh = np.sqrt(kh2)
    ref = np.float32(kh * TWOPI / TWOPI + 0.5).astype(np.int64)
    assert np.array_equal(np.floor(kh + 0.5).astype(np.int64), ref)


def test_we_vave_single_modes():
    """we_vave (diagnostics.f90:750-813) of BR holding one mode a at (kx, ky)
    and BI one mode b at (0, ky'): c2r gives 2 Re(a e^{i(kx x + ky y)}) (the
    realit mirror supplies the kx = 0 conjugate), so the map is
    0.5 (f_R^2 + f_I^2) / n3 at every grid point."""
    g = O.Grid(16, 16, 5)
    BR = np.zeros((g.n3, g.ikty, g.iktx), np.complex128)
    BI = np.zeros_like(BR)
    a, b = 0.3 - 0.4j, -0.2 + 0.7j
    BR[2, g.ikty - 3, 2] = a  # (kx, ky) = (2, -3)
    BI[4, 5, 0] = b           # (0, 5)
    S = _state(g, BRk=BR, BIk=BI)
    x = TWOPI * np.arange(16) / 16
    X, Y = np.meshgrid(x, x, indexing="xy")  # (y, x)
    fR = 2 * (a * np.exp(1j * (2 * X - 3 * Y))).real
    fI = 2 * (b * np.exp(1j * (5 * Y))).real
    np.testing.assert_allclose(O.we_vave(g, S), 0.5 * (fR ** 2 + fI ** 2) / g.n3, rtol=1e-13, atol=1e-15)
