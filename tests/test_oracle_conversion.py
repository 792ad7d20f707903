"""Pins the oracle's wave energy conversion terms (we_conversion,
diagnostics.f90:967-1076; gamma_conversion, :1080-1256), which the oracle
forms as physical-space sums of c2r fields, with closed forms derived by
Parseval on single modes:

* for fields f = c2r(F), g = c2r(G) holding one L = 1 mode k with kh2 > 0
  (kx = 0, ky > 0 included: realit supplies the mirror image),
  sum_x f g = 2 n1 n2 Re(F_k conj(G_k)); modes at different k give 0;
* so with nA = -kh2 A: we ca = -kh2 Re(AR nBR* + AI nBI*) / (n3 Bu) per mode,
  we cr = -kh2 Re(AI rBR* - AR rBI*) / (2 n3 Bu), gamma ca = -we ca,
  gamma cr = -we cr (the reference inverts the signs and drops the factor 2
  of the normalisation), and gamma cd = kh2 c Re(AR BR* + AI BI*) / (n3 Bu) with
  c = int_factor_w / delt.
Real and imaginary parts are given different values so that a swapped
pairing (AR with nBI, a lost sign of the refraction term, a missing kh2 or a
missing realit mirror) fails.
"""
from types import SimpleNamespace

import numpy as np
import pytest

from oracle import qg_oracle as O


def _state(g, tend, **fields):
    z = np.zeros((g.n3, g.ikty, g.iktx), np.complex128)
    kw = {f: z.copy() for f in ("psik", "qk", "qok", "BRk", "BIk", "BRok", "BIok", "ARk", "AIk", "CRk", "CIk")}
    kw.update(fields)
    return O.State(**kw, tend=tend)


def _params(g, Bu=0.7, dt=2e-3):
    rng = np.random.default_rng(11)
    return SimpleNamespace(Bu=Bu, dt=dt, ifac_w=dt * (0.5 + rng.random((g.ikty, g.iktx))))


def _iky(g, ky):
    return ky if ky >= 0 else g.ikty + ky


@pytest.mark.parametrize("kx,ky,lev", [(3, 2, 1), (0, 4, 0), (3, -2, 2)])
def test_single_mode_closed_forms(kx, ky, lev):
    g = O.Grid(16, 16, 4)
    P = _params(g)
    iy = _iky(g, ky)
    assert g.Lb[iy, kx]
    kh2 = float(kx * kx + ky * ky)
    z = np.zeros((g.n3, g.ikty, g.iktx), np.complex128)
    f = {}
    vals = {"AR": 0.3 - 0.8j, "AI": -0.5 + 0.2j, "nBR": 0.7 + 0.1j, "nBI": -0.4 + 0.9j,
            "rBR": 0.25 - 0.6j, "rBI": 0.55 + 0.35j, "BR": -0.2 + 0.45j, "BI": 0.65 - 0.15j}
    for k, v in vals.items():
        f[k] = z.copy()
        f[k][lev, iy, kx] = v
    S = _state(g, (f["nBR"], f["nBI"], f["rBR"], f["rBI"]), ARk=f["AR"], AIk=f["AI"], BRk=f["BR"], BIk=f["BI"])
    got = O.wave_conversion(g, P, S)

    def re(a, b):
        return (a * np.conj(b)).real

    v = vals
    n3, Bu = g.n3, P.Bu
    c = P.ifac_w[iy, kx] / P.dt
    ca = -kh2 * (re(v["AR"], v["nBR"]) + re(v["AI"], v["nBI"])) / (n3 * Bu)
    cr = -kh2 * (re(v["AI"], v["rBR"]) - re(v["AR"], v["rBI"])) / (2 * n3 * Bu)
    cd = kh2 * c * (re(v["AR"], v["BR"]) + re(v["AI"], v["BI"])) / (n3 * Bu)
    want = np.array([ca, cr, -ca, -cr, cd])
    np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-15)


def test_different_modes_do_not_pair():
    g = O.Grid(16, 16, 3)
    P = _params(g)
    z = np.zeros((g.n3, g.ikty, g.iktx), np.complex128)
    A, B = z.copy(), z.copy()
    A[1, 2, 3] = 0.4 + 0.3j
    B[1, 3, 2] = 0.9 - 0.2j  # (kx, ky) = (2, 3) vs (3, 2)
    S = _state(g, (B, B, B, B), ARk=A, AIk=A, BRk=B, BIk=B)
    assert np.max(np.abs(O.wave_conversion(g, P, S))) < 1e-15


def test_bilinear_in_A_and_tendencies():
    g = O.Grid(16, 16, 4)
    P = _params(g)
    rng = np.random.default_rng(5)
    L = g.Lb

    def rnd():
        a = rng.standard_normal((g.n3, g.ikty, g.iktx)) + 1j * rng.standard_normal((g.n3, g.ikty, g.iktx))
        return np.where(L, a, 0)

    A1, A2, AI = rnd(), rnd(), rnd()
    T = tuple(rnd() for _ in range(4))
    B = rnd()
    w1 = O.wave_conversion(g, P, _state(g, T, ARk=A1, AIk=AI, BRk=B, BIk=B))
    w2 = O.wave_conversion(g, P, _state(g, T, ARk=A2, AIk=AI, BRk=B, BIk=B))
    w0 = O.wave_conversion(g, P, _state(g, T, ARk=0 * A1, AIk=AI, BRk=B, BIk=B))
    w12 = O.wave_conversion(g, P, _state(g, T, ARk=A1 + 2.0 * A2, AIk=AI, BRk=B, BIk=B))
    np.testing.assert_allclose(w12 - w0, (w1 - w0) + 2.0 * (w2 - w0), rtol=1e-10, atol=1e-12)
