"""CPU checks of the C-ABI library and the host-side tables (no GPU calls)."""
import ctypes
import os
import re

import numpy as np
import pytest

from oracle import qg_oracle as O
import qgybj
from qgybj import params as qp

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _header_functions():
    txt = open(os.path.join(ROOT, "include", "qgybj.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(qg_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol(lib):
    declared = _header_functions()
    assert len(declared) >= 20
    for name in declared:
        assert hasattr(lib, name), name
    assert sorted(qgybj.SYMBOLS) == declared


def test_version(lib):
    assert lib.qg_version().decode().startswith("qgybj")


@pytest.mark.parametrize("n", [8, 16, 32, 64, 128, 256, 512, 1024])
def test_library_mask_matches_oracle(n):
    g = O.Grid(n, n, 2)
    np.testing.assert_array_equal(qgybj.mask(n), g.L)


@pytest.mark.parametrize("n", [16, 64, 256, 512])
def test_box_contains_every_retained_mode(n):
    kmax, KX, nlo, nhi = qgybj.box(n)
    g = O.Grid(n, n, 2)
    kx_ret = g.kx[g.Lb]
    ky_ret = g.ky[g.Lb]
    assert kx_ret.max() == kmax and KX == kmax + 1
    assert ky_ret.max() == kmax and ky_ret.min() == -kmax
    assert nlo == kmax + 1 and nhi == kmax


@pytest.mark.parametrize("n,ilap", [(16, 4), (32, 2), (64, 6)])
def test_int_factor_tables_match_oracle(n, ilap):
    nu1 = qp.nuh_coeff(0.01, n, 2)
    nu2 = qp.nuh_coeff(10.0, n, ilap)
    g = O.Grid(n, n, 2)
    a = qp.int_factor(n, 1e-3, nu1, 2, nu2, ilap)
    b = O.int_factor_table(g, 1e-3, nu1, 2, nu2, ilap)
    np.testing.assert_array_equal(a, b)


def test_f32_powers_are_correctly_rounded():
    # exact for small cases, and equal to numpy's float32 rounding of the
    # exact integer where the integer fits a double exactly
    for k in range(0, 130):
        for e in (4, 8):
            v = k ** e
            got = O.f32_pow_int(k, e)
            if v < 2 ** 53:
                assert got == np.float32(float(v)), (k, e)


def test_gamma_is_the_float32_literal():
    P = qp.make_params(16, 4)
    assert P.gamma == float(np.float32(1e-3)) == 0.0010000000474974513


def test_kx_ky_match_oracle():
    for n in (8, 64):
        kxa, kya = qp.kx_ky(n)
        g = O.Grid(n, n, 2)
        np.testing.assert_array_equal(kxa, g.kxa)
        np.testing.assert_array_equal(kya, g.kya)


def test_create_rejects_bad_parameters(lib):
    P = qp.make_params(16, 4)
    keep = [np.ascontiguousarray(a) for a in (P.ifac_q, P.ifac_w, P.N2u, P.N2s)]
    dp = [a.ctypes.data_as(ctypes.POINTER(ctypes.c_double)) for a in keep]
    h = ctypes.c_void_p()
    # n1 != n2 and unsupported sizes are refused before any device work
    for n1, n2, n3 in ((16, 32, 4), (24, 24, 4), (16, 16, 1), (2048, 2048, 4)):
        prm = qgybj._Params(n1, n2, n3, 0, P.dt, P.gamma, P.Ro, P.Bu, P.W2F, P.nuz, *dp)
        assert lib.qg_create(ctypes.byref(prm), ctypes.byref(h)) == 1
    prm = qgybj._Params(16, 16, 4, 0, -1.0, P.gamma, P.Ro, P.Bu, P.W2F, P.nuz, *dp)
    assert lib.qg_create(ctypes.byref(prm), ctypes.byref(h)) == 1
    prm = qgybj._Params(16, 16, 4, 0, P.dt, P.gamma, P.Ro, P.Bu, P.W2F, P.nuz, None, *dp[1:])
    assert lib.qg_create(ctypes.byref(prm), ctypes.byref(h)) == 1
