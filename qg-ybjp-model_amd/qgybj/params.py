"""Model parameters in the form the C ABI takes them (qg_params).

Mirrors the derivations of the reference's compile-time parameters module
(parameters_test1.f90:45-113) and the integrating-factor tables of the loop
(main_waqg.f90:246-247, 371-372), including the places where the reference
evaluates in REAL(4): default-real literals (``gamma = 1e-3``), the
hyperviscosity coefficients ``nuh`` and the powers ``(1.*kx)**(2.*ilap)``.
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np

TWOPI = 4.0 * math.asin(1.0)

FIXED_FLOW = 1
NO_FEEDBACK = 2
PASSIVE_SCALAR = 4
LINEAR = 8
NO_DISPERSION = 16
INVISCID = 32
YBJ_REGULAR = 64  # ybj_plus = 0: the regular YBJ operator instead of YBJ+


def _f32(x) -> float:
    return float(np.float32(x))


def _f32_ipow(x: np.float32, n: int) -> np.float32:
    """REAL(4) ** INTEGER by binary powering (each product rounded to REAL(4))."""
    x = np.float32(x)
    if n == 0:
        return np.float32(1.0)
    result = None
    base = x
    while n:
        if n & 1:
            result = base if result is None else np.float32(result * base)
        n >>= 1
        if n:
            base = np.float32(base * base)
    return np.float32(result)


def nuh_coeff(coeff: float, n1: int, ilap: int) -> float:
    """coeff*(64./(1.*n1))**(4./3.)*(3./n1)**(2*(ilap-1))
    (parameters_test1.f90:106-109) with the REAL(4) sub-expressions."""
    a = np.float32(np.float32(64.0) / np.float32(n1))
    pa = np.float32(a ** np.float32(np.float32(4.0) / np.float32(3.0)))
    b = np.float32(np.float32(3.0) / np.float32(n1))
    pb = _f32_ipow(b, 2 * (ilap - 1))
    return (float(coeff) * float(pa)) * float(pb)


def _round_int_to_f32(v: int) -> float:
    """Nearest float32 (ties to even) of a non-negative integer."""
    if v == 0:
        return 0.0
    b = v.bit_length()
    if b <= 24:
        return float(v)
    shift = b - 24
    q = v >> shift
    rem = v - (q << shift)
    half = 1 << (shift - 1)
    if rem > half or (rem == half and (q & 1)):
        q += 1
    return float(np.float32(q << shift))


def kx_ky(n1: int):
    """kxa, kya of init.f90:32-41 (Nyquist row relabelled 0)."""
    iktx = n1 // 2 + 1
    kxa = np.arange(iktx, dtype=np.int64)
    kya = np.array([j if j < n1 // 2 else (0 if j == n1 // 2 else j - n1) for j in range(n1)], dtype=np.int64)
    return kxa, kya


def int_factor(n1: int, delt: float, nu1: float, ilap1: int, nu2: float, ilap2: int) -> np.ndarray:
    """int_factor(kx,ky) as the reference forms it (main_waqg.f90:371):
    REAL(4) powers (correctly rounded) summed in REAL(4), the rest in REAL(8).
    Returned as (ikty, iktx) = Fortran (iktx, ikty) column-major."""
    kxa, kya = kx_ky(n1)
    cache = {}

    def p(k, e):
        key = (abs(int(k)), e)
        if key not in cache:
            cache[key] = np.float32(_round_int_to_f32(abs(int(k)) ** e))
        return cache[key]

    out = np.empty((len(kya), len(kxa)), dtype=np.float64)
    for iy, ky in enumerate(kya):
        for ix, kx in enumerate(kxa):
            s1 = np.float32(p(kx, 2 * ilap1) + p(ky, 2 * ilap1))
            s2 = np.float32(p(kx, 2 * ilap2) + p(ky, 2 * ilap2))
            out[iy, ix] = delt * (nu1 * float(s1) + nu2 * float(s2))
    return out


def n2_profile(z, kind: str, ratio: float = 100.0):
    """Nondimensional N^2(z): 'constant' (init.f90:186-188), 'skewed_gaussian'
    (init.f90:189-191 with the default-REAL constants of
    parameters_test1.f90:58-63 and the REAL(4) sqrt(2.)), 'exponential'
    (surface value 1 at z = L3, bottom value 1/ratio at z = 0)."""
    z = np.asarray(z, dtype=np.float64)
    if kind == "constant":
        return np.ones_like(z)
    if kind == "skewed_gaussian":
        from scipy.special import erf
        n02, n12, sig, z0, alpha = (_f32(c) for c in (0.537713935783168, 2.684198470106461, 0.648457170048730,
                                                       6.121537923499139, -5.338431587899242))
        sqrt2 = _f32(np.sqrt(np.float32(2.0)))
        return n12 * np.exp(-((z - z0) ** 2) / (sig ** 2)) * (1.0 + erf(alpha * (z - z0) / (sig * sqrt2))) + n02
    if kind == "exponential":
        return np.exp(math.log(ratio) * (z - TWOPI) / TWOPI)
    raise ValueError(kind)


@dataclasses.dataclass
class QGParams:
    n1: int
    n3: int
    dt: float
    gamma: float
    Ro: float
    Bu: float
    W2F: float
    nuz: float
    flags: int
    ifac_q: np.ndarray
    ifac_w: np.ndarray
    N2u: np.ndarray
    N2s: np.ndarray
    stratification: str = "constant"

    @property
    def n2(self) -> int:
        return self.n1

    @property
    def dz(self) -> float:
        return TWOPI / self.n3


def make_params(n1: int, n3: int, *, dt: float = 1e-3, gamma32: float = 1e-3, Ro: float = 0.1, Bu: float = 1.0,
                W2F: float = 1e-2, coeff1: float = 0.0, ilap1: int = 4, coeff2: float = 10.0, ilap2: int = 4,
                coeff1w: float = 0.0, ilap1w: int = 4, coeff2w: float = 10.0, ilap2w: int = 4,
                coeffz: float = 0.0, flags: int = 0, stratification: str = "constant",
                n2_ratio: float = 100.0) -> QGParams:
    """Parameters for one member.  ``gamma32`` is a REAL(4) literal as in the
    reference (``gamma = 1e-3`` is 0.0010000000474974513 in FP64)."""
    nu1 = nuh_coeff(coeff1, n1, ilap1)
    nu2 = nuh_coeff(coeff2, n1, ilap2)
    nu1w = nuh_coeff(coeff1w, n1, ilap1w)
    nu2w = nuh_coeff(coeff2w, n1, ilap2w)
    a = np.float32(np.float32(64.0) / np.float32(n1))
    nuz = float(coeffz) * float(np.float32(a ** np.float32(np.float32(4.0) / np.float32(3.0))))
    dz = TWOPI / n3
    za = TWOPI * np.arange(1, n3 + 1, dtype=np.float64) / n3
    zas = za - dz / 2
    return QGParams(
        n1=n1, n3=n3, dt=float(dt), gamma=_f32(gamma32), Ro=float(Ro), Bu=float(Bu), W2F=float(W2F),
        nuz=nuz, flags=int(flags),
        ifac_q=int_factor(n1, dt, nu1, ilap1, nu2, ilap2),
        ifac_w=int_factor(n1, dt, nu1w, ilap1w, nu2w, ilap2w),
        N2u=n2_profile(za, stratification, n2_ratio), N2s=n2_profile(zas, stratification, n2_ratio),
        stratification=stratification)


# The workloads named by the benchmark configurations
def config_plain_small(**kw) -> QGParams:
    """64^2 x 32, constant N^2, Ro ~ 0.1, Bu ~ 1, dt = 1e-3, gamma = 1e-3,
    del^8 hyperdiffusion."""
    return make_params(64, 32, **kw)


def config_plain_production(**kw) -> QGParams:
    """256^2 x 128, exponential N^2 (surface/bottom ratio 100)."""
    kw.setdefault("stratification", "exponential")
    return make_params(256, 128, **kw)
