"""Shared test helpers: build matching oracle / library parameters and
initial states."""
import numpy as np

from oracle import qg_oracle as O
from qgybj import ic
from qgybj.params import QGParams, make_params


def oracle_params(P: QGParams) -> O.Params:
    return O.Params(n1=P.n1, n3=P.n3, dt=P.dt, gamma=P.gamma, Ro=P.Ro, Bu=P.Bu, W2F=P.W2F, nuz=P.nuz,
                    flags=P.flags, ifac_q=P.ifac_q, ifac_w=P.ifac_w, a_ell_ut=P.Bu / P.N2u, N2u=P.N2u, N2s=P.N2s)


def physical_ics(P: QGParams, k0: int = 3, seed: int = 1, structured_waves: bool = True):
    psi = ic.eddy_psi(P.n1, P.n3, k0, seed, Bu=P.Bu)
    if structured_waves:
        BR, BI = ic.perturbed_waves(P.n1, P.n3, seed)
    else:
        BR, BI = ic.wave_b(P.n1, P.n3)
    return psi, BR, BI


def small_params(n1=32, n3=8, **kw) -> QGParams:
    kw.setdefault("W2F", 0.05)
    kw.setdefault("dt", 2e-3)
    return make_params(n1, n3, **kw)


def rel_l2(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    nb = np.linalg.norm(b.ravel())
    d = np.linalg.norm((a - b).ravel())
    return d / nb if nb > 0 else d


FIELDS = ["psik", "qk", "qok", "BRk", "BIk", "BRok", "BIok", "ARk", "AIk"]


def compare_state(got: dict, S: O.State, tol: float, fields=FIELDS):
    """Relative L2 of every state field (all modes) against the oracle."""
    worst = {}
    for f in fields:
        ref = getattr(S, f)
        r = rel_l2(got[f], ref)
        worst[f] = r
        if np.linalg.norm(ref) == 0:
            assert np.max(np.abs(got[f])) <= tol, (f, np.max(np.abs(got[f])))
        else:
            assert r <= tol, (f, r)
    return worst
