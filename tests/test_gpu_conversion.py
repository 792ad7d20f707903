"""Kept tendencies and wave energy conversion terms through the C ABI:

* qg_get_tendencies returns the last step's nBRk, nBIk (convol_waqg,
  derivatives.f90:607-803) and rBRk, rBIk (refraction_waqg :810-900), as the
  reference holds them after the step, equal to the oracle's to 1e-11
  (relative L2) after the Euler step and after leapfrog steps, under the
  special-case flags and the regular YBJ operator;
* they reproduce the update the step made (main_waqg.f90:375-377) from the
  state before and after it;
* qg_wave_conversion (we_conversion diagnostics.f90:967-1076,
  gamma_conversion :1080-1256), formed on the device by Parseval, equals the
  oracle's physical-space sums to 1e-10;
* keeping tendencies leaves the trajectory bit-identical, batch members equal
  single members bit for bit, ensemble graph replays keep them too, and the
  calls refuse without kept tendencies.
"""
import numpy as np
import pytest

import qgybj
from helpers import oracle_params, physical_ics, rel_l2
from oracle import qg_oracle as O
from qgybj.params import make_params

pytestmark = pytest.mark.gpu

NAMES = ("nBRk", "nBIk", "rBRk", "rBIk")


def _cmp_tend(got, want, tol):
    for name, a, b in zip(NAMES, got, want):
        if np.linalg.norm(b) == 0:
            assert np.max(np.abs(a)) == 0, name
        else:
            assert rel_l2(a, b) <= tol, (name, rel_l2(a, b))


@pytest.mark.parametrize("flags,dt", [
    (0, 2e-3),
    (qgybj.params.LINEAR, 2e-3),
    (qgybj.params.PASSIVE_SCALAR, 2e-3),
    (qgybj.params.NO_FEEDBACK | qgybj.params.NO_DISPERSION, 2e-3),
    (qgybj.params.YBJ_REGULAR, 2e-5),
])
def test_tendencies_and_conversion_match_oracle(flags, dt):
    n1, n3 = 32, 16
    P = make_params(n1, n3, dt=dt, W2F=0.05, stratification="exponential", flags=flags)
    g = O.Grid(n1, n1, n3)
    OP = oracle_params(P)
    psi, BR, BI = physical_ics(P, k0=3, seed=4)
    S = O.initial_state(g, OP, psi, BR, BI)
    m = qgybj.Member(P)
    m.set_state_physical(psi, BR, BI)
    m.keep_tendencies(True)
    with pytest.raises(qgybj.QGError):
        m.tendencies()  # no step yet
    for it in range(6):
        S = O.step(g, OP, S)
        m.step(1)
        if it in (0, 1, 5):
            _cmp_tend(m.tendencies(), S.tend, 1e-11)
            want = O.wave_conversion(g, OP, S)
            got = m.wave_conversion()
            if flags & qgybj.params.PASSIVE_SCALAR:
                # A = 0: every term vanishes
                assert np.all(got == 0) and np.all(want == 0)
            else:
                np.testing.assert_allclose(got, want, rtol=1e-10, atol=1e-13 * np.abs(want).max())
    m.close()


@pytest.mark.parametrize("n1,n3", [(32, 8), (256, 16)])
def test_tendencies_reproduce_the_update(n1, n3):
    """B^{n+1} = L [B^{n-1} e^{-2fw} - 2 dt (nBR + c kh2 AI^n - rBI/2) e^{-fw}]
    (main_waqg.f90:376) from the states around one leapfrog step, and the
    separated refraction terms equal the oracle's refraction_waqg of the state
    before the step (1e-11), also at the production width n1 = 256."""
    P = make_params(n1, n3, dt=2e-3, W2F=0.05)
    psi, BR, BI = physical_ics(P, k0=3, seed=6)
    m = qgybj.Member(P)
    m.set_state_physical(psi, BR, BI)
    m.keep_tendencies(True)
    m.step(3)
    before = m.get_state()
    m.step(1)
    after = m.get_state()
    nBR, nBI, rBR, rBI = m.tendencies()
    g = O.Grid(n1, n1, n3)
    L = g.Lb
    wR, wI = O.refraction_waqg(g, before["BRk"], before["BIk"], before["psik"])
    assert rel_l2(rBR, wR) < 1e-11 and rel_l2(rBI, wI) < 1e-11
    kh2 = g.kh2.astype(np.float64)
    c = 0.5 / (P.Bu * P.Ro)
    e1, e2 = np.exp(-P.ifac_w), np.exp(-2 * P.ifac_w)
    tdt = 2 * P.dt
    BRn = before["BRok"] * e2 - (tdt * ((nBR + (c * kh2) * before["AIk"]) - 0.5 * rBI)) * e1
    BIn = before["BIok"] * e2 - (tdt * ((nBI - (c * kh2) * before["ARk"]) + 0.5 * rBR)) * e1
    assert rel_l2(np.where(L, BRn, 0), after["BRk"]) < 1e-13
    assert rel_l2(np.where(L, BIn, 0), after["BIk"]) < 1e-13
    m.close()


def test_keeping_tendencies_leaves_the_run_unchanged():
    P = make_params(64, 16, dt=2e-3, W2F=0.05, stratification="exponential")
    ic = physical_ics(P, k0=3, seed=2)
    a, b = qgybj.Member(P), qgybj.Member(P)
    for m in (a, b):
        m.set_state_physical(*ic)
    a.keep_tendencies(True)
    a.step(4)
    b.step(4)
    sa, sb = a.get_state(), b.get_state()
    for f in sa:
        assert np.array_equal(sa[f], sb[f]), f
    a.keep_tendencies(False)
    with pytest.raises(qgybj.QGError):
        a.wave_conversion()
    a.step(3)
    b.step(3)
    sa, sb = a.get_state(), b.get_state()
    for f in sa:
        assert np.array_equal(sa[f], sb[f]), f
    # on again: nothing kept until the next step; a new state discards them
    a.keep_tendencies(True)
    with pytest.raises(qgybj.QGError):
        a.tendencies()
    a.step(1)
    a.tendencies()
    a.set_state_physical(*ic)
    with pytest.raises(qgybj.QGError):
        a.wave_conversion()
    a.close()
    b.close()


def test_batch_members_and_ensemble_replays():
    P = make_params(32, 8, dt=2e-3, W2F=0.05)
    ics = [physical_ics(P, k0=3, seed=s) for s in (3, 9)]
    bt = qgybj.Member(P, members=2)
    bt.set_state_physical(*[np.stack([x[i] for x in ics]) for i in range(3)])
    bt.keep_tendencies(True)
    bt.step(5)
    tb = bt.tendencies()
    cb = bt.wave_conversion()
    assert cb.shape == (2, 5)
    singles = []
    for j, ic in enumerate(ics):
        one = qgybj.Member(P)
        one.set_state_physical(*ic)
        one.keep_tendencies(True)
        one.step(5)
        t1 = one.tendencies()
        for a, b in zip(t1, tb):
            assert np.array_equal(a, b[j])
        assert np.array_equal(one.wave_conversion(), cb[j])
        singles.append(one)
    # the same two members stepped from graphs: 5 + 12 steps equal eager steps
    ens = qgybj.Ensemble(singles)
    ens.set_graph_steps(4)
    ens.step(12)
    bt.step(12)
    tb = bt.tendencies()
    cb = bt.wave_conversion()
    for j, one in enumerate(singles):
        for a, b in zip(one.tendencies(), tb):
            assert np.array_equal(a, b[j])
        assert np.array_equal(one.wave_conversion(), cb[j])
    ens.close()
    for one in singles:
        one.close()
    bt.close()
