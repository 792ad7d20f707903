"""Device diagnostics through the C ABI against the oracle's restatements:

* wave spectra (qg_hspec_waves, diagnostics.f90:547-638) at the reference's
  spectrum heights (parameters_test1.f90:188): shell counts exactly, the
  kinetic and potential parts to 1e-12; deterministic (bitwise equal on
  repeat); batch members equal single members bit for bit; more levels than
  one launch takes; bad levels rejected;
* the vertically averaged wave energy density (qg_wave_vave, we_vave
  diagnostics.f90:750-813) to 1e-12, batch members equal single members;
* the omega equation (qg_omega, derivatives.f90:1469-1586 and
  elliptic.f90:463-567) on a running state to 1e-11; the run continues
  bit-identically; batch members equal single members.
"""
import numpy as np
import pytest

import qgybj
from helpers import oracle_params, physical_ics
from oracle import qg_oracle as O
from qgybj.params import make_params

pytestmark = pytest.mark.gpu


def _heights(n3):
    return [1, n3 // 8, n3 // 4, 3 * n3 // 8, n3 // 2, 5 * n3 // 8, 3 * n3 // 4, 7 * n3 // 8, n3 - 2, n3]


def _oracle_state(st):
    z = np.zeros_like(st["psik"])
    return O.State(psik=st["psik"], qk=st["qk"], qok=st["qok"], BRk=st["BRk"], BIk=st["BIk"], BRok=st["BRok"],
                   BIok=st["BIok"], ARk=st["ARk"], AIk=st["AIk"], CRk=z, CIk=z)


@pytest.mark.parametrize("n1,n3", [(64, 32), (128, 16)])
def test_hspec_waves_matches_oracle(n1, n3):
    P = make_params(n1, n3, dt=2e-3, W2F=0.05, stratification="exponential")
    psi, BR, BI = physical_ics(P, k0=3, seed=5)
    m = qgybj.Member(P)
    m.set_state_physical(psi, BR, BI)
    m.step(7)
    hs = _heights(n3)
    got = m.hspec_waves(hs)
    assert got.shape == (len(hs), n1 // 2, 3)
    g = O.Grid(n1, n1, n3)
    OP = oracle_params(P)
    S = _oracle_state(m.get_state())
    for i, h in enumerate(hs):
        want = O.hspec_waves(g, OP, S, h)
        assert np.array_equal(got[i, :, 2], want[:, 2]), h
        for j in range(2):
            np.testing.assert_allclose(got[i, :, j], want[:, j], rtol=1e-12, atol=1e-14 * np.abs(want[:, j]).max())
        assert want[:, 1].max() > 0 and want[:, 0].max() > 0
    assert np.array_equal(m.hspec_waves(hs), got)
    for bad in ([0], [n3 + 1], []):
        with pytest.raises(qgybj.QGError):
            m.hspec_waves(bad)
    m.close()


def test_hspec_waves_batch_and_many_levels():
    P = make_params(32, 40, dt=2e-3, W2F=0.05)
    ics = [physical_ics(P, k0=3, seed=s) for s in (7, 8)]
    b = qgybj.Member(P, members=2)
    b.set_state_physical(*[np.stack([x[i] for x in ics]) for i in range(3)])
    b.step(5)
    levels = list(range(1, 41))  # two launches of at most 32 levels
    gb = b.hspec_waves(levels)
    assert gb.shape == (2, 40, 16, 3)
    for j, ic in enumerate(ics):
        one = qgybj.Member(P)
        one.set_state_physical(*ic)
        one.step(5)
        assert np.array_equal(one.hspec_waves(levels), gb[j])
        one.close()
    b.close()


@pytest.mark.parametrize("n1,n3,strat", [(64, 16, "exponential"), (32, 8, "constant")])
def test_omega_matches_oracle_and_leaves_the_run_unchanged(n1, n3, strat):
    """qg_omega (omega equation, elliptic.f90:463-567) against the oracle on a
    running state; a twin member that never calls it stays bit-identical."""
    P = make_params(n1, n3, dt=2e-3, W2F=0.05, stratification=strat)
    ics = physical_ics(P, k0=3, seed=6)
    m, twin = qgybj.Member(P), qgybj.Member(P)
    for x in (m, twin):
        x.set_state_physical(*ics)
        x.step(4)
    got = m.omega()
    g = O.Grid(n1, n1, n3)
    want = O.omega_equation(g, P.Bu / P.N2u, O.omega_eqn_rhs(g, m.get_state()["psik"]))
    assert np.linalg.norm(want) > 0
    assert np.linalg.norm(got - want) <= 1e-11 * np.linalg.norm(want), np.linalg.norm(got - want) / np.linalg.norm(want)
    assert np.all(got[-1] == 0)
    sel = g.Lb & (g.kh2 > 0)
    assert np.all(got[:, ~sel] == 0)
    m.step(3)
    twin.step(3)
    a, b = m.get_state(), twin.get_state()
    for f in a:
        assert np.array_equal(a[f], b[f]), f
    m.close()
    twin.close()


def test_omega_batch_members():
    P = make_params(32, 8, dt=2e-3, W2F=0.05)
    ics = [physical_ics(P, k0=3, seed=s) for s in (9, 10)]
    b = qgybj.Member(P, members=2)
    b.set_state_physical(*[np.stack([x[i] for x in ics]) for i in range(3)])
    b.step(3)
    wb = b.omega()
    for j, ic in enumerate(ics):
        one = qgybj.Member(P)
        one.set_state_physical(*ic)
        one.step(3)
        assert np.array_equal(one.omega(), wb[j])
        one.close()
    b.close()


def test_we_vave_matches_oracle_and_batch():
    """qg_wave_vave (we_vave, diagnostics.f90:750-813) on a running state
    against the oracle to 1e-12; batch members equal single members bit for
    bit; the run continues unchanged."""
    P = make_params(64, 16, dt=2e-3, W2F=0.05, stratification="exponential")
    ics = [physical_ics(P, k0=3, seed=s) for s in (5, 6)]
    g = O.Grid(64, 64, 16)
    b = qgybj.Member(P, members=2)
    b.set_state_physical(*[np.stack([x[i] for x in ics]) for i in range(3)])
    b.step(6)
    vb = b.wave_vave()
    assert vb.shape == (2, 64, 64)
    for j, ic in enumerate(ics):
        one, twin = qgybj.Member(P), qgybj.Member(P)
        for m in (one, twin):
            m.set_state_physical(*ic)
            m.step(6)
        v = one.wave_vave()
        assert np.array_equal(v, vb[j])
        want = O.we_vave(g, _oracle_state(one.get_state()))
        np.testing.assert_allclose(v, want, rtol=1e-12, atol=1e-14 * want.max())
        one.step(3)
        twin.step(3)
        s1, s2 = one.get_state(), twin.get_state()
        for f in s1:
            assert np.array_equal(s1[f], s2[f]), f
        one.close()
        twin.close()
    b.close()
