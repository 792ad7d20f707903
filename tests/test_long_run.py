"""1000-step parity of the energies and enstrophy (64^2 x 32, feedback on).

The oracle's trajectory summary is a committed fixture
(tests/golden/long_run_64x32.npz, written by tests/golden/make_long_run.py).
CPU: the fixture is what the oracle produces (first 100 steps recomputed).
GPU: the library's KE, PE, WKE, WPE, WCE and enstrophy after every 100 steps
agree with the oracle's to 1e-9 relative, and psi-hat at the bottom and top
levels after 1000 steps to 1e-9 relative L2.
"""
import os

import numpy as np
import pytest

import qgybj
from helpers import physical_ics, rel_l2

HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURE = os.path.join(HERE, "golden", "long_run_64x32.npz")


def _golden():
    import importlib.util
    spec = importlib.util.spec_from_file_location("make_long_run", os.path.join(HERE, "golden", "make_long_run.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_fixture_is_the_oracle_trajectory():
    gold = np.load(FIXTURE)
    mk = _golden()
    d, _ = mk.run(nsteps=100, every=100)
    np.testing.assert_allclose(d, gold["diag"][:2], rtol=1e-13, atol=0)


@pytest.mark.gpu
def test_thousand_steps_energy_and_enstrophy():
    gold = np.load(FIXTURE)
    mk = _golden()
    P = mk.params()
    psi, BR, BI = physical_ics(P, k0=3, seed=2)
    m = qgybj.Member(P)
    m.set_state_physical(psi, BR, BI)
    diags = [m.diag()]
    for _ in range(mk.NSTEPS // mk.EVERY):
        m.step(mk.EVERY)
        diags.append(m.diag())
    diags = np.array(diags)
    # the wave potential energy is exactly zero at step 0 (A = 0)
    ref = gold["diag"]
    rel = np.abs(diags - ref) / np.where(ref == 0, 1.0, np.abs(ref))
    print("max relative diagnostic difference per 100 steps:", rel.max(axis=1))
    assert rel.max() <= 1e-9, rel
    st = m.get_state()
    for name, lev in (("psik_bottom", 0), ("psik_top", -1)):
        r = rel_l2(st["psik"][lev], gold[name])
        print(name, r)
        assert r <= 1e-9, (name, r)
    m.close()
