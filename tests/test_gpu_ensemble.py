"""Batches, ensembles and the tuner against single-member runs.

A member computed inside a batch (qg_create_batch) or advanced by an
ensemble's captured graph runs exactly the kernels of a single-member
qg_step, so the states must agree bit for bit; single-member runs are in
turn checked against the oracle elsewhere (and here once, for a batch)."""
import numpy as np
import pytest

import qgybj
from helpers import FIELDS, compare_state, oracle_params, physical_ics
from oracle import qg_oracle as O
from qgybj.params import make_params

pytestmark = pytest.mark.gpu


def _ics(P, seeds, k0=3):
    out = [physical_ics(P, k0=k0, seed=s) for s in seeds]
    return [np.stack([o[i] for o in out]) for i in range(3)]


def _equal_states(a, b, what):
    for f in FIELDS:
        assert np.array_equal(a[f], b[f]), (what, f, np.max(np.abs(a[f] - b[f])))


def test_batch_equals_single_members_and_oracle():
    P = make_params(32, 8, W2F=0.05, dt=2e-3, stratification="exponential")
    seeds = [1, 2, 3]
    psi, BR, BI = _ics(P, seeds)
    b = qgybj.Member(P, members=3)
    b.set_state_physical(psi, BR, BI)
    b.step(6)
    sb = b.get_state()
    db = b.diag()
    assert db.shape == (3, 6)
    for m, s in enumerate(seeds):
        one = qgybj.Member(P)
        one.set_state_physical(psi[m], BR[m], BI[m])
        one.step(6)
        so = one.get_state()
        _equal_states({f: sb[f][m] for f in FIELDS}, so, f"member {m}")
        assert np.array_equal(db[m], one.diag())
        one.close()
    g = O.Grid(32, 32, 8)
    OP = oracle_params(P)
    S = O.run(g, OP, O.initial_state(g, OP, psi[1], BR[1], BI[1]), 6)
    compare_state({f: sb[f][1] for f in FIELDS}, S, 1e-11)
    b.close()


@pytest.mark.parametrize("nsteps", [1, 5, 13])
def test_ensemble_graph_equals_eager(nsteps):
    """Mixed sizes, a batch, both A placements; 13 steps = eager start, one
    8-step graph, a 2-step remainder graph; then eager steps continue."""
    cfg = [(make_params(32, 8, W2F=0.05, dt=2e-3), 2, [0, 0, 0, 0, 0, 3, 0]),
           (make_params(64, 16, W2F=0.05, dt=2e-3, stratification="exponential"), 1, [1, 2, 1, 2, 1, 0, 0]),
           (make_params(16, 40, W2F=0.05, dt=2e-3), 1, [0, 0, 0, 0, 0, 1, 1])]
    ens_m, ref_m = [], []
    for i, (P, nm, var) in enumerate(cfg):
        ics = _ics(P, [10 + i + j for j in range(nm)]) if nm > 1 else physical_ics(P, k0=3, seed=10 + i)
        for lst in (ens_m, ref_m):
            m = qgybj.Member(P, members=nm)
            m.set_variants(var)
            m.set_state_physical(*ics)
            lst.append(m)
    ens = qgybj.Ensemble(ens_m)
    ens.step(nsteps)
    for m in ref_m:
        m.step(nsteps)
    for i, (a, b) in enumerate(zip(ens_m, ref_m)):
        assert a.iter == b.iter == nsteps
        _equal_states(a.get_state(), b.get_state(), f"context {i} after {nsteps}")
    # a second ensemble call reuses the graphs; then plain steps continue
    ens.step(nsteps)
    for m in ref_m:
        m.step(nsteps)
    for a in ens_m:
        a.step(2)
    for b in ref_m:
        b.step(2)
    for i, (a, b) in enumerate(zip(ens_m, ref_m)):
        _equal_states(a.get_state(), b.get_state(), f"context {i} continued")
    ens.close()
    for m in ens_m + ref_m:
        m.close()


def test_ensemble_recaptures_after_variant_change():
    P = make_params(32, 8, W2F=0.05, dt=2e-3)
    ics = physical_ics(P, k0=3, seed=4)
    a, b = qgybj.Member(P), qgybj.Member(P)
    for m in (a, b):
        m.set_state_physical(*ics)
    ens = qgybj.Ensemble([a])
    ens.step(10)
    b.step(10)
    a.set_variants([1, 3, 1, 3, 1, 2, 1])
    b.set_variants([1, 3, 1, 3, 1, 2, 1])
    ens.step(10)
    b.step(10)
    _equal_states(a.get_state(), b.get_state(), "after variant change")
    ens.close()
    a.close()
    b.close()


def test_tuner_leaves_state_unchanged_and_run_correct():
    P = make_params(64, 16, W2F=0.05, dt=2e-3)
    psi, BR, BI = physical_ics(P, k0=3, seed=5)
    m = qgybj.Member(P)
    m.set_state_physical(psi, BR, BI)
    m.step(4)
    before = m.get_state()
    best, ms = m.tune(steps=3)
    assert ms > 0 and len(best) == qgybj.NVARIANTS
    assert list(m.get_variants()) == list(best)
    assert m.iter == 4
    _equal_states(m.get_state(), before, "after tuning")
    m.step(4)
    g = O.Grid(64, 64, 16)
    OP = oracle_params(P)
    S = O.run(g, OP, O.initial_state(g, OP, psi, BR, BI), 8)
    compare_state(m.get_state(), S, 1e-11)
    m.close()


def test_tuner_from_initial_state():
    P = make_params(32, 8, W2F=0.05, dt=2e-3)
    psi, BR, BI = physical_ics(P, k0=3, seed=6)
    m = qgybj.Member(P)
    m.set_state_physical(psi, BR, BI)
    m.tune(steps=2)
    assert m.iter == 0
    m.step(5)
    g = O.Grid(32, 32, 8)
    OP = oracle_params(P)
    S = O.run(g, OP, O.initial_state(g, OP, psi, BR, BI), 5)
    compare_state(m.get_state(), S, 1e-11)
    m.close()
