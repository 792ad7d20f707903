"""Tangent-linear and adjoint of the step (qg_adj_*), through the C ABI.

- the one-step tangent-linear map equals the central difference of two
  forward runs: the step is quadratic in psi0, so the difference is exact up
  to rounding;
- dot-product test: <TLM dpsi, sbar> = <dpsi, ADJ sbar> over several steps and
  checkpoint segments, to 1e-12;
- the gradient matches central finite differences of the objective along
  random directions to 1e-6, does not depend on the checkpoint interval, and
  agrees with the tangent-linear dJ;
- the torch.autograd.Function harness returns the same gradient.
"""
import numpy as np
import pytest

import qgybj
from helpers import physical_ics
from qgybj.params import make_params

pytestmark = pytest.mark.gpu


def _setup(n1=32, n3=16, seed=3, dt=2e-3, W2F=0.05, flags=0):
    P = make_params(n1, n3, dt=dt, W2F=W2F, flags=flags, stratification="exponential")
    psi, BR, BI = physical_ics(P, k0=3, seed=seed)
    return P, psi, BR, BI


def _spec_dot(P, a, b):
    L = qgybj.mask(P.n1).astype(bool)  # (ikty, iktx)
    return float(np.sum(np.real(np.conj(a) * b)[..., L]))


def _fwd_state(P, psi, BR, BI, nsteps):
    m = qgybj.Member(P)
    m.set_state_physical(psi, BR, BI)
    m.step(nsteps)
    st = m.get_state()
    m.close()
    return np.stack([st[f] for f in qgybj.Member.FIELDS])


@pytest.mark.parametrize("flags", [0, qgybj.params.NO_FEEDBACK])
def test_one_step_tlm_is_central_difference(flags):
    P, psi, BR, BI = _setup(flags=flags)
    rng = np.random.default_rng(1)
    dpsi = rng.standard_normal(psi.shape)
    m = qgybj.Member(P)
    adj = qgybj.Adjoint(m, checkpoint_every=4)
    _, dst = adj.tlm(psi, BR, BI, 1, dpsi, want_state=True)
    s = 1e-3 * np.linalg.norm(psi) / np.linalg.norm(dpsi)
    cd = (_fwd_state(P, psi + s * dpsi, BR, BI, 1) - _fwd_state(P, psi - s * dpsi, BR, BI, 1)) / (2 * s)
    for i, f in enumerate(qgybj.Member.FIELDS):
        ref = cd[i]
        if np.linalg.norm(ref) == 0:
            assert np.linalg.norm(dst[i]) == 0, f
            continue
        rel = np.linalg.norm(dst[i] - ref) / np.linalg.norm(ref)
        print(f, rel)
        assert rel < 1e-9, (f, rel)
    adj.close()
    m.close()


@pytest.mark.parametrize("nsteps,K", [(1, 1), (5, 2), (7, 7)])
def test_dot_product(nsteps, K):
    P, psi, BR, BI = _setup(seed=4)
    rng = np.random.default_rng(10 + nsteps)
    dpsi = rng.standard_normal(psi.shape)
    m = qgybj.Member(P)
    adj = qgybj.Adjoint(m, checkpoint_every=K)
    L = qgybj.mask(P.n1).astype(bool)
    sbar = (rng.standard_normal((9,) + m.shape) + 1j * rng.standard_normal((9,) + m.shape)) * L
    _, dst = adj.tlm(psi, BR, BI, nsteps, dpsi, want_state=True)
    lhs = _spec_dot(P, dst, sbar)
    psibar = adj.apply(psi, BR, BI, nsteps, sbar)
    rhs = float(np.sum(dpsi * psibar))
    print(nsteps, K, lhs, rhs, abs(lhs - rhs) / abs(lhs))
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs)
    adj.close()
    m.close()


def test_gradient_matches_finite_differences():
    P, psi, BR, BI = _setup(seed=5)
    nsteps = 20
    m = qgybj.Member(P)
    adj = qgybj.Adjoint(m, checkpoint_every=6)
    J, g = adj.gradient(psi, BR, BI, nsteps)
    assert J > 0
    rng = np.random.default_rng(7)
    for trial in range(3):
        d = rng.standard_normal(psi.shape)
        d /= np.linalg.norm(d)
        eps = 1e-4 * np.linalg.norm(psi)
        fd = (adj.objective(psi + eps * d, BR, BI, nsteps) - adj.objective(psi - eps * d, BR, BI, nsteps)) / (2 * eps)
        an = float(np.sum(g * d))
        print(trial, fd, an, abs(fd - an) / abs(an))
        assert abs(fd - an) <= 1e-6 * abs(an)
    adj.close()
    m.close()


def test_gradient_independent_of_checkpoint_interval_and_matches_tlm():
    P, psi, BR, BI = _setup(seed=6)
    nsteps = 9
    m = qgybj.Member(P)
    grads = []
    for K in (1, 4, 9, 20):
        adj = qgybj.Adjoint(m, checkpoint_every=K)
        J, g = adj.gradient(psi, BR, BI, nsteps)
        grads.append((J, g))
        adj.close()
    for J, g in grads[1:]:
        assert J == grads[0][0]
        assert np.array_equal(g, grads[0][1])
    adj = qgybj.Adjoint(m, checkpoint_every=4)
    d = np.random.default_rng(3).standard_normal(psi.shape)
    dJ, _ = adj.tlm(psi, BR, BI, nsteps, d)
    an = float(np.sum(grads[0][1] * d))
    assert abs(dJ - an) <= 1e-12 * abs(an), (dJ, an)
    # the member is left at the final state of the last run
    ref = _fwd_state(P, psi, BR, BI, nsteps)
    st = m.get_state()
    assert np.array_equal(np.stack([st[f] for f in qgybj.Member.FIELDS]), ref)
    adj.close()
    m.close()


def test_torch_autograd_harness():
    torch = pytest.importorskip("torch")
    from qgybj.autograd import wave_energy
    P, psi, BR, BI = _setup(seed=8)
    nsteps = 6
    m = qgybj.Member(P)
    adj = qgybj.Adjoint(m, checkpoint_every=3)
    x = torch.tensor(psi, dtype=torch.float64, requires_grad=True)
    J = wave_energy(x, adj, BR, BI, nsteps)
    (2.0 * J).backward()
    J2, g = adj.gradient(psi, BR, BI, nsteps)
    assert float(J.detach()) == J2
    np.testing.assert_array_equal(x.grad.numpy(), 2.0 * g)
    adj.close()
    m.close()


def test_rejects_unsupported_configurations():
    for flags in (qgybj.params.FIXED_FLOW, qgybj.params.PASSIVE_SCALAR, qgybj.params.LINEAR):
        P = make_params(16, 4, flags=flags)
        m = qgybj.Member(P)
        with pytest.raises(qgybj.QGError):
            qgybj.Adjoint(m)
        m.close()
    m = qgybj.Member(make_params(16, 4, coeffz=1.0))
    with pytest.raises(qgybj.QGError):
        qgybj.Adjoint(m)
    m.close()


def test_batch_gradients_equal_single_members():
    """A batch context gives each member's objective and gradient, exactly as
    single-member runs do."""
    P, _, _, _ = _setup(seed=1)
    ics = [physical_ics(P, k0=3, seed=s) for s in (11, 12, 13)]
    stack = [np.stack([x[i] for x in ics]) for i in range(3)]
    b = qgybj.Member(P, members=3)
    ab = qgybj.Adjoint(b, checkpoint_every=4)
    Jb, gb = ab.gradient(*stack, 9)
    assert Jb.shape == (3,) and gb.shape == (3,) + psi_shape(P)
    for j, (psi, BR, BI) in enumerate(ics):
        m = qgybj.Member(P)
        a = qgybj.Adjoint(m, checkpoint_every=4)
        J, g = a.gradient(psi, BR, BI, 9)
        assert J == Jb[j]
        assert np.array_equal(g, gb[j])
        a.close()
        m.close()
    ab.close()
    b.close()


def psi_shape(P):
    return (P.n3, P.n1, P.n1)


@pytest.mark.parametrize("n1,n3,flags", [(32, 16, 0), (64, 8, qgybj.params.NO_FEEDBACK), (16, 40, 0), (128, 8, 0),
                                          (256, 4, qgybj.params.NO_FEEDBACK)])
def test_fused_adjoint_matches_generic(n1, n3, flags):
    """The fused adjoint kernels and the generic (pointwise + separate
    transforms) adjoint compute the same map (n1 >= 128 runs the row kernel
    with 4 elements per thread)."""
    P, psi, BR, BI = _setup(n1=n1, n3=n3, flags=flags, seed=9)
    nsteps = 7
    m = qgybj.Member(P)
    adj = qgybj.Adjoint(m, checkpoint_every=3)
    adj.set_fused(False)
    J0, g0 = adj.gradient(psi, BR, BI, nsteps)
    L = qgybj.mask(P.n1).astype(bool)
    rng = np.random.default_rng(5)
    sbar = (rng.standard_normal((9,) + m.shape) + 1j * rng.standard_normal((9,) + m.shape)) * L
    a0 = adj.apply(psi, BR, BI, nsteps, sbar)
    adj.set_fused(True)
    J1, g1 = adj.gradient(psi, BR, BI, nsteps)
    a1 = adj.apply(psi, BR, BI, nsteps, sbar)
    assert J0 == J1
    for x, y in ((g1, g0), (a1, a0)):
        rel = np.linalg.norm(x - y) / np.linalg.norm(y)
        print("fused vs generic", rel)
        assert rel < 1e-12
    adj.close()
    m.close()
