"""GPU parity: the HIP library (through the C ABI) against the oracle on the
same seeded inputs.  Tolerances: transforms 1e-13 relative (FP64 rounding of
different but exact algorithms); time steps 1e-11 relative L2 after 10 steps
(the parity bar of the reference comparison); energy and enstrophy 1e-9."""
import numpy as np
import pytest

from oracle import qg_oracle as O
import qgybj
from qgybj.params import make_params
from helpers import FIELDS, compare_state, oracle_params, physical_ics, rel_l2, small_params

pytestmark = pytest.mark.gpu


def _member(P):
    return qgybj.Member(P)


@pytest.mark.parametrize("n,nz", [(16, 3), (32, 4), (64, 32), (128, 6), (256, 4), (512, 2)])
def test_r2c_matches_oracle(n, nz):
    P = make_params(n, max(nz, 2))
    m = _member(P)
    rng = np.random.default_rng(n + nz)
    r = rng.standard_normal((nz, n, n))
    got = m.r2c(r)
    ref = O.fft_r2c(r)
    assert rel_l2(got, ref) < 1e-14
    m.close()


@pytest.mark.parametrize("n,nz", [(16, 3), (32, 4), (64, 32), (128, 6), (256, 4), (512, 2)])
def test_c2r_matches_oracle(n, nz):
    """Arbitrary (non-Hermitian) half spectra: realit on kx = 0, imaginary
    parts of the kx = 0 and n/2 entries after the y transform ignored."""
    P = make_params(n, max(nz, 2))
    m = _member(P)
    rng = np.random.default_rng(7 * n + nz)
    k = rng.standard_normal((nz, n, n // 2 + 1)) + 1j * rng.standard_normal((nz, n, n // 2 + 1))
    got = m.c2r(k)
    ref = O.fft_c2r(k)
    assert rel_l2(got, ref) < 1e-14
    # the input is left untouched
    m.close()


def test_c2r_with_padded_leading_dimension():
    n = 32
    m = _member(make_params(n, 4))
    rng = np.random.default_rng(3)
    k = rng.standard_normal((4, n, n // 2 + 1)) + 1j * rng.standard_normal((4, n, n // 2 + 1))
    got = m.c2r(k, ldx=n + 2)
    assert rel_l2(got, O.fft_c2r(k)) < 1e-14
    m.close()


@pytest.mark.parametrize("kind", ["constant", "exponential"])
def test_init_q_matches_oracle(kind):
    P = make_params(64, 16, stratification=kind)
    m = _member(P)
    g = O.Grid(64, 64, 16)
    rng = np.random.default_rng(5)
    psi = rng.standard_normal((16, 64, 33)) + 1j * rng.standard_normal((16, 64, 33))
    got = m.init_q(psi)
    ref = O.init_q(g, P.Bu / P.N2u, psi)
    assert rel_l2(got, ref) < 1e-14
    m.close()


def _run_pair(P, nsteps, k0=3, seed=1, structured=True, check_every=False, tol=1e-11):
    g = O.Grid(P.n1, P.n1, P.n3)
    OP = oracle_params(P)
    psi, BR, BI = physical_ics(P, k0=k0, seed=seed, structured_waves=structured)
    S = O.initial_state(g, OP, psi, BR, BI)
    m = _member(P)
    m.set_state_physical(psi, BR, BI)
    got = m.get_state()
    # r2c rounding (1e-16) amplified by kh2 in init_q for q
    compare_state(got, S, 1e-12)
    worst = {}
    for it in range(nsteps):
        S = O.step(g, OP, S)
        m.step(1)
        if check_every or it == nsteps - 1:
            worst = compare_state(m.get_state(), S, tol)
    return m, S, g, OP, worst


def test_first_steps_64x32_feedback_on():
    """Euler step, first leapfrog step and the one after, each checked."""
    P = make_params(64, 32, W2F=0.05, dt=2e-3)
    m, S, g, OP, worst = _run_pair(P, 3, check_every=True)
    print("worst rel L2 after 3 steps:", worst)
    m.close()


def test_ten_steps_64x32_config1():
    """The CPU-runnable configuration: 64^2 x 32, constant N^2, dt = 1e-3,
    gamma = 1e-3 (REAL(4)), del^8 hyperdiffusion, eddies peaked at |k| = 3."""
    P = make_params(64, 32, dt=1e-3, W2F=0.01)
    m, S, g, OP, worst = _run_pair(P, 10, k0=3, seed=2)
    print("worst rel L2 after 10 steps:", worst)
    d_gpu = m.diag()
    d_ref = O.diagnostics(g, OP, S)
    np.testing.assert_allclose(d_gpu, d_ref, rtol=1e-12, atol=1e-300)
    m.close()


@pytest.mark.parametrize("flags", [
    qgybj.params.FIXED_FLOW, qgybj.params.NO_FEEDBACK, qgybj.params.PASSIVE_SCALAR, qgybj.params.LINEAR,
    qgybj.params.NO_DISPERSION, qgybj.params.INVISCID,
    qgybj.params.FIXED_FLOW | qgybj.params.PASSIVE_SCALAR,
])
def test_flag_variants(flags):
    P = small_params(32, 8, flags=flags)
    m, S, g, OP, worst = _run_pair(P, 4, check_every=True)
    m.close()


@pytest.mark.parametrize("n1,n3,flags,strat", [
    (32, 16, 0, "constant"),
    (64, 8, qgybj.params.NO_FEEDBACK, "exponential"),
    (32, 8, qgybj.params.LINEAR, "exponential"),
    (32, 8, qgybj.params.NO_DISPERSION, "constant"),
    (32, 8, qgybj.params.PASSIVE_SCALAR, "constant"),
])
def test_regular_ybj_operator(n1, n3, flags, strat):
    """ybj_plus = 0 (QG_YBJ_REGULAR): sumB, compute_sigma and compute_A
    (derivatives.f90:1037-1231) in place of the YBJ+ inversion, every state
    field checked after every step, including the Euler step and the steps
    whose B planes are re-formed from the mean-free B.  The regular operator
    is only conditionally stable (dispersion frequency kh^2 N^2 / (2 Bu Ro m^2)
    for vertical wavenumber m: the leapfrog needs a far smaller dt than
    YBJ+, the oracle itself overflows at dt = 2e-3 on these grids), so the
    runs use dt = 2e-5."""
    P = make_params(n1, n3, dt=2e-5, W2F=0.05, stratification=strat, flags=flags | qgybj.params.YBJ_REGULAR)
    m, S, g, OP, worst = _run_pair(P, 10, check_every=True)
    print("worst rel L2 after 6 steps:", worst)
    if not flags & qgybj.params.PASSIVE_SCALAR:
        # B stays mean-free in z at every L = 1 mode with kh2 > 0
        sel = g.Lb & (g.kh2 > 0)
        st = m.get_state()
        for f in ("BRk", "BIk"):
            assert np.max(np.abs(st[f].sum(axis=0)[sel])) < 1e-12 * np.abs(st[f]).max()
    m.close()


def test_unknown_flag_rejected():
    with pytest.raises(qgybj.QGError):
        qgybj.Member(make_params(16, 4, flags=128))


def test_vertical_diffusion_nuz():
    """nuz != 0 (coeffz > 0) exercises dissipation_q_nv including the
    Euler-step read of the old q."""
    P = small_params(32, 8, coeffz=0.5)
    assert P.nuz > 0
    m, S, g, OP, worst = _run_pair(P, 4, check_every=True)
    m.close()


def test_restart_from_state_mid_run():
    """qg_set_state with iter > 0 continues with leapfrog steps exactly as the
    uninterrupted run."""
    P = small_params(32, 8)
    g = O.Grid(32, 32, 8)
    OP = oracle_params(P)
    psi, BR, BI = physical_ics(P)
    S = O.initial_state(g, OP, psi, BR, BI)
    S = O.run(g, OP, S, 3)
    m = _member(P)
    m.set_state(S.psik, S.qk, S.qok, S.BRk, S.BIk, S.BRok, S.BIok, S.ARk, S.AIk, iter=S.iter)
    S = O.run(g, OP, S, 3)
    m.step(3)
    compare_state(m.get_state(), S, 1e-11)
    assert m.iter == S.iter
    m.close()


def test_set_state_spectral_iter0_matches_physical():
    P = small_params(32, 8)
    g = O.Grid(32, 32, 8)
    OP = oracle_params(P)
    psi, BR, BI = physical_ics(P)
    S = O.initial_state(g, OP, psi, BR, BI)
    m = _member(P)
    m.set_state(S.psik, S.qk, S.qok, S.BRk, S.BIk, S.BRok, S.BIok, S.ARk, S.AIk, iter=0)
    S = O.run(g, OP, S, 2)
    m.step(2)
    compare_state(m.get_state(), S, 1e-11)
    m.close()


def test_c_is_az():
    P = small_params(32, 8)
    m, S, g, OP, worst = _run_pair(P, 2)
    st = m.get_state(with_c=True)
    assert rel_l2(st["CRk"], S.CRk) < 1e-11
    assert rel_l2(st["CIk"], S.CIk) < 1e-11
    m.close()


@pytest.mark.parametrize("n,n3,steps", [(128, 16, 3), (256, 8, 2), (512, 4, 2)])
def test_larger_grids(n, n3, steps):
    P = make_params(n, n3, W2F=0.05, stratification="exponential")
    m, S, g, OP, worst = _run_pair(P, steps, k0=5)
    m.close()


def test_nonfinite_is_reported():
    """A NaN planted in the state (host side) is reported, not propagated
    silently."""
    P = small_params(16, 4)
    g = O.Grid(16, 16, 4)
    OP = oracle_params(P)
    psi, BR, BI = physical_ics(P)
    S = O.initial_state(g, OP, psi, BR, BI)
    S.BRk[1, 2, 3] = np.nan
    m = _member(P)
    m.set_state(S.psik, S.qk, S.qok, S.BRk, S.BIk, S.BRok, S.BIok, S.ARk, S.AIk, iter=0)
    with pytest.raises(qgybj.QGError) as e:
        m.step(1)
    assert e.value.code == 4
    m.close()


@pytest.mark.parametrize("n3", [2, 3, 9, 40, 64, 100, 128, 200, 256])
def test_vertical_partitions(n3):
    """Every chunking of the partitioned vertical solve (full and padded
    chunks, 1 to 16 chunks per column) against the sequential oracle."""
    P = make_params(16, n3, W2F=0.05, dt=2e-3, stratification="exponential")
    m, S, g, OP, worst = _run_pair(P, 3, k0=2, check_every=True)
    m.close()


@pytest.mark.parametrize("n,n3", [(32, 4), (64, 8), (128, 4), (256, 4), (512, 2)])
@pytest.mark.parametrize("variant", [(0, 0, 0, 0, 0), (1, 1, 1, 1, 1), (0, 2, 0, 2, 0), (1, 3, 1, 3, 1),
                                     (0, 0, 2, 0, 2), (1, 1, 3, 1, 3)])
def test_kernel_variants(n, n3, variant):
    """Every transform-kernel variant the tuner can select gives the oracle's
    result."""
    P = make_params(n, n3, W2F=0.05, dt=2e-3)
    g = O.Grid(n, n, n3)
    OP = oracle_params(P)
    psi, BR, BI = physical_ics(P, k0=3)
    S = O.initial_state(g, OP, psi, BR, BI)
    m = _member(P)
    m.set_variants(variant)
    assert list(m.get_variants())[:5] == list(variant)
    m.set_state_physical(psi, BR, BI)
    S = O.run(g, OP, S, 3)
    m.step(3)
    compare_state(m.get_state(), S, 1e-11)
    m.close()


@pytest.mark.parametrize("n3", [2, 3, 40, 128, 300, 520])
@pytest.mark.parametrize("shape", [0, 1, 2, 3])
def test_solve_shapes(n3, shape):
    """Every vertical-solve shape (levels per chunk, columns per block) and
    the sequential sweep beyond 512 levels, with the A inversions on the side
    stream and on the member's stream, give the oracle's result."""
    P = make_params(16, n3, W2F=0.05, dt=2e-3, stratification="exponential")
    g = O.Grid(16, 16, n3)
    OP = oracle_params(P)
    psi, BR, BI = physical_ics(P, k0=2)
    S0 = O.initial_state(g, OP, psi, BR, BI)
    S = O.run(g, OP, S0, 3)
    for a_main in (0, 1):
        m = _member(P)
        m.set_variants([0, 0, 0, 0, 0, shape, a_main])
        assert list(m.get_variants()) == [0, 0, 0, 0, 0, shape, a_main]
        m.set_state_physical(psi, BR, BI)
        m.step(3)
        compare_state(m.get_state(), S, 1e-11)
        m.close()


@pytest.mark.parametrize("flags", [qgybj.params.FIXED_FLOW, qgybj.params.PASSIVE_SCALAR,
                                   qgybj.params.NO_FEEDBACK])
def test_a_inversions_on_member_stream_with_flags(flags):
    P = make_params(32, 16, W2F=0.05, dt=2e-3, flags=flags)
    g = O.Grid(32, 32, 16)
    OP = oracle_params(P)
    psi, BR, BI = physical_ics(P, k0=3)
    S = O.run(g, OP, O.initial_state(g, OP, psi, BR, BI), 4)
    m = _member(P)
    m.set_variants([0, 0, 0, 0, 0, 1, 1])
    m.set_state_physical(psi, BR, BI)
    m.step(4)
    compare_state(m.get_state(), S, 1e-11)
    m.close()


def test_set_variants_rejects_out_of_range():
    P = make_params(16, 4)
    m = _member(P)
    before = list(m.get_variants())
    with pytest.raises(qgybj.QGError):
        m.set_variants([0, 0, 0, 0, 0, 4])
    with pytest.raises(qgybj.QGError):
        m.set_variants([0, 0, 0, 0, 0, 0, 2])
    assert list(m.get_variants()) == before
    m.close()


def test_kernel_timing_and_sampling():
    """qg_timing times the selected kernels with events on their streams;
    qg_timing_sample(k) times every k-th launch; timing never changes the
    results."""
    P = make_params(64, 8, W2F=0.05, dt=2e-3)
    psi, BR, BI = physical_ics(P, k0=3)
    ref = _member(P)
    ref.set_state_physical(psi, BR, BI)
    ref.step(11)
    m = _member(P)
    m.set_state_physical(psi, BR, BI)
    m.step(2)
    m.kernel_times()
    m.timing(True)
    m.step(3)
    kt = m.kernel_times()
    assert kt["row_nl"][1] == 3 and kt["col_upd"][1] == 3 and kt["row_nl"][0] > 0
    m.timing_sample(3)
    m.timing("col_upd")
    m.step(6)
    m.timing(False)
    m.timing_sample(1)
    kt = m.kernel_times()
    assert kt["col_upd"][1] == 2 and kt["row_nl"][1] == 0
    a, b = m.get_state(), ref.get_state()
    for f in qgybj.Member.FIELDS:
        assert np.array_equal(a[f], b[f]), f
    with pytest.raises(qgybj.QGError):
        m.timing_sample(0)
    m.close()
    ref.close()
