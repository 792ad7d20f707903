"""Python (ctypes) harness over the C ABI in include/qgybj.h.

The product is the HIP library ``lib/libqgybj.so``; this module only binds
it for tests, benchmarks and the verification harness.  Nothing here computes
a time step: if the library is missing, loading fails loudly.

Arrays: spectral fields are numpy complex128 arrays of shape (n3, n2, n1/2+1)
(byte-identical to the reference's Fortran (iktx, ikty, n3)); physical fields
are float64 arrays (nz, n2, n1).
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from .params import TWOPI, QGParams, make_params  # noqa: F401

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "lib", "libqgybj.so")

QG_OK = 0
ERRORS = {1: "QG_ERR_ARG", 2: "QG_ERR_HIP", 3: "QG_ERR_NOMEM", 4: "QG_ERR_NONFINITE", 5: "QG_ERR_STATE"}
NVARIANTS = 7  # QG_NVARIANTS
KERNEL_NAMES = ["col_inv", "row_nl", "col_upd", "row_qw", "col_qw", "solve", "solve_a", "aux"]

# exported symbols declared in include/qgybj.h
SYMBOLS = [
    "qg_create", "qg_create_batch", "qg_members", "qg_destroy", "qg_last_error", "qg_set_state", "qg_set_state_physical", "qg_get_state",
    "qg_step", "qg_iter", "qg_diag", "qg_hspec_waves", "qg_omega", "qg_wave_vave", "qg_keep_tendencies",
    "qg_get_tendencies", "qg_wave_conversion", "qg_init_q", "qg_fft_r2c", "qg_fft_c2r", "qg_device_field", "qg_device_pitch", "qg_stream",
    "qg_sync", "qg_timing", "qg_timing_sample", "qg_kernel_times", "qg_kernel_name", "qg_set_variants", "qg_get_variants",
    "qg_ensemble_create", "qg_ensemble_destroy", "qg_ensemble_error", "qg_ensemble_set_graph_steps",
    "qg_ensemble_step", "qg_tune", "qg_adj_create", "qg_adj_destroy", "qg_adj_error", "qg_adj_set_fused", "qg_adj_objective",
    "qg_adj_tlm", "qg_adj_apply", "qg_adj_gradient", "qg_mask", "qg_box", "qg_version",
]


class QGError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


class _Params(C.Structure):
    _fields_ = [
        ("n1", C.c_int32), ("n2", C.c_int32), ("n3", C.c_int32), ("flags", C.c_int32),
        ("dt", C.c_double), ("gamma", C.c_double), ("Ro", C.c_double), ("Bu", C.c_double),
        ("W2F", C.c_double), ("nuz", C.c_double),
        ("ifac_q", C.POINTER(C.c_double)), ("ifac_w", C.POINTER(C.c_double)),
        ("N2_u", C.POINTER(C.c_double)), ("N2_s", C.POINTER(C.c_double)),
    ]


_lib = None


def load(path: str = LIB_PATH):
    """Load libqgybj.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise FileNotFoundError(f"{path} not built: run qg-ybjp-model_amd/build.py")
    lib = C.CDLL(path)
    dp = C.POINTER(C.c_double)
    vp = C.c_void_p
    lib.qg_create.argtypes = [C.POINTER(_Params), C.POINTER(vp)]
    lib.qg_create_batch.argtypes = [C.POINTER(_Params), C.c_int32, C.POINTER(vp)]
    lib.qg_members.argtypes = [vp, C.POINTER(C.c_int32)]
    lib.qg_destroy.argtypes = [vp]
    lib.qg_last_error.argtypes = [vp]
    lib.qg_last_error.restype = C.c_char_p
    lib.qg_set_state.argtypes = [vp] + [dp] * 9 + [C.c_int32]
    lib.qg_set_state_physical.argtypes = [vp, dp, dp, dp, C.c_int32]
    lib.qg_get_state.argtypes = [vp] + [dp] * 11
    lib.qg_step.argtypes = [vp, C.c_int32]
    lib.qg_iter.argtypes = [vp, C.POINTER(C.c_int32)]
    lib.qg_diag.argtypes = [vp, dp]
    lib.qg_hspec_waves.argtypes = [vp, C.POINTER(C.c_int32), C.c_int32, dp]
    lib.qg_omega.argtypes = [vp, dp]
    lib.qg_wave_vave.argtypes = [vp, dp]
    lib.qg_keep_tendencies.argtypes = [vp, C.c_int32]
    lib.qg_get_tendencies.argtypes = [vp, dp, dp, dp, dp]
    lib.qg_wave_conversion.argtypes = [vp, dp]
    lib.qg_init_q.argtypes = [vp, dp, dp]
    lib.qg_fft_r2c.argtypes = [vp, dp, C.c_int32, dp, C.c_int32]
    lib.qg_fft_c2r.argtypes = [vp, dp, dp, C.c_int32, C.c_int32]
    lib.qg_device_field.argtypes = [vp, C.c_int32, C.POINTER(vp)]
    lib.qg_device_pitch.argtypes = [vp, C.POINTER(C.c_int32)]
    lib.qg_stream.argtypes = [vp, C.POINTER(vp)]
    lib.qg_sync.argtypes = [vp]
    lib.qg_timing.argtypes = [vp, C.c_int32]
    lib.qg_timing_sample.argtypes = [vp, C.c_int32]
    lib.qg_kernel_times.argtypes = [vp, dp, C.POINTER(C.c_int32), C.c_int32]
    lib.qg_kernel_name.argtypes = [C.c_int32]
    lib.qg_kernel_name.restype = C.c_char_p
    lib.qg_set_variants.argtypes = [vp, C.POINTER(C.c_int32), C.c_int32]
    lib.qg_get_variants.argtypes = [vp, C.POINTER(C.c_int32), C.c_int32]
    lib.qg_ensemble_create.argtypes = [C.POINTER(vp), C.c_int32, C.POINTER(vp)]
    lib.qg_ensemble_destroy.argtypes = [vp]
    lib.qg_ensemble_error.argtypes = [vp]
    lib.qg_ensemble_error.restype = C.c_char_p
    lib.qg_ensemble_set_graph_steps.argtypes = [vp, C.c_int32]
    lib.qg_ensemble_step.argtypes = [vp, C.c_int32]
    lib.qg_tune.argtypes = [vp, C.c_int32, C.POINTER(C.c_int32), dp]
    lib.qg_adj_create.argtypes = [vp, C.c_int32, C.c_double, C.POINTER(vp)]
    lib.qg_adj_destroy.argtypes = [vp]
    lib.qg_adj_error.argtypes = [vp]
    lib.qg_adj_error.restype = C.c_char_p
    lib.qg_adj_set_fused.argtypes = [vp, C.c_int32]
    lib.qg_adj_objective.argtypes = [vp, dp, dp, dp, C.c_int32, C.c_int32, dp]
    lib.qg_adj_tlm.argtypes = [vp, dp, dp, dp, C.c_int32, C.c_int32, dp, dp, dp]
    lib.qg_adj_apply.argtypes = [vp, dp, dp, dp, C.c_int32, C.c_int32, dp, dp]
    lib.qg_adj_gradient.argtypes = [vp, dp, dp, dp, C.c_int32, C.c_int32, dp, dp]
    lib.qg_mask.argtypes = [C.c_int32, C.c_int32, C.POINTER(C.c_int32)]
    lib.qg_box.argtypes = [C.c_int32, C.c_int32, C.POINTER(C.c_int32)]
    lib.qg_version.restype = C.c_char_p
    _lib = lib
    return lib


def _dp(a):
    return None if a is None else a.ctypes.data_as(C.POINTER(C.c_double))


def mask(n1: int) -> np.ndarray:
    """The library's dealiasing mask, shape (ikty, iktx) (no GPU needed)."""
    lib = load()
    L = np.zeros((n1, n1 // 2 + 1), dtype=np.int32)
    rc = lib.qg_mask(n1, n1, L.ctypes.data_as(C.POINTER(C.c_int32)))
    if rc:
        raise QGError(rc, "qg_mask")
    return L


def box(n1: int):
    lib = load()
    out = (C.c_int32 * 4)()
    rc = lib.qg_box(n1, n1, out)
    if rc:
        raise QGError(rc, "qg_box")
    return tuple(out)


class Member:
    """One QG-YBJ+ member, or a batch of members with the same parameters,
    on the current HIP device (a qg_ctx).  With members > 1 every array gets a
    leading member axis."""

    FIELDS = ["psik", "qk", "qok", "BRk", "BIk", "BRok", "BIok", "ARk", "AIk"]

    def __init__(self, P: QGParams, members: int = 1):
        self.lib = load()
        self.P = P
        self.members = int(members)
        self.n1, self.n3 = P.n1, P.n3
        self.iktx = P.n1 // 2 + 1
        lead = () if self.members == 1 else (self.members,)
        self.shape = lead + (P.n3, P.n1, self.iktx)
        self.rshape = lead + (P.n3, P.n1, P.n1)
        self._keep = [np.ascontiguousarray(a, dtype=np.float64) for a in (P.ifac_q, P.ifac_w, P.N2u, P.N2s)]
        prm = _Params(P.n1, P.n1, P.n3, P.flags, P.dt, P.gamma, P.Ro, P.Bu, P.W2F, P.nuz,
                      *[_dp(a) for a in self._keep])
        self._h = C.c_void_p()
        rc = self.lib.qg_create_batch(C.byref(prm), self.members, C.byref(self._h))
        if rc:
            raise QGError(rc, "qg_create_batch failed")

    def _chk(self, rc, what):
        if rc:
            raise QGError(rc, f"{what}: {self.lib.qg_last_error(self._h).decode()}")

    def close(self):
        if self._h:
            self.lib.qg_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _c(self, a):
        if a is None:
            return None
        a = np.ascontiguousarray(a, dtype=np.complex128)
        assert a.shape == self.shape, (a.shape, self.shape)
        return a

    def set_state(self, psik=None, qk=None, qok=None, BRk=None, BIk=None, BRok=None, BIok=None,
                  ARk=None, AIk=None, iter=0):
        arrs = [self._c(a) for a in (psik, qk, qok, BRk, BIk, BRok, BIok, ARk, AIk)]
        self._chk(self.lib.qg_set_state(self._h, *[_dp(a) for a in arrs], int(iter)), "qg_set_state")

    def set_state_physical(self, psi_r, BR_r, BI_r):
        arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (psi_r, BR_r, BI_r)]
        for a in arrs:
            assert a.shape == self.rshape, (a.shape, self.rshape)
        self._chk(self.lib.qg_set_state_physical(self._h, *[_dp(a) for a in arrs], self.n1),
                  "qg_set_state_physical")

    def get_state(self, with_c=False) -> dict:
        out = {k: np.empty(self.shape, np.complex128) for k in self.FIELDS}
        cr = np.empty(self.shape, np.complex128) if with_c else None
        ci = np.empty(self.shape, np.complex128) if with_c else None
        self._chk(self.lib.qg_get_state(self._h, *[_dp(out[k]) for k in self.FIELDS], _dp(cr), _dp(ci)),
                  "qg_get_state")
        if with_c:
            out["CRk"] = cr
            out["CIk"] = ci
        return out

    def step(self, n=1):
        self._chk(self.lib.qg_step(self._h, int(n)), "qg_step")

    @property
    def iter(self) -> int:
        v = C.c_int32()
        self._chk(self.lib.qg_iter(self._h, C.byref(v)), "qg_iter")
        return v.value

    def diag(self) -> np.ndarray:
        """[KE, PE, WKE, WPE, WCE, ENS] (per member: shape (members, 6))."""
        out = np.zeros((self.members, 6))
        self._chk(self.lib.qg_diag(self._h, _dp(out)), "qg_diag")
        return out[0] if self.members == 1 else out

    def hspec_waves(self, levels) -> np.ndarray:
        """Wave spectra at the given 1-based levels (diagnostics.f90:547-638):
        shape (len(levels), n1/2, 3) = [KE, PE, number of modes] per shell
        (a leading members axis for a batch)."""
        lv = np.ascontiguousarray(levels, dtype=np.int32)
        out = np.zeros((self.members, len(lv), self.n1 // 2, 3))
        self._chk(self.lib.qg_hspec_waves(self._h, lv.ctypes.data_as(C.POINTER(C.c_int32)), len(lv),
                                          _dp(out)), "qg_hspec_waves")
        return out[0] if self.members == 1 else out

    def omega(self) -> np.ndarray:
        """QG vertical velocity w (omega equation, elliptic.f90:463-567) of the
        current state, shaped like a state field (w at z_u(k) = k dz, 0 at the
        top level)."""
        w = np.empty(self.shape, np.complex128)
        self._chk(self.lib.qg_omega(self._h, _dp(w)), "qg_omega")
        return w

    def wave_vave(self) -> np.ndarray:
        """Vertically averaged wave energy density 0.5 (BR^2 + BI^2) / n3 per
        physical point (we_vave, diagnostics.f90:750-813), shape (n2, n1)
        (a leading members axis for a batch), without Uw_scale^2."""
        lead = () if self.members == 1 else (self.members,)
        out = np.zeros(lead + (self.n1, self.n1))
        self._chk(self.lib.qg_wave_vave(self._h, _dp(out)), "qg_wave_vave")
        return out

    def keep_tendencies(self, on: bool = True) -> None:
        """Keep each step's wave tendencies nB, rB (for tendencies() and
        wave_conversion())."""
        self._chk(self.lib.qg_keep_tendencies(self._h, 1 if on else 0), "qg_keep_tendencies")

    def tendencies(self):
        """(nBRk, nBIk, rBRk, rBIk) of the last step (convol_waqg /
        refraction_waqg), shaped like state fields."""
        out = [np.empty(self.shape, np.complex128) for _ in range(4)]
        self._chk(self.lib.qg_get_tendencies(self._h, *[_dp(a) for a in out]), "qg_get_tendencies")
        return tuple(out)

    def wave_conversion(self) -> np.ndarray:
        """[we ca, we cr, gamma ca, gamma cr, gamma cd] (diagnostics.f90:967-1256)
        without the dimensional factors (per member: shape (members, 5))."""
        out = np.zeros((self.members, 5))
        self._chk(self.lib.qg_wave_conversion(self._h, _dp(out)), "qg_wave_conversion")
        return out[0] if self.members == 1 else out

    def init_q(self, psik):
        psik = self._c(psik)
        q = np.empty(self.shape, np.complex128)
        self._chk(self.lib.qg_init_q(self._h, _dp(psik), _dp(q)), "qg_init_q")
        return q

    def r2c(self, r):
        r = np.ascontiguousarray(r, dtype=np.float64)
        nz = r.shape[0]
        assert r.ndim == 3
        k = np.empty((nz, self.n1, self.iktx), np.complex128)
        self._chk(self.lib.qg_fft_r2c(self._h, _dp(r), r.shape[-1], _dp(k), nz), "qg_fft_r2c")
        return k

    def c2r(self, k, ldx=None):
        k = np.ascontiguousarray(k, dtype=np.complex128)
        nz = k.shape[0]
        ldx = ldx or self.n1
        r = np.empty((nz, self.n1, ldx), np.float64)
        self._chk(self.lib.qg_fft_c2r(self._h, _dp(k), _dp(r), ldx, nz), "qg_fft_c2r")
        return r[..., : self.n1]

    def set_variants(self, v):
        """v[0..4]: transform-kernel variants, v[5]: solve shape, v[6]: A
        inversions on the member's stream (1) or a side stream (0)."""
        a = np.ascontiguousarray(v, dtype=np.int32)
        self._chk(self.lib.qg_set_variants(self._h, a.ctypes.data_as(C.POINTER(C.c_int32)), len(a)),
                  "qg_set_variants")

    def get_variants(self):
        a = np.zeros(NVARIANTS, dtype=np.int32)
        self._chk(self.lib.qg_get_variants(self._h, a.ctypes.data_as(C.POINTER(C.c_int32)), NVARIANTS),
                  "qg_get_variants")
        return a

    def sync(self):
        self._chk(self.lib.qg_sync(self._h), "qg_sync")

    def tune(self, steps: int = 5):
        """Run the tuner (qg_tune) from the current state; returns the chosen
        variants and the best time per step in ms.  The state is unchanged."""
        best = np.zeros(NVARIANTS, dtype=np.int32)
        ms = C.c_double()
        self._chk(self.lib.qg_tune(self._h, int(steps), best.ctypes.data_as(C.POINTER(C.c_int32)), C.byref(ms)),
                  "qg_tune")
        return best, ms.value

    def timing(self, on=True):
        """on: True (every kernel), False, a kernel name or an id bit mask."""
        if on is True:
            mask = 0xFF
        elif on is False or on is None:
            mask = 0
        elif isinstance(on, str):
            mask = 1 << KERNEL_NAMES.index(on)
        else:
            mask = int(on)
        self._chk(self.lib.qg_timing(self._h, mask), "qg_timing")

    def timing_sample(self, every: int = 1):
        """Time only every `every`-th launch of each timed kernel."""
        self._chk(self.lib.qg_timing_sample(self._h, int(every)), "qg_timing_sample")

    def kernel_times(self):
        n = len(KERNEL_NAMES)
        ms = np.zeros(n)
        cnt = np.zeros(n, dtype=np.int32)
        self._chk(self.lib.qg_kernel_times(self._h, _dp(ms), cnt.ctypes.data_as(C.POINTER(C.c_int32)), n),
                  "qg_kernel_times")
        return {KERNEL_NAMES[i]: (float(ms[i]), int(cnt[i])) for i in range(n)}


class Ensemble:
    """Contexts (members or batches) advanced together from captured HIP
    graphs (qg_ensemble_*)."""

    def __init__(self, members):
        self.lib = load()
        self.members = list(members)
        arr = (C.c_void_p * len(self.members))(*[m._h.value for m in self.members])
        self._h = C.c_void_p()
        rc = self.lib.qg_ensemble_create(arr, len(self.members), C.byref(self._h))
        if rc:
            raise QGError(rc, "qg_ensemble_create failed")

    def _chk(self, rc, what):
        if rc:
            raise QGError(rc, f"{what}: {self.lib.qg_ensemble_error(self._h).decode()}")

    def set_graph_steps(self, k: int):
        self._chk(self.lib.qg_ensemble_set_graph_steps(self._h, int(k)), "qg_ensemble_set_graph_steps")

    def step(self, n: int = 1):
        self._chk(self.lib.qg_ensemble_step(self._h, int(n)), "qg_ensemble_step")

    def close(self):
        if self._h:
            self.lib.qg_ensemble_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Adjoint:
    """Tangent-linear and adjoint of the step bound to a single-member context
    (qg_adj_*).  Objective: domain integral of |grad_h A|^2 / 2 over the
    levels with z_s < z_deep after nsteps; control: the physical initial
    streamfunction psi (n3, n1, n1)."""

    def __init__(self, member: "Member", checkpoint_every: int = 20, z_deep: float | None = None):
        self.lib = load()
        self.m = member
        self.z_deep = 0.8 * TWOPI if z_deep is None else float(z_deep)
        self._h = C.c_void_p()
        rc = self.lib.qg_adj_create(member._h, int(checkpoint_every), self.z_deep, C.byref(self._h))
        if rc:
            raise QGError(rc, f"qg_adj_create: {member.lib.qg_last_error(member._h).decode()}")

    def _chk(self, rc, what):
        if rc:
            raise QGError(rc, f"{what}: {self.lib.qg_adj_error(self._h).decode()}")

    def _phys(self, *arrs):
        out = [np.ascontiguousarray(a, dtype=np.float64) for a in arrs]
        for a in out:
            assert a.shape == self.m.rshape, (a.shape, self.m.rshape)
        return out

    def set_fused(self, fused: bool):
        self._chk(self.lib.qg_adj_set_fused(self._h, 1 if fused else 0), "qg_adj_set_fused")

    def _out(self, arr):
        return float(arr[0]) if self.m.members == 1 else arr

    def objective(self, psi, BR, BI, nsteps: int):
        """J (a float, or one per member for a batch)."""
        psi, BR, BI = self._phys(psi, BR, BI)
        J = np.zeros(self.m.members)
        self._chk(self.lib.qg_adj_objective(self._h, _dp(psi), _dp(BR), _dp(BI), self.m.n1, int(nsteps), _dp(J)),
                  "qg_adj_objective")
        return self._out(J)

    def tlm(self, psi, BR, BI, nsteps: int, dpsi, want_state: bool = False):
        """Returns dJ along dpsi and (optionally) the final-state perturbation
        as a (9, n3, n1, iktx) complex array (psi, q, qo, BR, BI, BRo, BIo, AR, AI)."""
        psi, BR, BI, dpsi = self._phys(psi, BR, BI, dpsi)
        dJ = np.zeros(self.m.members)
        st = np.empty((9,) + self.m.shape, np.complex128) if want_state else None
        self._chk(self.lib.qg_adj_tlm(self._h, _dp(psi), _dp(BR), _dp(BI), self.m.n1, int(nsteps), _dp(dpsi),
                                      _dp(dJ), _dp(st)), "qg_adj_tlm")
        return self._out(dJ), st

    def apply(self, psi, BR, BI, nsteps: int, state_bar):
        """Adjoint of the map psi -> final state, applied to state_bar (9, n3, n1, iktx)."""
        psi, BR, BI = self._phys(psi, BR, BI)
        sb = np.ascontiguousarray(state_bar, dtype=np.complex128)
        assert sb.shape == (9,) + self.m.shape
        out = np.empty(self.m.rshape, np.float64)
        self._chk(self.lib.qg_adj_apply(self._h, _dp(psi), _dp(BR), _dp(BI), self.m.n1, int(nsteps), _dp(sb),
                                        _dp(out)), "qg_adj_apply")
        return out

    def gradient(self, psi, BR, BI, nsteps: int):
        psi, BR, BI = self._phys(psi, BR, BI)
        J = np.zeros(self.m.members)
        g = np.empty(self.m.rshape, np.float64)
        self._chk(self.lib.qg_adj_gradient(self._h, _dp(psi), _dp(BR), _dp(BI), self.m.n1, int(nsteps),
                                           _dp(J), _dp(g)), "qg_adj_gradient")
        return self._out(J), g

    def close(self):
        if self._h:
            self.lib.qg_adj_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
