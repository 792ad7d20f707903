// C ABI of the QG-YBJ+ step (include/qgybj.h): member contexts, device
// buffers, host-computed tables and the per-step kernel sequence.
#include "qgybj.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "qg_common.hpp"
#include "qg_ctx.hpp"
#include "qg_fft.hpp"

using qg::Dev;
using qg::Geo;

namespace qg {

Geo make_geo(int n1, int n2, int n3, int nm) {
  Geo g{};
  g.n1 = n1;
  g.n2 = n2;
  g.n3 = n3;
  g.nm = nm;
  g.nz = nm * n3;
  g.iktx = n1 / 2 + 1;
  g.ikty = n2;
  g.sp = (g.iktx + 7) & ~7;
  // init.f90:50,72: L = 0 where (kh*L1/twopi) > ifix(float(n1)/3.+0.5)-0.5,
  // kh = REAL(4) sqrt(kx^2+ky^2).  The test is monotone in kh2, so it is an
  // integer bound kh2 <= kh2max, found here with the same arithmetic.
  const double twopi = 4.0 * std::asin(1.0);
  const int ci = (int)((float)n1 / 3.0f + 0.5f);
  const float cut = (float)ci - 0.5f;
  int m = 0;
  while (true) {
    const double kh = (double)std::sqrt((float)(m + 1));
    if ((kh * twopi) / twopi > (double)cut) break;
    ++m;
  }
  g.kh2max = m;
  int k = 0;
  while ((k + 1) * (k + 1) <= g.kh2max) ++k;
  g.kmax = k;
  g.KX = k + 1;
  g.nlo = k + 1;
  g.nhi = k;
  g.nbox = g.nlo + g.nhi;
  return g;
}

}  // namespace qg


namespace qgi {

int fail(qg_ctx* c, int code, const std::string& msg) {
  if (c) c->err = msg;
  return code;
}
int fail_hip(qg_ctx* c, hipError_t e, const char* what) {
  return fail(c, QG_ERR_HIP, std::string("HIP error '") + hipGetErrorString(e) + "' in " + what);
}


template <class T>
int dalloc(qg_ctx* c, T** p, size_t n) {
  void* v = nullptr;
  hipError_t e = hipMalloc(&v, n * sizeof(T) > 0 ? n * sizeof(T) : 16);
  if (e != hipSuccess) return fail(c, QG_ERR_NOMEM, std::string("hipMalloc failed: ") + hipGetErrorString(e));
  c->allocs.push_back(v);
  *p = static_cast<T*>(v);
  return QG_OK;
}

size_t field_elems(const Geo& g) { return (size_t)g.nz * g.ikty * g.sp; }

// host (iktx, ikty, nz) complex <-> device rows of pitch sp, through a
// contiguous device staging buffer (one host transfer, the re-pitching as a
// device-side 2D copy)
hipError_t h2d_field(const Geo& g, double2* dst, const double* src, int nz, hipStream_t st, double2* stage) {
  const size_t n = (size_t)g.iktx * g.ikty * nz;
  hipError_t e = hipMemcpyAsync(stage, src, n * sizeof(double2), hipMemcpyHostToDevice, st);
  if (e != hipSuccess) return e;
  return hipMemcpy2DAsync(dst, (size_t)g.sp * sizeof(double2), stage, (size_t)g.iktx * sizeof(double2),
                          (size_t)g.iktx * sizeof(double2), (size_t)g.ikty * nz, hipMemcpyDeviceToDevice, st);
}
hipError_t d2h_field(const Geo& g, double* dst, const double2* src, int nz, hipStream_t st, double2* stage) {
  const size_t n = (size_t)g.iktx * g.ikty * nz;
  hipError_t e = hipMemcpy2DAsync(stage, (size_t)g.iktx * sizeof(double2), src, (size_t)g.sp * sizeof(double2),
                                  (size_t)g.iktx * sizeof(double2), (size_t)g.ikty * nz, hipMemcpyDeviceToDevice, st);
  if (e == hipSuccess) e = hipMemcpyAsync(dst, stage, n * sizeof(double2), hipMemcpyDeviceToHost, st);
  if (e != hipSuccess) return e;
  return hipStreamSynchronize(st);
}

// exp(-2 pi i m / TWN), from long double, with exact quadrant symmetry
void make_twiddles(std::vector<double2>& tw) {
  const int n = qg::TWN;
  const long double pi = 3.141592653589793238462643383279502884L;
  tw.resize(n);
  for (int m = 0; m < n; ++m) {
    const int q = m / (n / 4), r = m % (n / 4);
    const long double th = 2.0L * pi * (long double)r / (long double)n;
    long double c = std::cos(th), s = std::sin(th);
    if (r == 0) { c = 1.0L; s = 0.0L; }
    long double co, si;
    switch (q) {
      case 0: co = c; si = s; break;
      case 1: co = -s; si = c; break;
      case 2: co = -c; si = -s; break;
      default: co = s; si = -c; break;
    }
    tw[m] = make_double2((double)co, (double)(-si));
  }
}

struct Timed {
  qg_ctx* c;
  int id;
  hipStream_t st;
  hipEvent_t a = nullptr;
  Timed(qg_ctx* cc, int i, hipStream_t s) : c(cc), id(i), st(s) {
    if (c->capturing || !((c->timing >> id) & 1u)) return;
    if (c->tseen[id]++ % c->timing_every != 0) return;
    if (hipEventCreate(&a) == hipSuccess) hipEventRecord(a, st);
  }
  void done() {
    if (a) {
      hipEvent_t b = nullptr;
      if (hipEventCreate(&b) == hipSuccess) {
        hipEventRecord(b, st);
        c->events.push_back({id, a, b});
      } else {
        hipEventDestroy(a);
      }
      a = nullptr;
    }
  }
};

#define KRUN_ON(c, s, id, call)                                 \
  do {                                                          \
    Timed _t((c), (id), (s));                                   \
    hipError_t _e = (call);                                     \
    _t.done();                                                  \
    if (_e != hipSuccess) return fail_hip((c), _e, #call);      \
  } while (0)
#define KRUN(c, id, call) KRUN_ON(c, (c)->stream, id, call)

int step_once(qg_ctx* c) {
  Dev& d = c->d;
  hipStream_t st = c->stream;
  const Geo& g = d.g;
  const int flags = d.flags;
  const bool euler = (c->iter == 0);
  const int has_dq = (d.nuz != 0.0 && !(flags & qg::F_INVISCID)) ? 1 : 0;
  const size_t bytes = field_elems(g) * sizeof(double2);
  // d^q from the old q as it stands before this step's assignments
  // (main_waqg.f90:206 then :232 in the Euler step, :335 in leapfrog)
  if (has_dq) KRUN(c, qg::K_AUX, qg::launch_dq(d, st));
  if (euler) {
    // qok = qk, BRok = BRk, BIok = BIk (main_waqg.f90:232-234)
    HIPCHK(c, hipMemcpyAsync(d.qo, d.q, bytes, hipMemcpyDeviceToDevice, st));
    HIPCHK(c, hipMemcpyAsync(d.BRo, d.BR, bytes, hipMemcpyDeviceToDevice, st));
    HIPCHK(c, hipMemcpyAsync(d.BIo, d.BI, bytes, hipMemcpyDeviceToDevice, st));
  }
  const bool full = euler || !c->p_valid;
  // regular YBJ (ybj_plus = 0) with dispersion: A from sumB / compute_sigma /
  // compute_A instead of the YBJ+ inversion
  const bool regular = (flags & qg::F_YBJ_REGULAR) && !(flags & qg::F_PASSIVE);
  const int W = full ? g.iktx : g.KX;
  const int box = full ? 0 : 1;
  // u, v, zeta ingredients from psi (compute_velo derivatives.f90:1302-1311,
  // refraction_waqg :831-845)
  {
    qg::ColInvJobs jobs{};
    jobs.j[0] = qg::ColInvJob{d.psi, {d.I[1], d.I[3], d.I[5]}, W, box, 1, g.nz};
    jobs.n = 1;
    if (full) {
      // unmasked fields as the state holds them (the initial B is un-dealiased)
      jobs.j[1] = qg::ColInvJob{d.q, {d.I[0], nullptr, nullptr}, W, 0, 0, g.nz};
      jobs.j[2] = qg::ColInvJob{d.BR, {d.I[2], nullptr, nullptr}, W, 0, 0, g.nz};
      jobs.j[3] = qg::ColInvJob{d.BI, {d.I[4], nullptr, nullptr}, W, 0, 0, g.nz};
      jobs.n = 4;
    } else if (regular) {
      // sumB changed B after the update formed these planes: form them again
      // from the mean-free B (the same masked box transform the update uses)
      jobs.j[1] = qg::ColInvJob{d.BR, {d.I[2], nullptr, nullptr}, W, 1, 1, g.nz};
      jobs.j[2] = qg::ColInvJob{d.BI, {d.I[4], nullptr, nullptr}, W, 1, 1, g.nz};
      jobs.n = 3;
    }
    KRUN(c, qg::K_COL_INV, qg::launch_col_inv(d, st, jobs));
  }
  KRUN(c, qg::K_ROW_NL, qg::launch_row_nl(d, st, W));
  // the update reads A^n and overwrites B in place: wait for the previous
  // step's A inversions on the side stream
  if (c->evA_live) HIPCHK(c, hipStreamWaitEvent(st, c->evA, 0));
  KRUN(c, qg::K_COL_UPD, qg::launch_col_upd(d, st, euler ? 1 : 0, has_dq));
  c->p_valid = 1;
  const bool solve_a = !(flags & qg::F_PASSIVE) && !regular;
  const bool a_side = solve_a && !d.tune.a_main;
  if (a_side) {
    // A^{n+1} from B^{n+1} (A_solver_ybj_plus) on the side stream; it is
    // needed by the next update only
    HIPCHK(c, hipEventRecord(c->evB, st));
    HIPCHK(c, hipStreamWaitEvent(c->stream2, c->evB, 0));
    KRUN_ON(c, c->stream2, qg::K_SOLVE_A, qg::launch_solve(d, c->stream2, 1, 2, nullptr));
    HIPCHK(c, hipEventRecord(c->evA, c->stream2));
    c->evA_live = 1;
  }
  // the inversions left for the member's stream: psi (field 0) unless the
  // flow is fixed, then AR, AI (fields 1, 2) unless they ran on the side
  const int f_lo = (flags & qg::F_FIXED_FLOW) ? 1 : 0;
  const int f_hi = (solve_a && !a_side) ? 2 : 0;
  if (!(flags & qg::F_FIXED_FLOW)) {
    const double2* rhs = d.q;
    if (!(flags & qg::F_NO_FEEDBACK)) {
      KRUN(c, qg::K_ROW_QW, qg::launch_row_qw(d, st));
      KRUN(c, qg::K_COL_QW, qg::launch_col_qw(d, st));
      rhs = d.qs;
    }
    KRUN(c, qg::K_SOLVE, qg::launch_solve(d, st, f_lo, f_hi, rhs));
  } else if (f_hi > 0) {
    KRUN(c, qg::K_SOLVE_A, qg::launch_solve(d, st, f_lo, f_hi, nullptr));
  }
  // after compute_qw has used B^{n+1} (main_waqg.f90:449-461)
  if (regular) KRUN(c, qg::K_SOLVE_A, qg::launch_ybj_regular(d, st));
  // kept tendencies: after the regular operator has summed the combined ones
  if (d.Z[0]) {
    KRUN(c, qg::K_AUX, qg::launch_tend_split(d, st));
    c->tend_valid = 1;
  }
  if (c->clear_pending > 0) {
    if (c->evA_live) HIPCHK(c, hipStreamWaitEvent(st, c->evA, 0));
    // modes outside the dealiasing box are zero in the reference after the
    // masked update / filter / inversions; the hot loop never touches them
    KRUN(c, qg::K_AUX, qg::launch_clear_outside(d, st, d.q, 0));
    KRUN(c, qg::K_AUX, qg::launch_clear_outside(d, st, d.BR, 0));
    KRUN(c, qg::K_AUX, qg::launch_clear_outside(d, st, d.BI, 0));
    if (!(flags & qg::F_FIXED_FLOW)) KRUN(c, qg::K_AUX, qg::launch_clear_outside(d, st, d.psi, 0));
    const int all = (flags & qg::F_PASSIVE) ? 1 : 0;
    KRUN(c, qg::K_AUX, qg::launch_clear_outside(d, st, d.AR, all));
    KRUN(c, qg::K_AUX, qg::launch_clear_outside(d, st, d.AI, all));
    if (!euler) {
      KRUN(c, qg::K_AUX, qg::launch_clear_outside(d, st, d.qo, 0));
      KRUN(c, qg::K_AUX, qg::launch_clear_outside(d, st, d.BRo, 0));
      KRUN(c, qg::K_AUX, qg::launch_clear_outside(d, st, d.BIo, 0));
    }
    c->clear_pending--;
  }
  c->iter++;
  return QG_OK;
}

int check_ctx(qg_ctx* c) {
  if (!c) return QG_ERR_ARG;
  hipError_t e = hipSetDevice(c->device);
  if (e != hipSuccess) return fail_hip(c, e, "hipSetDevice");
  return QG_OK;
}

}  // namespace qgi

using namespace qgi;

extern "C" {

const char* qg_version(void) { return "qgybj-mi355x 0.1"; }

const char* qg_kernel_name(int32_t id) { return qg::kernel_name(id); }

int qg_box(int32_t n1, int32_t n2, int32_t out[4]) {
  if (!out || n1 != n2 || n1 < 4 || (n1 & 1)) return QG_ERR_ARG;
  Geo g = qg::make_geo(n1, n2, 2);
  out[0] = g.kmax;
  out[1] = g.KX;
  out[2] = g.nlo;
  out[3] = g.nhi;
  return QG_OK;
}

int qg_mask(int32_t n1, int32_t n2, int32_t* L) {
  if (!L || n1 != n2 || n1 < 4 || (n1 & 1)) return QG_ERR_ARG;
  Geo g = qg::make_geo(n1, n2, 2);
  for (int iky = 0; iky < g.ikty; ++iky)
    for (int kx = 0; kx < g.iktx; ++kx) L[(size_t)iky * g.iktx + kx] = qg::lmask(kx, iky, g) ? 1 : 0;
  return QG_OK;
}

const char* qg_last_error(const qg_ctx* c) { return c ? c->err.c_str() : "null context"; }

int qg_destroy(qg_ctx* c) {
  if (!c) return QG_OK;
  hipSetDevice(c->device);
  if (c->stream) hipStreamSynchronize(c->stream);
  if (c->stream2) hipStreamSynchronize(c->stream2);
  for (auto& e : c->events) {
    hipEventDestroy(e.a);
    hipEventDestroy(e.b);
  }
  for (void* p : c->allocs) hipFree(p);
  if (c->evA) hipEventDestroy(c->evA);
  if (c->evB) hipEventDestroy(c->evB);
  if (c->stream2) hipStreamDestroy(c->stream2);
  if (c->stream) hipStreamDestroy(c->stream);
  delete c;
  return QG_OK;
}

int qg_create(const qg_params* p, qg_ctx** out) { return qg_create_batch(p, 1, out); }

int qg_members(const qg_ctx* c, int32_t* members) {
  if (!c || !members) return QG_ERR_ARG;
  *members = c->d.g.nm;
  return QG_OK;
}

int qg_create_batch(const qg_params* p, int32_t members, qg_ctx** out) {
  if (!p || !out) return QG_ERR_ARG;
  *out = nullptr;
  if (p->n1 != p->n2 || !qg::supported_n(p->n1) || p->n3 < 2) return QG_ERR_ARG;
  if (members < 1 || (int64_t)members * p->n3 > (1 << 20)) return QG_ERR_ARG;
  if (!p->ifac_q || !p->ifac_w || !p->N2_u || !p->N2_s) return QG_ERR_ARG;
  if (!(p->dt > 0.0) || !std::isfinite(p->gamma) || !(p->Ro > 0.0) || !(p->Bu > 0.0)) return QG_ERR_ARG;
  if (p->flags & ~qg::F_ALL) return QG_ERR_ARG;
  qg_ctx* c = new qg_ctx();
  int rc;
  hipError_t e = hipGetDevice(&c->device);
  if (e != hipSuccess) {
    qg_destroy(c);
    return QG_ERR_HIP;
  }
  e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->stream2, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&c->evA, hipEventDisableTiming);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&c->evB, hipEventDisableTiming);
  if (e == hipSuccess) e = hipEventRecord(c->evA, c->stream2);
  if (e != hipSuccess) {
    qg_destroy(c);
    return QG_ERR_HIP;
  }
  Dev& d = c->d;
  d.g = qg::make_geo(p->n1, p->n2, p->n3, members);
  const Geo& g = d.g;
  // measured defaults at 256^2 x 128: 8-level chunks, 32 columns per block
  // (0.471 vs 0.474 ms per step with 16-level chunks), A inversions on the
  // side stream
  d.tune.solve = 1;
  d.tune.a_main = 0;
  d.tune.e16[2] = 2;  // col_upd with one exchange buffer (171 vs 182 us)
  d.tune.e16[4] = 2;  // col_qw with one exchange buffer (43 vs 48 us)
  // at 512^2 x 256, col_qw with 16 elements per thread: 327 vs 511 us
  if (p->n2 >= 512) d.tune.e16[4] = 3;
  d.flags = p->flags;
  d.dt = p->dt;
  d.gamma = p->gamma;
  d.Ro = p->Ro;
  d.Bu = p->Bu;
  d.W2F = p->W2F;
  d.nuz = p->nuz;
  d.dz = (4.0 * std::asin(1.0)) / p->n3;  // dz = L3/n3 (parameters_test1.f90:14)
  const size_t S = field_elems(g);
  const size_t plane = (size_t)g.ikty * g.iktx;
  double2** fields[] = {&d.psi, &d.q, &d.qo, &d.BR, &d.BI, &d.BRo, &d.BIo, &d.AR, &d.AI, &d.qs, &d.dq, &c->stage};
  for (auto f : fields)
    if ((rc = dalloc(c, f, S)) != QG_OK) { qg_destroy(c); return rc; }
  for (int i = 0; i < 6; ++i)
    if ((rc = dalloc(c, &d.I[i], (size_t)g.nz * g.n2 * g.sp)) != QG_OK) { qg_destroy(c); return rc; }
  if (d.flags & qg::F_YBJ_REGULAR) {
    if ((rc = dalloc(c, &d.tR, S)) != QG_OK) { qg_destroy(c); return rc; }
    if ((rc = dalloc(c, &d.tI, S)) != QG_OK) { qg_destroy(c); return rc; }
  }
  const size_t ncols = (size_t)g.KX * g.nbox;
  if ((rc = dalloc(c, &d.scratch, 3 * (size_t)g.nz * ncols * 2)) != QG_OK) { qg_destroy(c); return rc; }
  double2* tw = nullptr;
  if ((rc = dalloc(c, &tw, (size_t)qg::TWN)) != QG_OK) { qg_destroy(c); return rc; }
  d.tw = tw;
  double* tab = nullptr;
  if ((rc = dalloc(c, &tab, 5 * plane + 6 * (size_t)g.n3)) != QG_OK) { qg_destroy(c); return rc; }
  if ((rc = dalloc(c, &d.diag_part, (size_t)qg::diag_blocks(g) * g.nm * 6)) != QG_OK) { qg_destroy(c); return rc; }
  if ((rc = dalloc(c, &d.flag, 1)) != QG_OK) { qg_destroy(c); return rc; }

  // host tables
  std::vector<double2> htw;
  make_twiddles(htw);
  std::vector<double> h(5 * plane + 6 * (size_t)g.n3);
  for (size_t i = 0; i < plane; ++i) {
    h[i] = std::exp(-p->ifac_q[i]);
    h[plane + i] = std::exp(-2.0 * p->ifac_q[i]);
    h[2 * plane + i] = std::exp(-p->ifac_w[i]);
    h[3 * plane + i] = std::exp(-2.0 * p->ifac_w[i]);
    // the wave dissipation operator without delt (diagnostics.f90:1218-1219)
    h[4 * plane + i] = p->dt != 0.0 ? p->ifac_w[i] / p->dt : 0.0;
  }
  const int n3 = g.n3;
  double* hv = h.data() + 5 * plane;
  double* hdl = hv;
  double* hdu = hv + n3;
  double* hds = hv + 2 * n3;
  double* ha = hv + 3 * n3;
  double* hN2u = hv + 4 * n3;
  double* hN2s = hv + 5 * n3;
  for (int k = 0; k < n3; ++k) {
    ha[k] = p->Bu / p->N2_u[k];  // a_ell_ut (init.f90:245)
    hN2u[k] = p->N2_u[k];
    hN2s[k] = p->N2_s[k];
  }
  // elliptic.f90:177-189 with rho_ut = rho_st = 1
  for (int k = 0; k < n3; ++k) {
    hdl[k] = k < n3 - 1 ? ha[k] : 0.0;
    hdu[k] = k < n3 - 1 ? ha[k] : 0.0;
    if (k == 0) hds[k] = ha[0];
    else if (k == n3 - 1) hds[k] = ha[n3 - 2];
    else hds[k] = ha[k] + ha[k - 1];
  }
  e = hipMemcpy(tw, htw.data(), htw.size() * sizeof(double2), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(tab, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    qg_destroy(c);
    return QG_ERR_HIP;
  }
  d.e1q = tab;
  d.e2q = tab + plane;
  d.e1w = tab + 2 * plane;
  d.e2w = tab + 3 * plane;
  d.cw = tab + 4 * plane;
  const double* tv = tab + 5 * plane;
  d.dl = tv;
  d.du = tv + n3;
  d.dsum = tv + 2 * n3;
  d.a_ell = tv + 3 * n3;
  d.N2u = tv + 4 * n3;
  d.N2s = tv + 5 * n3;
  // zero state
  double2* zf[] = {d.psi, d.q, d.qo, d.BR, d.BI, d.BRo, d.BIo, d.AR, d.AI, d.qs, d.dq};
  for (double2* f : zf) {
    e = qg::launch_clear_outside(d, c->stream, f, 1);
    if (e != hipSuccess) { qg_destroy(c); return QG_ERR_HIP; }
  }
  e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) { qg_destroy(c); return QG_ERR_HIP; }
  *out = c;
  return QG_OK;
}

int qg_set_state(qg_ctx* c, const double* psik, const double* qk, const double* qok, const double* BRk,
                 const double* BIk, const double* BRok, const double* BIok, const double* ARk, const double* AIk,
                 int32_t iter) {
  int rc = check_ctx(c);
  if (rc) return rc;
  if (iter < 0) return fail(c, QG_ERR_ARG, "iter < 0");
  Dev& d = c->d;
  const double* src[] = {psik, qk, qok, BRk, BIk, BRok, BIok, ARk, AIk};
  double2* dst[] = {d.psi, d.q, d.qo, d.BR, d.BI, d.BRo, d.BIo, d.AR, d.AI};
  for (int i = 0; i < 9; ++i) {
    HIPCHK(c, qg::launch_clear_outside(d, c->stream, dst[i], 1));
    if (src[i]) HIPCHK(c, h2d_field(d.g, dst[i], src[i], d.g.nz, c->stream, c->stage));
  }
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->iter = iter;
  c->p_valid = 0;
  c->clear_pending = 2;
  c->tend_valid = 0;
  return QG_OK;
}

int qg_set_state_physical(qg_ctx* c, const double* psir, const double* BRr, const double* BIr, int32_t ldx) {
  int rc = check_ctx(c);
  if (rc) return rc;
  Dev& d = c->d;
  const Geo& g = d.g;
  if (!psir || !BRr || !BIr || ldx < g.n1) return fail(c, QG_ERR_ARG, "null input or ldx < n1");
  const size_t nr = (size_t)ldx * g.n2 * g.nz;
  double* tmp = nullptr;
  HIPCHK(c, hipMalloc(&tmp, nr * sizeof(double)));
  const double* srcs[] = {psir, BRr, BIr};
  double2* dsts[] = {d.psi, d.BR, d.BI};
  for (int i = 0; i < 3; ++i) {
    hipError_t e = hipMemcpyAsync(tmp, srcs[i], nr * sizeof(double), hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = qg::launch_r2c(d, c->stream, tmp, ldx, dsts[i], g.nz);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) {
      hipFree(tmp);
      return fail_hip(c, e, "r2c of initial conditions");
    }
  }
  hipFree(tmp);
  if (d.flags & qg::F_FIXED_FLOW) HIPCHK(c, qg::launch_clear_outside(d, c->stream, d.q, 1));
  else HIPCHK(c, qg::launch_init_q(d, c->stream, d.psi, d.q));
  double2* z[] = {d.qo, d.BRo, d.BIo, d.AR, d.AI};
  for (double2* f : z) HIPCHK(c, qg::launch_clear_outside(d, c->stream, f, 1));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->iter = 0;
  c->p_valid = 0;
  c->clear_pending = 2;
  c->tend_valid = 0;
  return QG_OK;
}

int qg_get_state(qg_ctx* c, double* psik, double* qk, double* qok, double* BRk, double* BIk, double* BRok,
                 double* BIok, double* ARk, double* AIk, double* CRk, double* CIk) {
  int rc = check_ctx(c);
  if (rc) return rc;
  Dev& d = c->d;
  double* dst[] = {psik, qk, qok, BRk, BIk, BRok, BIok, ARk, AIk};
  const double2* src[] = {d.psi, d.q, d.qo, d.BR, d.BI, d.BRo, d.BIo, d.AR, d.AI};
  for (int i = 0; i < 9; ++i)
    if (dst[i]) HIPCHK(c, d2h_field(d.g, dst[i], src[i], d.g.nz, c->stream, c->stage));
  if (CRk) {
    HIPCHK(c, qg::launch_cz(d, c->stream, d.AR, d.qs));
    HIPCHK(c, d2h_field(d.g, CRk, d.qs, d.g.nz, c->stream, c->stage));
  }
  if (CIk) {
    HIPCHK(c, qg::launch_cz(d, c->stream, d.AI, d.qs));
    HIPCHK(c, d2h_field(d.g, CIk, d.qs, d.g.nz, c->stream, c->stage));
  }
  return QG_OK;
}

int qg_step(qg_ctx* c, int32_t nsteps) {
  int rc = check_ctx(c);
  if (rc) return rc;
  if (nsteps < 0) return fail(c, QG_ERR_ARG, "nsteps < 0");
  HIPCHK(c, hipMemsetAsync(c->d.flag, 0, sizeof(int), c->stream));
  for (int s = 0; s < nsteps; ++s) {
    rc = step_once(c);
    if (rc) return rc;
  }
  int flag = 0;
  HIPCHK(c, hipStreamWaitEvent(c->stream, c->evA, 0));  // join the side stream
  HIPCHK(c, hipMemcpyAsync(&flag, c->d.flag, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (flag) return fail(c, QG_ERR_NONFINITE, "non-finite value in the updated fields");
  return QG_OK;
}

int qg_iter(const qg_ctx* c, int32_t* iter) {
  if (!c || !iter) return QG_ERR_ARG;
  *iter = c->iter;
  return QG_OK;
}

int qg_diag(qg_ctx* c, double* out) {
  int rc = check_ctx(c);
  if (rc) return rc;
  if (!out) return fail(c, QG_ERR_ARG, "null output");
  int nb = 0;
  HIPCHK(c, qg::launch_diag(c->d, c->stream, &nb));
  const int nm = c->d.g.nm;
  std::vector<double> part((size_t)nb * nm * 6);
  HIPCHK(c, hipStreamSynchronize(c->stream));
  HIPCHK(c, hipMemcpy(part.data(), c->d.diag_part, part.size() * sizeof(double), hipMemcpyDeviceToHost));
  for (int m = 0; m < nm; ++m)
    for (int j = 0; j < 6; ++j) {
      double s = 0.0;
      for (int b = 0; b < nb; ++b) s += part[((size_t)m * nb + b) * 6 + j];
      out[m * 6 + j] = s / c->d.g.n3;
    }
  return QG_OK;
}

int qg_keep_tendencies(qg_ctx* c, int32_t on) {
  int rc = check_ctx(c);
  if (rc) return rc;
  Dev& d = c->d;
  const bool regular = (d.flags & qg::F_YBJ_REGULAR) != 0;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream2));
  if (on) {
    const size_t S = field_elems(d.g);
    for (int i = 0; i < 2; ++i) {
      if (!c->kz[i]) {
        if ((rc = dalloc(c, &c->kz[i], S)) != QG_OK) return fail(c, rc, "no memory for the kept tendencies");
        // the columns kx >= KX of the planes are never written by the step
        HIPCHK(c, hipMemset(c->kz[i], 0, S * sizeof(double2)));
      }
      if (!c->kt[i]) {
        if (regular) {
          c->kt[i] = i == 0 ? d.tR : d.tI;
        } else {
          if ((rc = dalloc(c, &c->kt[i], S)) != QG_OK) return fail(c, rc, "no memory for the kept tendencies");
          HIPCHK(c, hipMemset(c->kt[i], 0, S * sizeof(double2)));
        }
      }
    }
    if (!c->keep_tend) c->tend_valid = 0;
    d.Z[0] = c->kz[0];
    d.Z[1] = c->kz[1];
    d.tR = c->kt[0];
    d.tI = c->kt[1];
  } else {
    d.Z[0] = d.Z[1] = nullptr;
    if (!regular) d.tR = d.tI = nullptr;
    c->tend_valid = 0;
  }
  c->keep_tend = on ? 1 : 0;
  c->generation++;  // captured launch sequences change
  return QG_OK;
}

int qg_get_tendencies(qg_ctx* c, double* nBRk, double* nBIk, double* rBRk, double* rBIk) {
  int rc = check_ctx(c);
  if (rc) return rc;
  if (!c->keep_tend || !c->tend_valid) return fail(c, QG_ERR_STATE, "no tendencies kept since the last state change");
  Dev& d = c->d;
  double* dst[] = {nBRk, nBIk, rBRk, rBIk};
  const double2* src[] = {d.tR, d.tI, d.Z[0], d.Z[1]};
  HIPCHK(c, hipStreamWaitEvent(c->stream, c->evA, 0));
  for (int i = 0; i < 4; ++i)
    if (dst[i]) HIPCHK(c, d2h_field(d.g, dst[i], src[i], d.g.nz, c->stream, c->stage));
  return QG_OK;
}

int qg_wave_conversion(qg_ctx* c, double* out) {
  int rc = check_ctx(c);
  if (rc) return rc;
  if (!out) return fail(c, QG_ERR_ARG, "null output");
  if (!c->keep_tend || !c->tend_valid) return fail(c, QG_ERR_STATE, "no tendencies kept since the last state change");
  int nb = 0;
  HIPCHK(c, hipStreamWaitEvent(c->stream, c->evA, 0));  // A^{n+1} from the side stream
  HIPCHK(c, qg::launch_conv(c->d, c->stream, &nb));
  const int nm = c->d.g.nm;
  std::vector<double> part((size_t)nb * nm * 6);
  HIPCHK(c, hipStreamSynchronize(c->stream));
  HIPCHK(c, hipMemcpy(part.data(), c->d.diag_part, part.size() * sizeof(double), hipMemcpyDeviceToHost));
  // volume means of the physical-space sums (sum_x = n1 n2 sum_k), then the
  // normalisations of diagnostics.f90:1071-1072 and :1249-1251 without the
  // dimensional scales Uw_scale^2 and Uw_scale^2 U_scale / L_scale
  const double n3 = c->d.g.n3, Bu = c->d.Bu;
  for (int m = 0; m < nm; ++m) {
    double S[3] = {0.0, 0.0, 0.0};
    for (int j = 0; j < 3; ++j)
      for (int b = 0; b < nb; ++b) S[j] += part[((size_t)m * nb + b) * 6 + j];
    for (int j = 0; j < 3; ++j) S[j] /= n3;
    double* o = out + (size_t)m * 5;
    o[0] = S[0] / (2.0 * Bu);          // we_conversion ca
    o[1] = 0.5 * S[1] / (2.0 * Bu);    // we_conversion cr
    o[2] = -0.5 * S[0] / Bu;           // gamma_conversion ca
    o[3] = -0.25 * S[1] / Bu;          // gamma_conversion cr
    o[4] = S[2] / Bu;                  // gamma_conversion cd
  }
  return QG_OK;
}

int qg_wave_vave(qg_ctx* c, double* out) {
  int rc = check_ctx(c);
  if (rc) return rc;
  if (!out) return fail(c, QG_ERR_ARG, "null output");
  Dev& d = c->d;
  const Geo& g = d.g;
  const size_t nsp = field_elems(g), nph = (size_t)g.n1 * g.n2 * g.nz;
  // the omega work fields: two physical fields for BR, BI and a third for the map
  if (!c->omega_blk && (rc = dalloc(c, &c->omega_blk, 5 * nsp + 2 * nph)) != QG_OK)
    return fail(c, rc, "no memory for the diagnostic work fields");
  double* pb = reinterpret_cast<double*>(c->omega_blk + 5 * nsp);
  hipError_t e = qg::launch_vave(d, c->stream, pb, pb + nph, pb + 2 * nph);
  if (e == hipSuccess)
    e = hipMemcpyAsync(out, pb + 2 * nph, (size_t)g.n1 * g.n2 * g.nm * sizeof(double), hipMemcpyDeviceToHost,
                       c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) return fail_hip(c, e, "qg_wave_vave");
  return QG_OK;
}

int qg_omega(qg_ctx* c, double* wk) {
  int rc = check_ctx(c);
  if (rc) return rc;
  if (!wk) return fail(c, QG_ERR_ARG, "null output");
  Dev& d = c->d;
  const Geo& g = d.g;
  const size_t nsp = field_elems(g), nph = (size_t)g.n1 * g.n2 * g.nz;
  // work fields of this diagnostic, kept for the next call (nph is even: the
  // four physical fields take 2 nph complex slots)
  if (!c->omega_blk && (rc = dalloc(c, &c->omega_blk, 5 * nsp + 2 * nph)) != QG_OK)
    return fail(c, rc, "no memory for the omega work fields");
  double2* sp[4];
  double* ph[4];
  double2* base = c->omega_blk;
  for (int i = 0; i < 4; ++i) sp[i] = base + i * nsp;
  double2* w = base + 4 * nsp;
  double* pb = reinterpret_cast<double*>(base + 5 * nsp);
  for (int i = 0; i < 4; ++i) ph[i] = pb + i * nph;
  hipError_t e = hipMemsetAsync(w, 0, nsp * sizeof(double2), c->stream);
  if (e == hipSuccess) e = qg::launch_omega(d, c->stream, sp, ph, w);
  if (e == hipSuccess) e = d2h_field(g, wk, w, g.nz, c->stream, c->stage);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) return fail_hip(c, e, "qg_omega");
  return QG_OK;
}

int qg_hspec_waves(qg_ctx* c, const int32_t* levels, int32_t nlev, double* out) {
  int rc = check_ctx(c);
  if (rc) return rc;
  if (!levels || !out || nlev < 1) return fail(c, QG_ERR_ARG, "bad spectra arguments");
  const Geo& g = c->d.g;
  for (int i = 0; i < nlev; ++i)
    if (levels[i] < 1 || levels[i] > g.n3) return fail(c, QG_ERR_ARG, "spectrum level outside 1 .. n3");
  const int nbin = g.n1 / 2, nm = g.nm;
  const size_t per = (size_t)nbin * 3;
  const size_t nout = (size_t)nm * qg::SpecLevels::MAXL * per;
  if (!c->spec_blk && (rc = dalloc(c, &c->spec_blk, nout * (1 + qg::hspec_chunks()))) != QG_OK)
    return fail(c, rc, "no memory for the spectra");
  double* dout = c->spec_blk;
  std::vector<double> h((size_t)nm * qg::SpecLevels::MAXL * per);
  hipError_t e = hipSuccess;
  for (int i0 = 0; i0 < nlev && e == hipSuccess; i0 += qg::SpecLevels::MAXL) {
    qg::SpecLevels L;
    L.n = std::min(nlev - i0, qg::SpecLevels::MAXL);
    for (int i = 0; i < L.n; ++i) L.lev[i] = levels[i0 + i];
    e = qg::launch_hspec_waves(c->d, c->stream, L, nbin, dout, dout + nout);
    if (e == hipSuccess) e = hipMemcpyAsync(h.data(), dout, (size_t)nm * L.n * per * sizeof(double), hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    for (int m = 0; e == hipSuccess && m < nm; ++m)
      std::copy(h.begin() + (size_t)m * L.n * per, h.begin() + (size_t)(m + 1) * L.n * per,
                out + ((size_t)m * nlev + i0) * per);
  }
  HIPCHK(c, e);
  return QG_OK;
}

int qg_init_q(qg_ctx* c, const double* psik, double* qk) {
  int rc = check_ctx(c);
  if (rc) return rc;
  if (!psik || !qk) return fail(c, QG_ERR_ARG, "null array");
  Dev& d = c->d;
  HIPCHK(c, qg::launch_clear_outside(d, c->stream, d.qs, 1));
  HIPCHK(c, h2d_field(d.g, d.qs, psik, d.g.nz, c->stream, c->stage));
  HIPCHK(c, qg::launch_init_q(d, c->stream, d.qs, d.dq));
  HIPCHK(c, d2h_field(d.g, qk, d.dq, d.g.nz, c->stream, c->stage));
  return QG_OK;
}

int qg_fft_r2c(qg_ctx* c, const double* r, int32_t ldx, double* k, int32_t nz) {
  int rc = check_ctx(c);
  if (rc) return rc;
  Dev& d = c->d;
  const Geo& g = d.g;
  if (!r || !k || ldx < g.n1 || nz < 1 || nz > g.nz) return fail(c, QG_ERR_ARG, "bad r2c arguments");
  const size_t nr = (size_t)ldx * g.n2 * nz;
  double* tmp = nullptr;
  HIPCHK(c, hipMalloc(&tmp, nr * sizeof(double)));
  hipError_t e = hipMemcpyAsync(tmp, r, nr * sizeof(double), hipMemcpyHostToDevice, c->stream);
  if (e == hipSuccess) e = qg::launch_r2c(d, c->stream, tmp, ldx, d.dq, nz);
  if (e == hipSuccess) e = d2h_field(g, k, d.dq, nz, c->stream, c->stage);
  hipFree(tmp);
  if (e != hipSuccess) return fail_hip(c, e, "qg_fft_r2c");
  c->p_valid = 0;
  return QG_OK;
}

int qg_fft_c2r(qg_ctx* c, const double* k, double* r, int32_t ldx, int32_t nz) {
  int rc = check_ctx(c);
  if (rc) return rc;
  Dev& d = c->d;
  const Geo& g = d.g;
  if (!r || !k || ldx < g.n1 || nz < 1 || nz > g.nz) return fail(c, QG_ERR_ARG, "bad c2r arguments");
  const size_t nr = (size_t)ldx * g.n2 * nz;
  double* tmp = nullptr;
  HIPCHK(c, hipMalloc(&tmp, nr * sizeof(double)));
  hipError_t e = hipMemsetAsync(tmp, 0, nr * sizeof(double), c->stream);
  if (e == hipSuccess) e = h2d_field(g, d.dq, k, nz, c->stream, c->stage);
  if (e == hipSuccess) e = qg::launch_c2r(d, c->stream, d.dq, tmp, ldx, nz);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e == hipSuccess) e = hipMemcpy(r, tmp, nr * sizeof(double), hipMemcpyDeviceToHost);
  hipFree(tmp);
  if (e != hipSuccess) return fail_hip(c, e, "qg_fft_c2r");
  c->p_valid = 0;
  return QG_OK;
}

int qg_device_field(qg_ctx* c, int32_t which, void** dptr) {
  if (!c || !dptr || which < 0 || which >= QG_NFIELDS) return QG_ERR_ARG;
  Dev& d = c->d;
  double2* f[] = {d.psi, d.q, d.qo, d.BR, d.BI, d.BRo, d.BIo, d.AR, d.AI};
  *dptr = f[which];
  return QG_OK;
}

int qg_set_variants(qg_ctx* c, const int32_t* v, int32_t n) {
  if (!c || !v || n < 0) return QG_ERR_ARG;
  for (int i = 0; i < n && i < QG_NVARIANTS; ++i) {
    const int hi = i < 6 ? 3 : 1;
    if (v[i] < 0 || v[i] > hi) return fail(c, QG_ERR_ARG, "qg_set_variants: value out of range");
  }
  // the A inversions change stream: let any in-flight side-stream work end
  if (n > 6) HIPCHK(c, hipStreamWaitEvent(c->stream, c->evA, 0));
  c->generation++;
  for (int i = 0; i < n && i < QG_NVARIANTS; ++i) {
    if (i < 5) c->d.tune.e16[i] = v[i];
    else if (i == 5) c->d.tune.solve = v[i];
    else c->d.tune.a_main = v[i];
  }
  return QG_OK;
}

int qg_get_variants(const qg_ctx* c, int32_t* v, int32_t n) {
  if (!c || !v || n < 0) return QG_ERR_ARG;
  for (int i = 0; i < n && i < QG_NVARIANTS; ++i)
    v[i] = i < 5 ? c->d.tune.e16[i] : (i == 5 ? c->d.tune.solve : c->d.tune.a_main);
  return QG_OK;
}

int qg_device_pitch(const qg_ctx* c, int32_t* pitch) {
  if (!c || !pitch) return QG_ERR_ARG;
  *pitch = c->d.g.sp;
  return QG_OK;
}

int qg_stream(qg_ctx* c, void** stream) {
  if (!c || !stream) return QG_ERR_ARG;
  *stream = (void*)c->stream;
  return QG_OK;
}

int qg_sync(qg_ctx* c) {
  int rc = check_ctx(c);
  if (rc) return rc;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return QG_OK;
}

int qg_timing(qg_ctx* c, int32_t mask) {
  if (!c) return QG_ERR_ARG;
  c->timing = (uint32_t)mask & ((1u << qg::K_NKERNELS) - 1u);
  return QG_OK;
}

int qg_timing_sample(qg_ctx* c, int32_t every) {
  if (!c || every < 1) return QG_ERR_ARG;
  c->timing_every = every;
  for (int i = 0; i < qg::K_NKERNELS; ++i) c->tseen[i] = 0;
  return QG_OK;
}

int qg_kernel_times(qg_ctx* c, double* ms, int32_t* counts, int32_t n) {
  int rc = check_ctx(c);
  if (rc) return rc;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  for (auto& e : c->events) {
    float t = 0.f;
    if (hipEventElapsedTime(&t, e.a, e.b) == hipSuccess && e.id >= 0 && e.id < qg::K_NKERNELS) {
      c->tms[e.id] += t;
      c->tcnt[e.id] += 1;
    }
    hipEventDestroy(e.a);
    hipEventDestroy(e.b);
  }
  c->events.clear();
  for (int i = 0; i < n && i < qg::K_NKERNELS; ++i) {
    if (ms) ms[i] = c->tms[i];
    if (counts) counts[i] = c->tcnt[i];
  }
  for (int i = 0; i < qg::K_NKERNELS; ++i) {
    c->tms[i] = 0.0;
    c->tcnt[i] = 0;
  }
  return QG_OK;
}

}  // extern "C"
