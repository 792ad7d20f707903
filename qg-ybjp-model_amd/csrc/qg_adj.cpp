// Tangent-linear and adjoint of the QG-YBJ+ step, with checkpointing.
//
// The forward step (main_waqg.f90:197-467) maps the state
// S = (psi, q, qo, BR, BI, BRo, BIo, AR, AI) to the next one; it is quadratic
// in S (bilinear products of convol_waqg / refraction_waqg / compute_qw, the
// rest linear).  Its linearisation about a forward trajectory is written out
// here stage by stage, and its adjoint is the stage-by-stage transpose under
// the inner products
//   spectral:  <F, G> = sum over L = 1 modes of Re(conj(F_k) G_k)
//   physical:  <f, g> = sum over grid points of f g.
// With these, for a half-spectrum mode k with Hermitian weight W_k (2, and 1
// at kx = ky = 0; the kx = 0, ky < 0 modes are L = 0 and refilled by realit):
//   c2r restricted to L (f = sum_k W_k Re(F_k e^{ik.x}))  has adjoint  N W L r2c
//   r2c restricted to L (F_k = N^-1 sum_x f e^{-ik.x})   has adjoint  N^-1 c2r(L F / W)
//   a spectral multiplier m(k)                           has adjoint  conj(m(k))
//   the psi and A inversions (symmetric tridiagonal T)   are self-adjoint
//   init_q (T / dz^2, symmetric)                          is self-adjoint
// with N = n1 n2.  Every forward physical field the linearisation needs is
// recomputed from stored spectral states (psi, q, BR, BI) of the trajectory.
//
// Checkpointing: the forward run stores the full member state every K steps;
// the adjoint sweep restores each checkpoint, recomputes its K steps with the
// forward kernels (storing psi, q, BR, BI of every step) and runs the adjoint
// steps of that segment backwards.
//
// Scope: nuz = 0, flags 0 or QG_NO_FEEDBACK; a batch context gives the
// members' independent gradients (and objectives) in one pass.
// No counterpart in the reference.
#include <hip/hip_runtime.h>

#include <cmath>
#include <string>
#include <vector>

#include "qg_adj.hpp"
#include "qg_common.hpp"
#include "qg_ctx.hpp"
#include "qgybj.h"

using qg::Dev;
using qg::PFmaArgs;
using qg::SMulArgs;
using qg::SMulTerm;

namespace {

enum { F_PSI = 0, F_Q, F_QO, F_BR, F_BI, F_BRO, F_BIO, F_AR, F_AI, NSTATE };

}  // namespace

struct qg_adj {
  qg_ctx* c = nullptr;
  Dev w{};  // the member's view with its own transform temporary
  hipStream_t st = nullptr;
  int K = 20;
  int zhi = 0;  // deep levels: [0, zhi)
  double zdeep = 0.0;
  size_t S = 0, NP = 0;
  std::vector<void*> allocs;
  // linearised state (tangent-linear perturbation or adjoint) and work
  double2* x[NSTATE] = {};
  double2* s[12] = {};
  double* p[24] = {};
  double2* F[6] = {};  // fused adjoint: forward (kx, y) planes of step n
  double2* F1[6] = {};  // and of B^{n+1} (entries 2..5: BR', ky BR', BI', ky BI')
  double2* A[6] = {};  // fused adjoint: adjoint (kx, y) planes
  int fused = 1;       // 1: fused adjoint kernels, 0: generic kernels
  double* part = nullptr;
  // trajectory storage: checkpoints (full member state) and one segment
  std::vector<double2*> ckpt;     // 12 fields each
  std::vector<int> ckpt_meta;     // iter, p_valid, clear_pending per checkpoint
  double2* seg = nullptr;         // (K + 1) x 4 fields: psi, q, BR, BI
  double2* fin = nullptr;         // final state (12 fields)
  int fin_meta[3] = {0, 0, 0};
  int last_seg = -1;  // checkpoint index whose segment the forward run already stored
  std::string err;
};

namespace {

int afail(qg_adj* a, int code, const std::string& m) {
  if (a) a->err = m;
  return code;
}
#define AHIP(a, call)                                                                                \
  do {                                                                                               \
    hipError_t _r = (call);                                                                          \
    if (_r != hipSuccess)                                                                            \
      return afail((a), QG_ERR_HIP, std::string("HIP error '") + hipGetErrorString(_r) + "' in " #call); \
  } while (0)
#define ARC(call)                 \
  do {                            \
    int _rc = (call);             \
    if (_rc != QG_OK) return _rc; \
  } while (0)

template <class T>
int aalloc(qg_adj* a, T** ptr, size_t n) {
  void* v = nullptr;
  hipError_t e = hipMalloc(&v, n * sizeof(T));
  if (e != hipSuccess) return afail(a, QG_ERR_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
  a->allocs.push_back(v);
  hipMemsetAsync(v, 0, n * sizeof(T), a->st);
  *ptr = static_cast<T*>(v);
  return QG_OK;
}

// ---- building blocks --------------------------------------------------------
SMulTerm term(const double2* in, double amul = 1.0) {
  SMulTerm t;
  t.in = in;
  t.a = amul;
  return t;
}
// multipliers: i kx, i ky, -kh2 (and their scalings via a)
SMulTerm t_ikx(const double2* in, double amul = 1.0) {
  SMulTerm t = term(in, amul);
  t.c0 = 0.0;
  t.ckx = 1.0;
  t.imag = 1;
  return t;
}
SMulTerm t_iky(const double2* in, double amul = 1.0) {
  SMulTerm t = term(in, amul);
  t.c0 = 0.0;
  t.cky = 1.0;
  t.imag = 1;
  return t;
}
SMulTerm t_kh2(const double2* in, double amul = 1.0) {
  SMulTerm t = term(in, amul);
  t.c0 = 0.0;
  t.ckh = 1.0;
  return t;
}
SMulTerm tab(SMulTerm t, int table) {
  t.tab = table;
  return t;
}
SMulTerm wpow(SMulTerm t, int wp) {
  t.wp = wp;
  return t;
}

struct Ops {
  qg_adj* a;
  const Dev& d() const { return a->c->d; }
  int n1() const { return a->w.g.n1; }
  double N() const { return (double)a->w.g.n1 * (double)a->w.g.n2; }

  int smul(double2* out, std::initializer_list<SMulTerm> ts, int accumulate = 0, int zlo = 0,
           int zhi = 1 << 30) const {
    SMulArgs s;
    int j = 0;
    for (const auto& t : ts) s.t[j++] = t;
    s.accumulate = accumulate;
    s.zlo = zlo;
    s.zhi = zhi;
    AHIP(a, qg::launch_smul(a->w, a->st, out, s));
    return QG_OK;
  }
  int pfma(double* out, std::initializer_list<const double*> xs, std::initializer_list<const double*> ys,
           std::initializer_list<double> ss, int accumulate = 0) const {
    PFmaArgs f;
    int j = 0;
    for (auto v : xs) f.x[j++] = v;
    j = 0;
    for (auto v : ys) f.y[j++] = v;
    j = 0;
    for (auto v : ss) f.s[j++] = v;
    f.accumulate = accumulate;
    AHIP(a, qg::launch_pfma(a->st, out, a->NP, f));
    return QG_OK;
  }
  int c2r(const double2* k, double* r) const {
    AHIP(a, qg::launch_c2r(a->w, a->st, k, r, n1(), a->w.g.nz));
    return QG_OK;
  }
  int r2c(const double* r, double2* k) const {
    AHIP(a, qg::launch_r2c(a->w, a->st, r, n1(), k, a->w.g.nz));
    return QG_OK;
  }
  // c2r of a spectral combination (through the spectral temporary s[11])
  int c2r_of(double* r, std::initializer_list<SMulTerm> ts) const {
    ARC(smul(a->s[11], ts));
    return c2r(a->s[11], r);
  }
  int solve_psi(const double2* in, double2* out) const {
    Dev dd = a->c->d;
    dd.psi = out;
    AHIP(a, qg::launch_solve(dd, a->st, 0, 0, in));
    return QG_OK;
  }
  int solve_a(const double2* inR, double2* outR, const double2* inI, double2* outI) const {
    Dev dd = a->c->d;
    dd.BR = const_cast<double2*>(inR);
    dd.AR = outR;
    dd.BI = const_cast<double2*>(inI);
    dd.AI = outI;
    AHIP(a, qg::launch_solve(dd, a->st, 1, 2, nullptr));
    return QG_OK;
  }
  int copy(double2* dst, const double2* src) const {
    AHIP(a, hipMemcpyAsync(dst, src, a->S * sizeof(double2), hipMemcpyDeviceToDevice, a->st));
    return QG_OK;
  }
  int zero(double2* f) const {
    AHIP(a, hipMemsetAsync(f, 0, a->S * sizeof(double2), a->st));
    return QG_OK;
  }
};

// physical fields of a stored state: u, v, zeta (from L psi), q, BR, BI ->
// p[0..5].  q, BR, BI as stored before the Euler step (the initial B is not
// dealiased), their L = 1 modes afterwards (stored slots keep only the box).
int fwd_fields_n(const Ops& o, const double2* psi, const double2* q, const double2* BR, const double2* BI,
                 bool euler) {
  qg_adj* a = o.a;
  ARC(o.c2r_of(a->p[0], {t_iky(psi, -1.0)}));
  ARC(o.c2r_of(a->p[1], {t_ikx(psi)}));
  ARC(o.c2r_of(a->p[2], {t_kh2(psi, -1.0)}));
  if (euler) {
    ARC(o.c2r(q, a->p[3]));
    ARC(o.c2r(BR, a->p[4]));
    ARC(o.c2r(BI, a->p[5]));
  } else {
    ARC(o.c2r_of(a->p[3], {term(q)}));
    ARC(o.c2r_of(a->p[4], {term(BR)}));
    ARC(o.c2r_of(a->p[5], {term(BI)}));
  }
  return QG_OK;
}
// gradient fields of the new B for the feedback term: BRx, BRy, BIx, BIy,
// BR, BI -> p[6..11]
int fwd_fields_w(const Ops& o, const double2* BR, const double2* BI) {
  qg_adj* a = o.a;
  ARC(o.c2r_of(a->p[6], {t_ikx(BR)}));
  ARC(o.c2r_of(a->p[7], {t_iky(BR)}));
  ARC(o.c2r_of(a->p[8], {t_ikx(BI)}));
  ARC(o.c2r_of(a->p[9], {t_iky(BI)}));
  ARC(o.c2r_of(a->p[10], {term(BR)}));
  ARC(o.c2r_of(a->p[11], {term(BI)}));
  return QG_OK;
}

struct Coef {
  double dt, c, g, rw;
  bool feedback;
};
Coef coefs(const qg_adj* a) {
  const Dev& d = a->c->d;
  return {d.dt, 0.5 / (d.Bu * d.Ro), d.gamma, d.Ro * d.W2F, !(d.flags & qg::F_NO_FEEDBACK)};
}

// ---- tangent-linear step ------------------------------------------------------
// x (perturbation at step n) -> x at step n+1, about the forward states
// (psi, q, BR, BI) at n and (BR, BI) at n+1.
int tlm_step(qg_adj* a, const double2* const* fn, const double2* BR1, const double2* BI1, bool euler) {
  Ops o{a};
  const Coef k = coefs(a);
  double2** x = a->x;
  double2** s = a->s;
  double** p = a->p;
  ARC(fwd_fields_n(o, fn[0], fn[1], fn[2], fn[3], euler));
  // perturbation fields p[12..17]: du, dv, dzeta, dq, dBR, dBI
  ARC(o.c2r_of(p[12], {t_iky(x[F_PSI], -1.0)}));
  ARC(o.c2r_of(p[13], {t_ikx(x[F_PSI])}));
  ARC(o.c2r_of(p[14], {t_kh2(x[F_PSI], -1.0)}));
  ARC(o.c2r(x[F_Q], p[15]));
  ARC(o.c2r(x[F_BR], p[16]));
  ARC(o.c2r(x[F_BI], p[17]));
  // n_X = L (i kx r2c(u X) + i ky r2c(v X)) linearised, X = q, BR, BI -> s[0..2]
  for (int j = 0; j < 3; ++j) {
    const double* X = p[3 + j];
    const double* dX = p[15 + j];
    ARC(o.pfma(p[18], {p[12], p[0]}, {X, dX}, {1.0, 1.0}));
    ARC(o.pfma(p[19], {p[13], p[1]}, {X, dX}, {1.0, 1.0}));
    ARC(o.r2c(p[18], s[9]));
    ARC(o.r2c(p[19], s[10]));
    ARC(o.smul(s[j], {t_ikx(s[9]), t_iky(s[10])}));
  }
  // r_BR, r_BI = L r2c(zeta B) linearised -> s[3], s[4]
  for (int j = 0; j < 2; ++j) {
    ARC(o.pfma(p[18], {p[14], p[2]}, {p[4 + j], p[16 + j]}, {1.0, 1.0}));
    ARC(o.r2c(p[18], s[9]));
    ARC(o.smul(s[3 + j], {term(s[9])}));
  }
  const double2 *nq = s[0], *nR = s[1], *nI = s[2], *rR = s[3], *rI = s[4];
  // update -> s[5] (q+), s[6] (BR+), s[7] (BI+)
  if (euler) {
    ARC(o.smul(s[5], {tab(term(x[F_Q]), qg::TAB_E1Q), tab(term(nq, -k.dt), qg::TAB_E1Q)}));
    ARC(o.smul(s[6], {tab(term(x[F_BR]), qg::TAB_E1W), tab(term(nR, -k.dt), qg::TAB_E1W),
                      tab(t_kh2(x[F_AI], -k.dt * k.c), qg::TAB_E1W), tab(term(rI, 0.5 * k.dt), qg::TAB_E1W)}));
    ARC(o.smul(s[7], {tab(term(x[F_BI]), qg::TAB_E1W), tab(term(nI, -k.dt), qg::TAB_E1W),
                      tab(t_kh2(x[F_AR], k.dt * k.c), qg::TAB_E1W), tab(term(rR, -0.5 * k.dt), qg::TAB_E1W)}));
    ARC(o.copy(x[F_QO], x[F_Q]));
    ARC(o.copy(x[F_BRO], x[F_BR]));
    ARC(o.copy(x[F_BIO], x[F_BI]));
  } else {
    ARC(o.smul(s[5], {tab(term(x[F_QO]), qg::TAB_E2Q), tab(term(nq, -2.0 * k.dt), qg::TAB_E1Q)}));
    ARC(o.smul(s[6], {tab(term(x[F_BRO]), qg::TAB_E2W), tab(term(nR, -2.0 * k.dt), qg::TAB_E1W),
                      tab(t_kh2(x[F_AI], -2.0 * k.dt * k.c), qg::TAB_E1W), tab(term(rI, k.dt), qg::TAB_E1W)}));
    ARC(o.smul(s[7], {tab(term(x[F_BIO]), qg::TAB_E2W), tab(term(nI, -2.0 * k.dt), qg::TAB_E1W),
                      tab(t_kh2(x[F_AR], 2.0 * k.dt * k.c), qg::TAB_E1W), tab(term(rR, -k.dt), qg::TAB_E1W)}));
    // Robert-Asselin filter
    ARC(o.smul(x[F_QO], {term(x[F_Q], 1.0 - 2.0 * k.g), term(x[F_QO], k.g), term(s[5], k.g)}));
    ARC(o.smul(x[F_BRO], {term(x[F_BR], 1.0 - 2.0 * k.g), term(x[F_BRO], k.g), term(s[6], k.g)}));
    ARC(o.smul(x[F_BIO], {term(x[F_BI], 1.0 - 2.0 * k.g), term(x[F_BIO], k.g), term(s[7], k.g)}));
  }
  ARC(o.copy(x[F_Q], s[5]));
  ARC(o.copy(x[F_BR], s[6]));
  ARC(o.copy(x[F_BI], s[7]));
  // feedback: q* = q - q^w -> s[8]
  if (k.feedback) {
    ARC(fwd_fields_w(o, BR1, BI1));
    ARC(o.c2r_of(p[12], {t_ikx(x[F_BR])}));
    ARC(o.c2r_of(p[13], {t_iky(x[F_BR])}));
    ARC(o.c2r_of(p[14], {t_ikx(x[F_BI])}));
    ARC(o.c2r_of(p[15], {t_iky(x[F_BI])}));
    ARC(o.c2r(x[F_BR], p[16]));
    ARC(o.c2r(x[F_BI], p[17]));
    // J = BRy BIx - BRx BIy ; E = BR^2 + BI^2
    ARC(o.pfma(p[18], {p[13], p[7], p[12], p[6]}, {p[8], p[14], p[9], p[15]}, {1.0, 1.0, -1.0, -1.0}));
    ARC(o.pfma(p[19], {p[10], p[11]}, {p[16], p[17]}, {2.0, 2.0}));
    ARC(o.r2c(p[18], s[9]));
    ARC(o.r2c(p[19], s[10]));
    ARC(o.smul(s[8], {term(x[F_Q]), term(s[9], -k.rw), t_kh2(s[10], 0.25 * k.rw)}));
  } else {
    ARC(o.smul(s[8], {term(x[F_Q])}));
  }
  ARC(o.solve_psi(s[8], x[F_PSI]));
  ARC(o.solve_a(x[F_BR], x[F_AR], x[F_BI], x[F_AI]));
  return QG_OK;
}

// ---- adjoint step -------------------------------------------------------------
// x (adjoint of the state at step n+1) -> x (adjoint of the state at step n).
int adj_step(qg_adj* a, const double2* const* fn, const double2* BR1, const double2* BI1, bool euler) {
  Ops o{a};
  const Coef k = coefs(a);
  const double N = o.N();
  double2** x = a->x;
  double2** s = a->s;
  double** p = a->p;
  // 1. inversions: qbar* = S_psi(psibar'), BR+bar = S_A(ARbar'), BI+bar = S_A(AIbar')
  ARC(o.solve_psi(x[F_PSI], s[8]));
  ARC(o.solve_a(x[F_AR], s[6], x[F_AI], s[7]));
  // 2. q* = L (q+ - q^w): q+bar = qbar' + qbar*, q^w bar = -qbar* (s[8] keeps qbar*)
  ARC(o.smul(s[5], {term(x[F_Q]), term(s[8])}));
  ARC(o.smul(s[6], {term(x[F_BR])}, 1));
  ARC(o.smul(s[7], {term(x[F_BI])}, 1));
  // 3. feedback: q^w = Ro W2F L (r2c(J) - kh2/4 r2c(E))
  if (k.feedback) {
    ARC(fwd_fields_w(o, BR1, BI1));
    // Jbar = N^-1 c2r(-Ro W2F qbar* / W), Ebar = N^-1 c2r(Ro W2F kh2/4 qbar* / W)
    ARC(o.c2r_of(p[12], {wpow(term(s[8], -k.rw / N), -1)}));
    ARC(o.c2r_of(p[13], {wpow(t_kh2(s[8], 0.25 * k.rw / N), -1)}));
    // BRybar = Jbar BIx, BIxbar = Jbar BRy, BRxbar = -Jbar BIy, BIybar = -Jbar BRx,
    // BRbar = 2 Ebar BR, BIbar = 2 Ebar BI
    ARC(o.pfma(p[14], {p[12]}, {p[9]}, {-1.0}));  // BRx bar
    ARC(o.pfma(p[15], {p[12]}, {p[8]}, {1.0}));   // BRy bar
    ARC(o.pfma(p[16], {p[13]}, {p[10]}, {2.0}));  // BR bar
    ARC(o.r2c(p[14], s[9]));
    ARC(o.r2c(p[15], s[10]));
    ARC(o.r2c(p[16], s[11]));
    ARC(o.smul(s[6], {wpow(t_ikx(s[9], -N), 1), wpow(t_iky(s[10], -N), 1), wpow(term(s[11], N), 1)}, 1));
    ARC(o.pfma(p[14], {p[12]}, {p[7]}, {1.0}));   // BIx bar
    ARC(o.pfma(p[15], {p[12]}, {p[6]}, {-1.0}));  // BIy bar
    ARC(o.pfma(p[16], {p[13]}, {p[11]}, {2.0}));  // BI bar
    ARC(o.r2c(p[14], s[9]));
    ARC(o.r2c(p[15], s[10]));
    ARC(o.r2c(p[16], s[11]));
    ARC(o.smul(s[7], {wpow(t_ikx(s[9], -N), 1), wpow(t_iky(s[10], -N), 1), wpow(term(s[11], N), 1)}, 1));
  }
  // 4./5. time update (and Robert filter): adjoints of the inputs q, qo, BR,
  // BRo, BI, BIo, AR, AI and of the tendencies nq, nBR, nBI, rBR, rBI
  double2 *nq = s[0], *nR = s[1], *nI = s[2], *rR = s[3], *rI = s[4];
  if (euler) {
    // q1 = L e1 (q - dt nq) ; qo' = q
    ARC(o.smul(nq, {tab(term(s[5], -k.dt), qg::TAB_E1Q)}));
    ARC(o.smul(x[F_Q], {term(x[F_QO]), tab(term(s[5]), qg::TAB_E1Q)}));
    ARC(o.smul(nR, {tab(term(s[6], -k.dt), qg::TAB_E1W)}));
    ARC(o.smul(rI, {tab(term(s[6], 0.5 * k.dt), qg::TAB_E1W)}));
    ARC(o.smul(nI, {tab(term(s[7], -k.dt), qg::TAB_E1W)}));
    ARC(o.smul(rR, {tab(term(s[7], -0.5 * k.dt), qg::TAB_E1W)}));
    ARC(o.smul(x[F_AI], {tab(t_kh2(s[6], -k.dt * k.c), qg::TAB_E1W)}));
    ARC(o.smul(x[F_AR], {tab(t_kh2(s[7], k.dt * k.c), qg::TAB_E1W)}));
    ARC(o.smul(x[F_BR], {term(x[F_BRO]), tab(term(s[6]), qg::TAB_E1W)}));
    ARC(o.smul(x[F_BI], {term(x[F_BIO]), tab(term(s[7]), qg::TAB_E1W)}));
    ARC(o.zero(x[F_QO]));
    ARC(o.zero(x[F_BRO]));
    ARC(o.zero(x[F_BIO]));
  } else {
    // Robert: qo' = L((1-2g) q + g qo + g q+) ; q' = q+ (already in s[5])
    ARC(o.smul(s[5], {term(x[F_QO], k.g)}, 1));
    ARC(o.smul(s[6], {term(x[F_BRO], k.g)}, 1));
    ARC(o.smul(s[7], {term(x[F_BIO], k.g)}, 1));
    ARC(o.smul(x[F_Q], {term(x[F_QO], 1.0 - 2.0 * k.g)}));
    ARC(o.smul(x[F_BR], {term(x[F_BRO], 1.0 - 2.0 * k.g)}));
    ARC(o.smul(x[F_BI], {term(x[F_BIO], 1.0 - 2.0 * k.g)}));
    // update: q+ = L(e2 qo - 2dt e1 nq), BR+ = L(e2w BRo - 2dt e1w (nBR + c kh2 AI - rBI/2)), ...
    ARC(o.smul(x[F_QO], {term(x[F_QO], k.g), tab(term(s[5]), qg::TAB_E2Q)}));
    ARC(o.smul(x[F_BRO], {term(x[F_BRO], k.g), tab(term(s[6]), qg::TAB_E2W)}));
    ARC(o.smul(x[F_BIO], {term(x[F_BIO], k.g), tab(term(s[7]), qg::TAB_E2W)}));
    ARC(o.smul(nq, {tab(term(s[5], -2.0 * k.dt), qg::TAB_E1Q)}));
    ARC(o.smul(nR, {tab(term(s[6], -2.0 * k.dt), qg::TAB_E1W)}));
    ARC(o.smul(rI, {tab(term(s[6], k.dt), qg::TAB_E1W)}));
    ARC(o.smul(nI, {tab(term(s[7], -2.0 * k.dt), qg::TAB_E1W)}));
    ARC(o.smul(rR, {tab(term(s[7], -k.dt), qg::TAB_E1W)}));
    ARC(o.smul(x[F_AI], {tab(t_kh2(s[6], -2.0 * k.dt * k.c), qg::TAB_E1W)}));
    ARC(o.smul(x[F_AR], {tab(t_kh2(s[7], 2.0 * k.dt * k.c), qg::TAB_E1W)}));
  }
  // 6. nonlinear terms: physical adjoints p[12..17] = ubar, vbar, zetabar,
  // qbar, BRbar, BIbar
  ARC(fwd_fields_n(o, fn[0], fn[1], fn[2], fn[3], euler));
  const double2* nX[3] = {nq, nR, nI};
  for (int j = 0; j < 3; ++j) {
    // a_x = N^-1 c2r(-i kx nX / W), a_y = N^-1 c2r(-i ky nX / W)
    ARC(o.c2r_of(p[18], {wpow(t_ikx(nX[j], -1.0 / N), -1)}));
    ARC(o.c2r_of(p[19], {wpow(t_iky(nX[j], -1.0 / N), -1)}));
    const double* X = p[3 + j];
    ARC(o.pfma(p[12], {p[18]}, {X}, {1.0}, j > 0));
    ARC(o.pfma(p[13], {p[19]}, {X}, {1.0}, j > 0));
    ARC(o.pfma(p[15 + j], {p[18], p[19]}, {p[0], p[1]}, {1.0, 1.0}));
  }
  // r_BR = L r2c(zeta BR), r_BI = L r2c(zeta BI)
  ARC(o.c2r_of(p[18], {wpow(term(rR, 1.0 / N), -1)}));
  ARC(o.c2r_of(p[19], {wpow(term(rI, 1.0 / N), -1)}));
  ARC(o.pfma(p[14], {p[18], p[19]}, {p[4], p[5]}, {1.0, 1.0}));
  ARC(o.pfma(p[16], {p[18]}, {p[2]}, {1.0}, 1));
  ARC(o.pfma(p[17], {p[19]}, {p[2]}, {1.0}, 1));
  // back to spectral space: psibar = N W L (i ky r2c(ubar) - i kx r2c(vbar) - kh2 r2c(zetabar))
  ARC(o.r2c(p[12], s[9]));
  ARC(o.r2c(p[13], s[10]));
  ARC(o.r2c(p[14], s[11]));
  ARC(o.smul(x[F_PSI], {wpow(t_iky(s[9], N), 1), wpow(t_ikx(s[10], -N), 1), wpow(t_kh2(s[11], -N), 1)}));
  ARC(o.r2c(p[15], s[9]));
  ARC(o.r2c(p[16], s[10]));
  ARC(o.r2c(p[17], s[11]));
  ARC(o.smul(x[F_Q], {wpow(term(s[9], N), 1)}, 1));
  ARC(o.smul(x[F_BR], {wpow(term(s[10], N), 1)}, 1));
  ARC(o.smul(x[F_BI], {wpow(term(s[11], N), 1)}, 1));
  return QG_OK;
}

// ---- fused adjoint step -------------------------------------------------------
// Same map as adj_step, with the transform work in the fused kernels of
// qg_adj_step.hip; the forward (kx, y) planes are recomputed from the stored
// spectral states with the step's own y-inverse kernel.
int adj_step_fused(qg_adj* a, const double2* const* fn, const double2* BR1, const double2* BI1, bool euler) {
  Ops o{a};
  const Coef k = coefs(a);
  const Dev& d = a->c->d;
  const auto& g = d.g;
  double2** x = a->x;
  double2** s = a->s;
  qg::AdjBufs b;
  for (int i = 0; i < 6; ++i) {
    b.F[i] = a->F[i];
    b.A[i] = a->A[i];
  }
  b.psi = x[F_PSI]; b.q = x[F_Q]; b.qo = x[F_QO]; b.BR = x[F_BR]; b.BI = x[F_BI];
  b.BRo = x[F_BRO]; b.BIo = x[F_BIO]; b.AR = x[F_AR]; b.AI = x[F_AI];
  b.qstar = s[8];
  b.sR = s[6];
  b.sI = s[7];
  {
    // the three inversions in one launch: psi from psibar', A from ARbar', AIbar'
    Dev dd = a->c->d;
    dd.psi = s[8];
    dd.BR = x[F_AR];
    dd.AR = s[6];
    dd.BI = x[F_AI];
    dd.AI = s[7];
    AHIP(a, qg::launch_solve(dd, a->st, 0, 2, x[F_PSI]));
  }
  // every forward (kx, y) plane the step needs, in one launch: those of step
  // n as row_nl reads them (the Euler step's fields are not yet
  // box-restricted) and those of B^{n+1} for the feedback term.  q, BR, BI
  // are adjacent in the segment slot, and so are F0, F2, F4; likewise BR',
  // BI' and F1[2], F1[4] and F1[3], F1[5].
  const int Win = euler ? g.iktx : g.KX;
  {
    qg::ColInvJobs jobs{};
    jobs.j[0] = qg::ColInvJob{fn[0], {a->F[1], a->F[3], a->F[5]}, g.KX, 1, 1, g.nz};
    jobs.j[1] = qg::ColInvJob{fn[1], {a->F[0], nullptr, nullptr}, Win, euler ? 0 : 1, 0, 3 * g.nz};
    jobs.n = 2;
    if (k.feedback) {
      jobs.j[2] = qg::ColInvJob{BR1, {a->F1[2], a->F1[3], nullptr}, g.KX, 1, 1, 2 * g.nz};
      jobs.n = 3;
    }
    AHIP(a, qg::launch_col_inv(d, a->st, jobs));
  }
  if (k.feedback) {
    AHIP(a, qg::launch_adj_col_qw(d, a->st, b));
    qg::AdjBufs b1 = b;
    for (int i = 2; i < 6; ++i) b1.F[i] = a->F1[i];
    AHIP(a, qg::launch_adj_row_qw(d, a->st, b1));
  }
  AHIP(a, qg::launch_adj_col_upd(d, a->st, b, euler ? 1 : 0, k.feedback ? 1 : 0));
  AHIP(a, qg::launch_adj_row_nl(d, a->st, b, Win, euler ? 1 : 0));
  AHIP(a, qg::launch_adj_col_out(d, a->st, b));
  return QG_OK;
}

// ---- trajectory ---------------------------------------------------------------
double2* const* member_fields12(qg_ctx* c, double2** out) {
  Dev& d = c->d;
  double2* f[12] = {d.psi, d.q, d.qo, d.BR, d.BI, d.BRo, d.BIo, d.AR, d.AI, d.I[0], d.I[2], d.I[4]};
  for (int i = 0; i < 12; ++i) out[i] = f[i];
  return out;
}

int save_state(qg_adj* a, double2* dst, int* meta) {
  qg_ctx* c = a->c;
  AHIP(a, hipStreamWaitEvent(c->stream, c->evA, 0));
  double2* f[12];
  member_fields12(c, f);
  for (int i = 0; i < 12; ++i)
    AHIP(a, hipMemcpyAsync(dst + (size_t)i * a->S, f[i], a->S * sizeof(double2), hipMemcpyDeviceToDevice, a->st));
  meta[0] = c->iter;
  meta[1] = c->p_valid;
  meta[2] = c->clear_pending;
  return QG_OK;
}

int load_state(qg_adj* a, const double2* src, const int* meta) {
  qg_ctx* c = a->c;
  AHIP(a, hipStreamWaitEvent(c->stream, c->evA, 0));
  double2* f[12];
  member_fields12(c, f);
  for (int i = 0; i < 12; ++i)
    AHIP(a, hipMemcpyAsync(f[i], src + (size_t)i * a->S, a->S * sizeof(double2), hipMemcpyDeviceToDevice, a->st));
  c->iter = meta[0];
  c->p_valid = meta[1];
  c->clear_pending = meta[2];
  return QG_OK;
}

// keep psi, q, BR, BI of the member in segment slot j: whole fields before
// the Euler step (the initial B is not yet box-restricted), the dealiasing
// box afterwards (the only part the adjoint step reads)
int keep_step(qg_adj* a, int j) {
  qg_ctx* c = a->c;
  AHIP(a, hipStreamWaitEvent(c->stream, c->evA, 0));
  const double2* f[4] = {c->d.psi, c->d.q, c->d.BR, c->d.BI};
  if (c->iter == 0) {
    for (int i = 0; i < 4; ++i)
      AHIP(a, hipMemcpyAsync(a->seg + ((size_t)j * 4 + i) * a->S, f[i], a->S * sizeof(double2),
                             hipMemcpyDeviceToDevice, a->st));
  } else {
    qg::CopyBoxArgs cb;
    for (int i = 0; i < 4; ++i) {
      cb.src[i] = f[i];
      cb.dst[i] = a->seg + ((size_t)j * 4 + i) * a->S;
    }
    cb.n = 4;
    AHIP(a, qg::launch_copy_box(c->d.g, a->st, cb));
  }
  return QG_OK;
}
const double2* seg_field(const qg_adj* a, int j, int i) { return a->seg + ((size_t)j * 4 + i) * a->S; }

int fstep(qg_adj* a) {
  int rc = qgi::step_once(a->c);
  if (rc != QG_OK) return afail(a, rc, "forward step: " + a->c->err);
  return QG_OK;
}

// forward run from physical initial conditions, checkpoints every K steps
int forward(qg_adj* a, const double* psir, const double* BRr, const double* BIr, int ldx, int nsteps,
            bool checkpoints) {
  qg_ctx* c = a->c;
  int rc = qg_set_state_physical(c, psir, BRr, BIr, ldx);
  if (rc != QG_OK) return afail(a, rc, "initial state: " + c->err);
  AHIP(a, hipMemsetAsync(c->d.flag, 0, sizeof(int), a->st));
  const int nck = checkpoints ? (nsteps + a->K - 1) / a->K : 0;
  while ((int)a->ckpt.size() < nck) {
    double2* b = nullptr;
    ARC(aalloc(a, &b, 12 * a->S));
    a->ckpt.push_back(b);
  }
  a->ckpt_meta.assign(3 * (size_t)nck, 0);
  // the last segment's states are kept as the forward run passes them, so
  // the adjoint sweep does not recompute that segment
  const int n_last = checkpoints ? (nck - 1) * a->K : nsteps;
  a->last_seg = -1;
  for (int n = 0; n < nsteps; ++n) {
    if (checkpoints && n % a->K == 0) ARC(save_state(a, a->ckpt[n / a->K], &a->ckpt_meta[3 * (n / a->K)]));
    if (n >= n_last) ARC(keep_step(a, n - n_last));
    ARC(fstep(a));
  }
  if (checkpoints) {
    ARC(keep_step(a, nsteps - n_last));
    a->last_seg = nck - 1;
  }
  AHIP(a, hipStreamWaitEvent(c->stream, c->evA, 0));
  int flag = 0;
  AHIP(a, hipMemcpyAsync(&flag, c->d.flag, sizeof(int), hipMemcpyDeviceToHost, a->st));
  AHIP(a, hipStreamSynchronize(a->st));
  if (flag) return afail(a, QG_ERR_NONFINITE, "non-finite value in the forward run");
  return QG_OK;
}

// J of every member: 1/2 (2 pi)^2 dz sum over the deep levels of W kh2 (|AR|^2 + |AI|^2)
int objective(qg_adj* a, const double2* AR, const double2* AI, double* J) {
  const auto& g = a->w.g;
  Dev dm = a->w;  // one member's view
  dm.g.nm = 1;
  dm.g.nz = g.n3;
  const int nb = qg::obj_blocks(dm.g);
  const size_t moff = (size_t)g.n3 * g.ikty * g.sp;
  for (int m = 0; m < g.nm; ++m)
    AHIP(a, qg::launch_obj(dm, a->st, AR + m * moff, AI + m * moff, 0, a->zhi, a->part + (size_t)m * nb));
  std::vector<double> h((size_t)nb * g.nm);
  AHIP(a, hipMemcpyAsync(h.data(), a->part, h.size() * sizeof(double), hipMemcpyDeviceToHost, a->st));
  AHIP(a, hipStreamSynchronize(a->st));
  const double twopi = 4.0 * std::asin(1.0);
  for (int m = 0; m < g.nm; ++m) {
    double sum = 0.0;
    for (int b = 0; b < nb; ++b) sum += h[(size_t)m * nb + b];
    J[m] = 0.5 * twopi * twopi * a->c->d.dz * sum;
  }
  return QG_OK;
}

// physical work fields of the generic (unfused) linearised step, allocated
// on first use
int ensure_generic(qg_adj* a) {
  for (int i = 0; i < 24; ++i)
    if (!a->p[i]) ARC(aalloc(a, &a->p[i], a->NP));
  return QG_OK;
}

// adjoint sweep over nsteps given x = adjoint of the final state; leaves the
// adjoint of the initial (psi-hat, q) in x[F_PSI], x[F_Q]
int sweep(qg_adj* a, int nsteps) {
  if (!a->fused) ARC(ensure_generic(a));
  const int nck = (nsteps + a->K - 1) / a->K;
  for (int cidx = nck - 1; cidx >= 0; --cidx) {
    const int n0 = cidx * a->K;
    const int len = std::min(a->K, nsteps - n0);
    if (cidx != a->last_seg) {
      ARC(load_state(a, a->ckpt[cidx], &a->ckpt_meta[3 * cidx]));
      ARC(keep_step(a, 0));
      for (int j = 1; j <= len; ++j) {
        ARC(fstep(a));
        ARC(keep_step(a, j));
      }
    }
    a->last_seg = -1;
    for (int j = len - 1; j >= 0; --j) {
      const double2* fn[4] = {seg_field(a, j, 0), seg_field(a, j, 1), seg_field(a, j, 2), seg_field(a, j, 3)};
      if (a->fused)
        ARC(adj_step_fused(a, fn, seg_field(a, j + 1, 2), seg_field(a, j + 1, 3), n0 + j == 0));
      else
        ARC(adj_step(a, fn, seg_field(a, j + 1, 2), seg_field(a, j + 1, 3), n0 + j == 0));
    }
  }
  return QG_OK;
}

// adjoint of the control map psi(x) -> (psi-hat, q = init_q(psi-hat)):
// psibar(x) = N^-1 c2r(L (psibar-hat + T qbar / dz^2) / W)
int control_adjoint(qg_adj* a, double* grad_host, int ldx) {
  Ops o{a};
  Dev dd = a->c->d;
  AHIP(a, qg::launch_init_q(dd, a->st, a->x[F_Q], a->s[9]));
  ARC(o.c2r_of(a->p[12], {wpow(term(a->x[F_PSI], 1.0 / o.N()), -1), wpow(term(a->s[9], 1.0 / o.N()), -1)}));
  const auto& g = a->w.g;
  if (ldx == g.n1)
    AHIP(a, hipMemcpyAsync(grad_host, a->p[12], a->NP * sizeof(double), hipMemcpyDeviceToHost, a->st));
  else
    AHIP(a, hipMemcpy2DAsync(grad_host, (size_t)ldx * sizeof(double), a->p[12], (size_t)g.n1 * sizeof(double),
                             (size_t)g.n1 * sizeof(double), (size_t)g.n2 * g.nz, hipMemcpyDeviceToHost, a->st));
  AHIP(a, hipStreamSynchronize(a->st));
  return QG_OK;
}

int check(qg_adj* a, const double* psir, const double* BRr, const double* BIr, int ldx, int nsteps) {
  if (!a) return QG_ERR_ARG;
  a->err.clear();
  if (!psir || !BRr || !BIr || ldx < a->w.g.n1 || nsteps < 1) return afail(a, QG_ERR_ARG, "bad arguments");
  AHIP(a, hipSetDevice(a->c->device));
  return QG_OK;
}

int h2d_state(qg_adj* a, const double* host) {
  const auto& g = a->w.g;
  for (int f = 0; f < NSTATE; ++f)
    AHIP(a, qgi::h2d_field(g, a->x[f], host + (size_t)f * g.iktx * g.ikty * g.nz * 2, g.nz, a->st, a->c->stage));
  // keep only L = 1 modes
  Ops o{a};
  for (int f = 0; f < NSTATE; ++f) ARC(o.smul(a->x[f], {term(a->x[f])}));
  return QG_OK;
}

int d2h_state(qg_adj* a, double* host) {
  const auto& g = a->w.g;
  for (int f = 0; f < NSTATE; ++f)
    AHIP(a, qgi::d2h_field(g, host + (size_t)f * g.iktx * g.ikty * g.nz * 2, a->x[f], g.nz, a->st, a->c->stage));
  return QG_OK;
}

}  // namespace

extern "C" {

int qg_adj_create(qg_ctx* c, int32_t checkpoint_every, double z_deep, qg_adj** out) {
  if (!c || !out || checkpoint_every < 1) return QG_ERR_ARG;
  *out = nullptr;
  if (c->d.nuz != 0.0 || (c->d.flags & ~qg::F_NO_FEEDBACK))
    return qgi::fail(c, QG_ERR_ARG, "qg_adj_create: needs nuz = 0 and flags 0 or QG_NO_FEEDBACK");
  int rc = qgi::check_ctx(c);
  if (rc) return rc;
  qg_adj* a = new qg_adj();
  a->c = c;
  a->st = c->stream;
  a->K = checkpoint_every;
  a->zdeep = z_deep;
  a->w = c->d;
  const auto& g = a->w.g;
  a->S = qgi::field_elems(g);
  a->NP = (size_t)g.n1 * g.n2 * g.nz;
  // deep levels: z_s(k) = (k + 1/2) dz < z_deep
  a->zhi = 0;
  while (a->zhi < g.n3 && (a->zhi + 0.5) * c->d.dz < z_deep) ++a->zhi;
  rc = QG_OK;
  for (int i = 0; i < NSTATE && rc == QG_OK; ++i) rc = aalloc(a, &a->x[i], a->S);
  for (int i = 0; i < 12 && rc == QG_OK; ++i) rc = aalloc(a, &a->s[i], a->S);
  rc = aalloc(a, &a->p[12], a->NP);  // control transfers; the rest on first use
  // forward planes in the order F0, F2, F4, F1, F3, F5 and then F1[2],
  // F1[4], F1[3], F1[5], so that the y-inverses of (q, BR, BI) and of
  // (BR', BI') are jobs over stacked levels
  double2* fblk = nullptr;
  if (rc == QG_OK) rc = aalloc(a, &fblk, 10 * a->S);
  if (rc == QG_OK) {
    const int order[6] = {0, 3, 1, 4, 2, 5};
    for (int i = 0; i < 6; ++i) a->F[i] = fblk + (size_t)order[i] * a->S;
    const int order1[4] = {6, 8, 7, 9};
    for (int i = 0; i < 4; ++i) a->F1[2 + i] = fblk + (size_t)order1[i] * a->S;
  }
  for (int i = 0; i < 6 && rc == QG_OK; ++i) rc = aalloc(a, &a->A[i], a->S);
  double2* tq = nullptr;
  if (rc == QG_OK) rc = aalloc(a, &tq, a->S);
  a->w.qs = tq;  // transform temporary of the raw r2c / c2r
  if (rc == QG_OK) {
    qg::Geo gm = g;
    gm.nm = 1;
    gm.nz = g.n3;
    rc = aalloc(a, &a->part, (size_t)qg::obj_blocks(gm) * g.nm);
  }
  if (rc == QG_OK) rc = aalloc(a, &a->seg, (size_t)(a->K + 1) * 4 * a->S);
  if (rc == QG_OK) rc = aalloc(a, &a->fin, 12 * a->S);
  if (rc == QG_OK && hipStreamSynchronize(a->st) != hipSuccess) rc = QG_ERR_HIP;
  if (rc != QG_OK) {
    c->err = a->err;
    qg_adj_destroy(a);
    return rc;
  }
  *out = a;
  return QG_OK;
}

int qg_adj_destroy(qg_adj* a) {
  if (!a) return QG_OK;
  hipSetDevice(a->c->device);
  hipStreamSynchronize(a->st);
  for (void* v : a->allocs) hipFree(v);
  delete a;
  return QG_OK;
}

const char* qg_adj_error(const qg_adj* a) { return a ? a->err.c_str() : "null adjoint context"; }

int qg_adj_set_fused(qg_adj* a, int32_t fused) {
  if (!a || fused < 0 || fused > 1) return QG_ERR_ARG;
  a->fused = fused;
  return QG_OK;
}

int qg_adj_objective(qg_adj* a, const double* psir, const double* BRr, const double* BIr, int32_t ldx,
                     int32_t nsteps, double* J) {
  ARC(check(a, psir, BRr, BIr, ldx, nsteps));
  if (!J) return afail(a, QG_ERR_ARG, "null J");
  ARC(forward(a, psir, BRr, BIr, ldx, nsteps, false));
  return objective(a, a->c->d.AR, a->c->d.AI, J);
}

int qg_adj_tlm(qg_adj* a, const double* psir, const double* BRr, const double* BIr, int32_t ldx, int32_t nsteps,
               const double* dpsir, double* dJ, double* dstate) {
  ARC(check(a, psir, BRr, BIr, ldx, nsteps));
  if (!dpsir) return afail(a, QG_ERR_ARG, "null perturbation");
  ARC(ensure_generic(a));
  qg_ctx* c = a->c;
  Ops o{a};
  int rc = qg_set_state_physical(c, psir, BRr, BIr, ldx);
  if (rc != QG_OK) return afail(a, rc, "initial state: " + c->err);
  AHIP(a, hipMemsetAsync(c->d.flag, 0, sizeof(int), a->st));
  // perturbation of the initial state: psi-hat = r2c(dpsi), q = init_q(psi-hat)
  const auto& g = a->w.g;
  for (int f = 0; f < NSTATE; ++f) ARC(o.zero(a->x[f]));
  if (ldx == g.n1)
    AHIP(a, hipMemcpyAsync(a->p[12], dpsir, a->NP * sizeof(double), hipMemcpyHostToDevice, a->st));
  else
    AHIP(a, hipMemcpy2DAsync(a->p[12], (size_t)g.n1 * sizeof(double), dpsir, (size_t)ldx * sizeof(double),
                             (size_t)g.n1 * sizeof(double), (size_t)g.n2 * g.nz, hipMemcpyHostToDevice, a->st));
  ARC(o.r2c(a->p[12], a->s[9]));
  // only the L = 1 modes of psi-hat act in the step (every use is masked);
  // the inversions write the box only, so the rest must stay zero
  ARC(o.smul(a->x[F_PSI], {term(a->s[9])}));
  AHIP(a, qg::launch_init_q(c->d, a->st, a->x[F_PSI], a->x[F_Q]));
  for (int n = 0; n < nsteps; ++n) {
    ARC(keep_step(a, 0));
    ARC(fstep(a));
    AHIP(a, hipStreamWaitEvent(c->stream, c->evA, 0));
    const double2* fn[4] = {seg_field(a, 0, 0), seg_field(a, 0, 1), seg_field(a, 0, 2), seg_field(a, 0, 3)};
    ARC(tlm_step(a, fn, c->d.BR, c->d.BI, n == 0));
  }
  // dJ = (2 pi)^2 dz sum W kh2 Re(conj(AR) dAR + conj(AI) dAI) over the deep levels
  if (dJ) {
    ARC(o.smul(a->s[9], {term(c->d.AR), term(a->x[F_AR])}));
    ARC(o.smul(a->s[10], {term(c->d.AR), term(a->x[F_AR], -1.0)}));
    ARC(o.smul(a->s[11], {term(c->d.AI), term(a->x[F_AI])}));
    ARC(o.smul(a->s[8], {term(c->d.AI), term(a->x[F_AI], -1.0)}));
    // |a+b|^2 - |a-b|^2 = 4 Re(conj(a) b)
    const int nm = a->w.g.nm;
    std::vector<double> j1(nm), j2(nm);
    ARC(objective(a, a->s[9], a->s[11], j1.data()));
    ARC(objective(a, a->s[10], a->s[8], j2.data()));
    for (int m = 0; m < nm; ++m) dJ[m] = 2.0 * (j1[m] - j2[m]) / 4.0;
  }
  if (dstate) ARC(d2h_state(a, dstate));
  AHIP(a, hipStreamSynchronize(a->st));
  return QG_OK;
}

int qg_adj_apply(qg_adj* a, const double* psir, const double* BRr, const double* BIr, int32_t ldx, int32_t nsteps,
                 const double* state_bar, double* psi_bar) {
  ARC(check(a, psir, BRr, BIr, ldx, nsteps));
  if (!state_bar || !psi_bar) return afail(a, QG_ERR_ARG, "null array");
  ARC(forward(a, psir, BRr, BIr, ldx, nsteps, true));
  ARC(save_state(a, a->fin, a->fin_meta));
  ARC(h2d_state(a, state_bar));
  ARC(sweep(a, nsteps));
  ARC(control_adjoint(a, psi_bar, ldx));
  ARC(load_state(a, a->fin, a->fin_meta));
  AHIP(a, hipStreamSynchronize(a->st));
  return QG_OK;
}

int qg_adj_gradient(qg_adj* a, const double* psir, const double* BRr, const double* BIr, int32_t ldx,
                    int32_t nsteps, double* J, double* grad) {
  ARC(check(a, psir, BRr, BIr, ldx, nsteps));
  if (!grad) return afail(a, QG_ERR_ARG, "null gradient");
  ARC(forward(a, psir, BRr, BIr, ldx, nsteps, true));
  std::vector<double> Jv(a->w.g.nm);
  ARC(objective(a, a->c->d.AR, a->c->d.AI, Jv.data()));
  if (J)
    for (int m = 0; m < a->w.g.nm; ++m) J[m] = Jv[m];
  ARC(save_state(a, a->fin, a->fin_meta));
  // dJ/dA at the final step: (2 pi)^2 dz W kh2 A on the deep levels
  Ops o{a};
  const double twopi = 4.0 * std::asin(1.0);
  const double f = twopi * twopi * a->c->d.dz;
  for (int i = 0; i < NSTATE; ++i) ARC(o.zero(a->x[i]));
  ARC(o.smul(a->x[F_AR], {wpow(t_kh2(a->c->d.AR, f), 1)}, 0, 0, a->zhi));
  ARC(o.smul(a->x[F_AI], {wpow(t_kh2(a->c->d.AI, f), 1)}, 0, 0, a->zhi));
  ARC(sweep(a, nsteps));
  ARC(control_adjoint(a, grad, ldx));
  ARC(load_state(a, a->fin, a->fin_meta));
  AHIP(a, hipStreamSynchronize(a->st));
  return QG_OK;
}

}  // extern "C"
