// Fused adjoint step: the transpose of the step's transform kernels, built
// like them from register-resident FFTs (qg_regfft.hpp) with the pointwise
// adjoint work between the transforms.  Order of one adjoint step (the
// forward order reversed), with the vertical solves of qg_kernels.hip:
//
//   solves      : qbar* = S_psi(psibar'), sR = S_A(ARbar'), sI = S_A(AIbar')
//   adj_col_qw  : qbar*            -> A0 = Fy^-1[L qbar*/W], A1 = Fy^-1[L ky^2 qbar*/W]
//   adj_row_qw  : F2..F5 (forward (kx,y) rows of BR', i ky BR', BI', i ky BI')
//                 and A0, A1       -> feedback adjoint rows A2..A5 (DFT_x,
//                                     pre-combined with -i kx)
//   adj_col_upd : A2..A5 + state   -> BR+bar, BI+bar, q+bar; the time-update
//                 adjoint in place; A0..A5 = Fy^-1[L T/W], Fy^-1[L (-i ky) T/W]
//                 for T = e1 q+bar, e1w BR+bar, e1w BI+bar
//   adj_row_nl  : F0..F5 (forward rows of q, psi, BR, psi_y, BI, psi_yy) and
//                 A0..A5           -> advection / refraction adjoint rows
//   adj_col_out : A0..A5           -> psibar, qbar, BRbar, BIbar
//
// With N = n1 n2 and W the Hermitian weight of a half-spectrum mode (2, and
// 1 at kx = ky = 0), the adjoint of c2r restricted to L is W L DFT (no 1/N)
// and the adjoint of r2c restricted to L is N^-1 c2r(L . / W); the
// derivation is in qg_adj.cpp, whose generic adjoint step computes the same
// quantities and is used to test these kernels.
#include "qg_adj.hpp"
#include "qg_common.hpp"
#include "qg_regfft.hpp"
#include "qg_tx.hpp"

namespace qg {
namespace {

__device__ __forceinline__ double hweight(int kx, int ky) { return (kx == 0 && ky == 0) ? 1.0 : 2.0; }

// realit fill of the kx = 0 column for sequences S0 .. S0+G-1 of v (after
// their multipliers): entries iky > N/2 become conj of entry N - iky
template <int N, int E, int S, int S0, int G>
__device__ __forceinline__ void realit_col0(double2 (&v)[S][E], const ColBuf<N, E>& xb, int t, int c) {
  constexpr int T = N / E;
  rf_publish_range<N, E, S, S0, G>(v, xb, t);
  if (c == 0) {
#pragma unroll
    for (int s = 0; s < G; ++s)
#pragma unroll
      for (int m = 0; m < E; ++m) {
        const int iky = t + T * m;
        if (iky > N / 2) v[S0 + s][m] = cconj(xb.pub(s, N - iky));
      }
  }
}
// the same for all S sequences through SB exchange buffers, SB at a time
template <int N, int E, int S, int SB, int S0 = 0>
__device__ __forceinline__ void realit_cols(double2 (&v)[S][E], const ColBuf<N, E>& xb, int t, int c) {
  constexpr int G = (S - S0) < SB ? (S - S0) : SB;
  realit_col0<N, E, S, S0, G>(v, xb, t, c);
  if constexpr (S0 + G < S) realit_cols<N, E, S, SB, S0 + G>(v, xb, t, c);
}

// ---------------------------------------------------------------------------
// adj_col_qw: y-inverses of the feedback adjoint seeds
// ---------------------------------------------------------------------------
template <int N, int E, int SB>
__global__ __launch_bounds__((SH<N, E>::COLT)) void k_adj_col_qw(Dev d, AdjBufs b) {
  constexpr int T = SH<N, E>::T, TX = SH<N, E>::TX;
  __shared__ double2 lds[col_xlen<N, E, SB, TX>()];
  const Geo& g = d.g;
  const int c = threadIdx.x % TX, t = threadIdx.x / TX;
  const TwSrc<N, E, (E <= 8)> tws(d.tw, t);  // before the data loads
  const int kx0 = blockIdx.x * TX, kx = kx0 + c, z = blockIdx.y;
  const ColBuf<N, E> xb{lds + c};
  const bool colok = kx < g.KX;
  const size_t sz = (size_t)z * g.ikty * g.sp, pz = (size_t)z * g.n2 * g.sp;
  // loads first, unconditionally (column 0 stands in for columns beyond KX),
  // then the mask: a zero-initialised branch around each load would make
  // the compiler wait on it
  double2 v[2][E];
#pragma unroll
  for (int m = 0; m < E; ++m) v[0][m] = b.qstar[sz + (size_t)(t + T * m) * g.sp + (colok ? kx : 0)];
#pragma unroll
  for (int m = 0; m < E; ++m) {
    const int iky = t + T * m;
    const int ky = ky_of(iky, g.n2);
    const bool on = colok && lmask(kx, iky, g);
    const double2 a = on ? cscale(v[0][m], 1.0 / hweight(kx, ky)) : zero2();
    v[0][m] = a;
    v[1][m] = cscale(a, (double)ky * (double)ky);
  }
  if (kx0 == 0) realit_cols<N, E, 2, SB>(v, xb, t, c);
  reg_fft_multi<N, E, 2, SB, true>(v, xb, t, tws.get());
  if (colok) {
#pragma unroll
    for (int m = 0; m < E; ++m) {
      const size_t o = pz + (size_t)(t + T * m) * g.sp + kx;
      st_nt(&b.A[0][o], v[0][m]);
      st_nt(&b.A[1][o], v[1][m]);
    }
  }
}

// ---------------------------------------------------------------------------
// adj_row_qw: per (z, y) row, c2r in x of BR', BI', their x and y
// derivatives and of the adjoint seeds Jbar = cJ c2r(A0), Ebar =
// cE c2r(kx^2 A0 + A1); the transposed products of J = BRy BIx - BRx BIy and
// |B|^2; DFT in x:
//   A2 = -i kx X(BRx bar) + X(BR bar), A3 = X(BRy bar),
//   A4 = -i kx X(BIx bar) + X(BI bar), A5 = X(BIy bar)
// ---------------------------------------------------------------------------
template <int N, int E, int RP, int SB>
__global__ __launch_bounds__(RP*(N / E)) void k_adj_row_qw(Dev d, AdjBufs b, double cJ, double cE) {
  constexpr int T = N / E, RS = row_xstride<N, E, SB>();
  constexpr int MK0 = ((N / 3 + 2) + T - 1) / T;
  constexpr int MK = MK0 < E ? MK0 : E;
  __shared__ double2 lds[RP * RS];
  const Geo& g = d.g;
  int t, p;
  row_lane<T, RP>(threadIdx.x, t, p);
  const TwSrc<N, E, (E <= 8)> tws(d.tw, t);  // before the data loads
  const RowBuf<N, E> xb{lds + p * RS};
  const size_t ro = ((size_t)blockIdx.x * RP + p) * g.sp;
  const int KX = g.KX;
  // loads first, unconditionally (entry 0 stands in for k >= KX), then the
  // mask
  double2 in[6][E];
#pragma unroll
  for (int m = 0; m < E; ++m) {
    const int pos = t + T * m;
    const int k = pos <= N / 2 ? pos : N - pos;
    const size_t a = ro + (k < KX ? k : 0);
    in[0][m] = b.F[2][a]; in[1][m] = b.F[3][a]; in[2][m] = b.F[4][a];
    in[3][m] = b.F[5][a]; in[4][m] = b.A[0][a]; in[5][m] = b.A[1][a];
  }
  double2 z[4][E];
#pragma unroll
  for (int m = 0; m < E; ++m) {
    const int pos = t + T * m;
    const int k = pos <= N / 2 ? pos : N - pos;
    const bool ok = k < KX;
    const double2 br = ok ? in[0][m] : zero2(), bry = ok ? in[1][m] : zero2(), bi = ok ? in[2][m] : zero2();
    const double2 biy = ok ? in[3][m] : zero2(), aj = ok ? in[4][m] : zero2(), ajy = ok ? in[5][m] : zero2();
    const double kk = (double)k;
    z[0][m] = pack2<N>(br, bi, pos);
    z[1][m] = pack2<N>(mul_ik(br, kk), mul_ik(bi, kk), pos);
    z[2][m] = pack2<N>(bry, biy, pos);
    const double2 je = make_double2(kk * kk * aj.x + ajy.x, kk * kk * aj.y + ajy.y);
    z[3][m] = pack2<N>(cscale(aj, cJ), cscale(je, cE), pos);
  }
  reg_fft_multi<N, E, 4, SB, true>(z, xb, t, tws.get());
  // z0 = (BR, BI), z1 = (BRx, BIx), z2 = (BRy, BIy), z3 = (Jbar, Ebar)
  double2 pv[3][E];
#pragma unroll
  for (int m = 0; m < E; ++m) {
    const double br = z[0][m].x, bi = z[0][m].y, brx = z[1][m].x, bix = z[1][m].y;
    const double bry = z[2][m].x, biy = z[2][m].y, jb = z[3][m].x, eb = z[3][m].y;
    pv[0][m] = make_double2(-jb * biy, jb * bry);           // (BRx bar, BIx bar)
    pv[1][m] = make_double2(jb * bix, -jb * brx);           // (BRy bar, BIy bar)
    pv[2][m] = make_double2(2.0 * eb * br, 2.0 * eb * bi);  // (BR bar, BI bar)
  }
  reg_fft_multi<N, E, 3, SB, false>(pv, xb, t, tws.get());
  double2 kx_r[MK], kx_i[MK];
  rf_publish_range<N, E, 3, 0, 2>(pv, xb, t);
#pragma unroll
  for (int m = 0; m < MK; ++m) {
    const int k = t + T * m;
    kx_r[m] = zero2();
    kx_i[m] = zero2();
    if (k < KX) {
      const int km = (N - k) & (N - 1);
      double2 Xrx, Xix, Xry, Xiy;
      split2(pv[0][m], xb.pub(0, km), Xrx, Xix);
      split2(pv[1][m], xb.pub(1, km), Xry, Xiy);
      const double kk = (double)k;
      kx_r[m] = mul_ik(Xrx, -kk);
      kx_i[m] = mul_ik(Xix, -kk);
      st_nt(&b.A[3][ro + k], Xry);
      st_nt(&b.A[5][ro + k], Xiy);
    }
  }
  rf_publish_range<N, E, 3, 2, 1>(pv, xb, t);
#pragma unroll
  for (int m = 0; m < MK; ++m) {
    const int k = t + T * m;
    if (k < KX) {
      double2 Xr, Xi;
      split2(pv[2][m], xb.pub(0, (N - k) & (N - 1)), Xr, Xi);
      st_nt(&b.A[2][ro + k], cadd(kx_r[m], Xr));
      st_nt(&b.A[4][ro + k], cadd(kx_i[m], Xi));
    }
  }
}

// ---------------------------------------------------------------------------
// adj_col_upd: finish BR+bar / BI+bar (DFT_y of the feedback rows, plus the
// A-inversion adjoints and the adjoints of the new BR, BI) and q+bar, apply
// the adjoint of the time update (Euler or leapfrog + Robert-Asselin) to
// the state adjoint in place, and form the y-inverses of the tendency seeds
// T/W and (-i ky) T/W.  blockIdx.z: 0 = q, 1 = BR, 2 = BI.
// ---------------------------------------------------------------------------
template <int N, int E, int SB>
__global__ __launch_bounds__((SH<N, E>::COLT)) void k_adj_col_upd(Dev d, AdjBufs b, int euler, int feedback) {
  constexpr int T = SH<N, E>::T, TX = SH<N, E>::TX;
  __shared__ double2 lds[col_xlen<N, E, SB, TX>()];
  const Geo& g = d.g;
  const int c = threadIdx.x % TX, t = threadIdx.x / TX;
  const TwSrc<N, E, (E <= 8)> tws(d.tw, t);  // before the data loads
  const int kx0 = blockIdx.x * TX, kx = kx0 + c, z = blockIdx.y, f = blockIdx.z;
  const ColBuf<N, E> xb{lds + c};
  const bool colok = kx < g.KX;
  const size_t sz = (size_t)z * g.ikty * g.sp, pz = (size_t)z * g.n2 * g.sp;
  const int toff = t * g.sp + kx;
  const size_t mstride = (size_t)T * g.sp;
  double2* Y0 = b.A[2 * f];
  double2* Y1 = b.A[2 * f + 1];
  double2 v[2][E];
  const bool fb = feedback && f > 0;
  if (fb) {
#pragma unroll
    for (int m = 0; m < E; ++m) {
      v[0][m] = colok ? (Y0 + pz + m * mstride)[toff] : zero2();
      v[1][m] = colok ? (Y1 + pz + m * mstride)[toff] : zero2();
    }
    reg_fft_multi<N, E, 2, SB, false>(v, xb, t, tws.get());
  }
  double2* X = f == 0 ? b.q : (f == 1 ? b.BR : b.BI);
  double2* Xo = f == 0 ? b.qo : (f == 1 ? b.BRo : b.BIo);
  double2* Aout = f == 1 ? b.AI : b.AR;
  const double2* sol = f == 0 ? b.qstar : (f == 1 ? b.sR : b.sI);
  const double* e1t = f == 0 ? d.e1q : d.e1w;
  const double* e2t = f == 0 ? d.e2q : d.e2w;
  const double cd = 0.5 / (d.Bu * d.Ro);
  const double gm = d.gamma;
  // dA = -s c kh2 e1w BR+bar for BR, +s c kh2 e1w BI+bar for BI (s = 2 dt, or dt
  // in the Euler step)
  const double sdt = euler ? d.dt : 2.0 * d.dt;
  const double sgn = f == 1 ? -1.0 : 1.0;
#pragma unroll
  for (int m = 0; m < E; ++m) {
    const int iky = t + T * m;
    double2 o0 = zero2(), o1 = zero2();
    if (colok && lmask(kx, iky, g)) {
      const int ky = ky_of(iky, g.n2);
      const double kyd = (double)ky, w = hweight(kx, ky);
      const size_t idx = sz + (size_t)iky * g.sp + kx;
      const size_t kidx = (size_t)iky * g.iktx + kx;
      double2 xp = cadd(X[idx], sol[idx]);  // adjoint of the new field, plus the inversion's
      if (fb) {
        const double2 F1 = v[0][m], F2 = v[1][m];
        // W (F1 + (-i ky) F2)
        xp.x += w * (F1.x + kyd * F2.y);
        xp.y += w * (F1.y - kyd * F2.x);
      }
      const double2 xo = Xo[idx];
      const double e1 = e1t[kidx];
      double2 xn, xon;
      if (euler) {
        xn = make_double2(xo.x + e1 * xp.x, xo.y + e1 * xp.y);
        xon = zero2();
      } else {
        xp = make_double2(xp.x + gm * xo.x, xp.y + gm * xo.y);
        const double e2 = e2t[kidx];
        xn = make_double2((1.0 - 2.0 * gm) * xo.x, (1.0 - 2.0 * gm) * xo.y);
        xon = make_double2(gm * xo.x + e2 * xp.x, gm * xo.y + e2 * xp.y);
      }
      X[idx] = xn;
      Xo[idx] = xon;
      const double2 tv = make_double2(e1 * xp.x, e1 * xp.y);
      if (f > 0) {
        const double kh2 = (double)kx * (double)kx + kyd * kyd;
        const double ca = sgn * (sdt * cd) * kh2;
        Aout[idx] = make_double2(ca * tv.x, ca * tv.y);
      }
      o0 = cscale(tv, 1.0 / w);
      o1 = mul_ik(o0, -kyd);
    } else if (colok && in_box_row(iky, g)) {
      const size_t idx = sz + (size_t)iky * g.sp + kx;
      X[idx] = zero2();
      Xo[idx] = zero2();
      if (f > 0) Aout[idx] = zero2();
    }
    v[0][m] = o0;
    v[1][m] = o1;
  }
  if (kx0 == 0) realit_cols<N, E, 2, SB>(v, xb, t, c);
  reg_fft_multi<N, E, 2, SB, true>(v, xb, t, tws.get());
  if (colok) {
#pragma unroll
    for (int m = 0; m < E; ++m) {
      st_nt(&(Y0 + pz + m * mstride)[toff], v[0][m]);
      st_nt(&(Y1 + pz + m * mstride)[toff], v[1][m]);
    }
  }
}

// ---------------------------------------------------------------------------
// adj_row_nl: per (z, y) row, c2r in x of the forward u, v, zeta, q, BR, BI
// (as row_nl forms them) and of the advection / refraction adjoint seeds
//   a_x(X) = s1 c2r(-i kx A_X), a_y(X) = s1 c2r(A_Xy)  (X = q, BR, BI)
//   b_R = -s2 c2r(A_BI), b_I = s2 c2r(A_BR)
// the transposed products, DFT in x and the pre-combination of the psi
// adjoint's multipliers:
//   A0 = X(ubar), A1 = -i kx X(vbar) - kx^2 X(zetabar), A2 = X(zetabar),
//   A3 = X(qbar), A4 = X(BRbar), A5 = X(BIbar)
// ---------------------------------------------------------------------------
template <int N, int E, int RP, int SB>
__global__ __launch_bounds__(RP*(N / E)) void k_adj_row_nl(Dev d, AdjBufs b, int Win, double s1, double s2) {
  constexpr int T = N / E, RS = row_xstride<N, E, SB>();
  constexpr int MK0 = ((N / 3 + 2) + T - 1) / T;
  constexpr int MK = MK0 < E ? MK0 : E;
  __shared__ double2 lds[RP * RS];
  const Geo& g = d.g;
  int t, p;
  row_lane<T, RP>(threadIdx.x, t, p);
  const TwSrc<N, E, (E <= 8)> tws(d.tw, t);  // before the data loads
  const RowBuf<N, E> xb{lds + p * RS};
  const size_t ro = ((size_t)blockIdx.x * RP + p) * g.sp;
  const int KX = g.KX;
  double2 z[7][E];
#pragma unroll
  for (int m = 0; m < E; ++m) {
    const int pos = t + T * m;
    const int k = pos <= N / 2 ? pos : N - pos;
    const double kk = (double)k, k2 = kk * kk;
    // unconditional loads (entry 0 stands in for masked positions), then the
    // mask: a zero-initialised branch around each load would make the
    // compiler wait on it
    const size_t af = ro + (k < Win ? k : 0), aa = ro + (k < KX ? k : 0);
    const double2 lps = b.F[1][af], lpy = b.F[3][af], lpyy = b.F[5][af];
    const double2 lpq = b.F[0][af], lpbr = b.F[2][af], lpbi = b.F[4][af];
    const double2 laq = b.A[0][aa], laqy = b.A[1][aa], lar = b.A[2][aa];
    const double2 lary = b.A[3][aa], lai = b.A[4][aa], laiy = b.A[5][aa];
    const bool okf = k < Win, oka = k < KX;
    const double2 ps = okf ? lps : zero2(), py = okf ? lpy : zero2(), pyy = okf ? lpyy : zero2();
    const double2 pq = okf ? lpq : zero2(), pbr = okf ? lpbr : zero2(), pbi = okf ? lpbi : zero2();
    const double2 U = make_double2(-py.x, -py.y);
    const double2 V = mul_ik(ps, kk);
    const double2 Z = make_double2(-k2 * ps.x + pyy.x, -k2 * ps.y + pyy.y);
    z[0][m] = pack2<N>(U, V, pos);
    z[1][m] = pack2<N>(Z, pq, pos);
    z[2][m] = pack2<N>(pbr, pbi, pos);
    const double2 aq = oka ? laq : zero2(), aqy = oka ? laqy : zero2(), ar = oka ? lar : zero2();
    const double2 ary = oka ? lary : zero2(), ai = oka ? lai : zero2(), aiy = oka ? laiy : zero2();
    z[3][m] = pack2<N>(mul_ik(aq, -s1 * kk), cscale(aqy, s1), pos);
    z[4][m] = pack2<N>(mul_ik(ar, -s1 * kk), cscale(ary, s1), pos);
    z[5][m] = pack2<N>(mul_ik(ai, -s1 * kk), cscale(aiy, s1), pos);
    z[6][m] = pack2<N>(cscale(ai, -s2), cscale(ar, s2), pos);
  }
  reg_fft_multi<N, E, 7, SB, true>(z, xb, t, tws.get());
  double2 pv[3][E];
#pragma unroll
  for (int m = 0; m < E; ++m) {
    const double u = z[0][m].x, v = z[0][m].y, ze = z[1][m].x, q = z[1][m].y, br = z[2][m].x, bi = z[2][m].y;
    const double axq = z[3][m].x, ayq = z[3][m].y, axr = z[4][m].x, ayr = z[4][m].y;
    const double axi = z[5][m].x, ayi = z[5][m].y, bR = z[6][m].x, bI = z[6][m].y;
    const double ub = (axq * q + axr * br) + axi * bi;
    const double vb = (ayq * q + ayr * br) + ayi * bi;
    const double zb = bR * br + bI * bi;
    const double qb = axq * u + ayq * v;
    const double rb = (axr * u + ayr * v) + bR * ze;
    const double ib = (axi * u + ayi * v) + bI * ze;
    pv[0][m] = make_double2(ub, vb);
    pv[1][m] = make_double2(zb, qb);
    pv[2][m] = make_double2(rb, ib);
  }
  reg_fft_multi<N, E, 3, SB, false>(pv, xb, t, tws.get());
  rf_publish_range<N, E, 3, 0, 2>(pv, xb, t);
#pragma unroll
  for (int m = 0; m < MK; ++m) {
    const int k = t + T * m;
    if (k < KX) {
      const int km = (N - k) & (N - 1);
      double2 Xu, Xv, Xz, Xq;
      split2(pv[0][m], xb.pub(0, km), Xu, Xv);
      split2(pv[1][m], xb.pub(1, km), Xz, Xq);
      const double kk = (double)k;
      const double2 a1 = mul_ik(Xv, -kk);
      st_nt(&b.A[0][ro + k], Xu);
      st_nt(&b.A[1][ro + k], make_double2(a1.x - kk * kk * Xz.x, a1.y - kk * kk * Xz.y));
      st_nt(&b.A[2][ro + k], Xz);
      st_nt(&b.A[3][ro + k], Xq);
    }
  }
  rf_publish_range<N, E, 3, 2, 1>(pv, xb, t);
#pragma unroll
  for (int m = 0; m < MK; ++m) {
    const int k = t + T * m;
    if (k < KX) {
      double2 Xr, Xi;
      split2(pv[2][m], xb.pub(0, (N - k) & (N - 1)), Xr, Xi);
      st_nt(&b.A[4][ro + k], Xr);
      st_nt(&b.A[5][ro + k], Xi);
    }
  }
}

// ---------------------------------------------------------------------------
// adj_col_out: DFT in y of the pre-combined rows and the psi / q / BR / BI
// adjoints: psibar = W L (i ky Y(A0) + Y(A1) - ky^2 Y(A2)) (written),
// qbar += W L Y(A3), BRbar += W L Y(A4), BIbar += W L Y(A5).
// blockIdx.z: 0 = psi, 1 = q, BR, BI.
// ---------------------------------------------------------------------------
template <int N, int E, int SB>
__global__ __launch_bounds__((SH<N, E>::COLT)) void k_adj_col_out(Dev d, AdjBufs b) {
  constexpr int T = SH<N, E>::T, TX = SH<N, E>::TX;
  __shared__ double2 lds[col_xlen<N, E, SB, TX>()];
  const Geo& g = d.g;
  const int c = threadIdx.x % TX, t = threadIdx.x / TX;
  const TwSrc<N, E, (E <= 8)> tws(d.tw, t);  // before the data loads
  const int kx0 = blockIdx.x * TX, kx = kx0 + c, z = blockIdx.y, f = blockIdx.z;
  const ColBuf<N, E> xb{lds + c};
  const bool colok = kx < g.KX;
  const size_t sz = (size_t)z * g.ikty * g.sp, pz = (size_t)z * g.n2 * g.sp;
  const int toff = t * g.sp + kx;
  const size_t mstride = (size_t)T * g.sp;
  double2 v[3][E];
#pragma unroll
  for (int s = 0; s < 3; ++s)
#pragma unroll
    for (int m = 0; m < E; ++m) v[s][m] = colok ? (b.A[3 * f + s] + pz + m * mstride)[toff] : zero2();
  reg_fft_multi<N, E, 3, SB, false>(v, xb, t, tws.get());
  if (!colok) return;
#pragma unroll
  for (int m = 0; m < E; ++m) {
    const int iky = t + T * m;
    if (!in_box_row(iky, g)) continue;
    const size_t idx = sz + (size_t)iky * g.sp + kx;
    const bool on = lmask(kx, iky, g);
    const int ky = ky_of(iky, g.n2);
    const double w = on ? hweight(kx, ky) : 0.0;
    if (f == 0) {
      const double kyd = (double)ky;
      const double2 a = mul_ik(v[0][m], kyd);
      b.psi[idx] = make_double2(w * ((a.x + v[1][m].x) - kyd * kyd * v[2][m].x),
                                w * ((a.y + v[1][m].y) - kyd * kyd * v[2][m].y));
    } else {
      b.q[idx] = make_double2(b.q[idx].x + w * v[0][m].x, b.q[idx].y + w * v[0][m].y);
      b.BR[idx] = make_double2(b.BR[idx].x + w * v[1][m].x, b.BR[idx].y + w * v[1][m].y);
      b.BI[idx] = make_double2(b.BI[idx].x + w * v[2][m].x, b.BI[idx].y + w * v[2][m].y);
    }
  }
}

template <int N>
struct LA {
  static constexpr int E = 8;
  using S = SH<N, E>;
  static constexpr int RP = S::template rp<64>();
  static constexpr int CSB = 1;  // exchange buffers of the column kernels (LDS for 3 waves)
  // the row_nl kernel holds seven transforms; at E = 8 they take every
  // register (1 wave per SIMD), at E = 4 half as many (2 waves)
  static constexpr int ENL = N >= 128 ? 4 : 8, RPNL = SH<N, ENL>::template rp<64>();
  static hipError_t col_qw(const Dev& d, hipStream_t st, const AdjBufs& b) {
    hipLaunchKernelGGL((k_adj_col_qw<N, E, CSB>), dim3((d.g.KX + S::TX - 1) / S::TX, d.g.nz), dim3(S::COLT), 0, st, d, b);
    return hipGetLastError();
  }
  static hipError_t row_qw(const Dev& d, hipStream_t st, const AdjBufs& b) {
    const double Nn = (double)d.g.n1 * (double)d.g.n2;
    const double rw = d.Ro * d.W2F;
    hipLaunchKernelGGL((k_adj_row_qw<N, E, RP, 2>), dim3(d.g.n2 * d.g.nz / RP), dim3(RP * (N / E)), 0, st, d, b,
                       -rw / Nn, rw / (4.0 * Nn));
    return hipGetLastError();
  }
  static hipError_t col_upd(const Dev& d, hipStream_t st, const AdjBufs& b, int euler, int feedback) {
    hipLaunchKernelGGL((k_adj_col_upd<N, E, CSB>), dim3((d.g.KX + S::TX - 1) / S::TX, d.g.nz, 3), dim3(S::COLT), 0, st, d,
                       b, euler, feedback);
    return hipGetLastError();
  }
  static hipError_t row_nl(const Dev& d, hipStream_t st, const AdjBufs& b, int Win, int euler) {
    const double Nn = (double)d.g.n1 * (double)d.g.n2;
    const double sg = euler ? d.dt : 2.0 * d.dt;
    hipLaunchKernelGGL((k_adj_row_nl<N, ENL, RPNL, 2>), dim3(d.g.n2 * d.g.nz / RPNL), dim3(RPNL * (N / ENL)), 0, st, d,
                       b, Win, -sg / Nn, sg / (2.0 * Nn));
    return hipGetLastError();
  }
  static hipError_t col_out(const Dev& d, hipStream_t st, const AdjBufs& b) {
    hipLaunchKernelGGL((k_adj_col_out<N, E, CSB>), dim3((d.g.KX + S::TX - 1) / S::TX, d.g.nz, 2), dim3(S::COLT), 0, st, d,
                       b);
    return hipGetLastError();
  }
};

}  // namespace

#define QG_ADISPATCH(n, call)                \
  switch (n) {                               \
    case 16: return LA<16>::call;            \
    case 32: return LA<32>::call;            \
    case 64: return LA<64>::call;            \
    case 128: return LA<128>::call;          \
    case 256: return LA<256>::call;          \
    case 512: return LA<512>::call;          \
    default: return hipErrorInvalidValue;    \
  }

hipError_t launch_adj_col_qw(const Dev& d, hipStream_t st, const AdjBufs& b) { QG_ADISPATCH(d.g.n2, col_qw(d, st, b)) }
hipError_t launch_adj_row_qw(const Dev& d, hipStream_t st, const AdjBufs& b) { QG_ADISPATCH(d.g.n2, row_qw(d, st, b)) }
hipError_t launch_adj_col_upd(const Dev& d, hipStream_t st, const AdjBufs& b, int euler, int feedback) {
  QG_ADISPATCH(d.g.n2, col_upd(d, st, b, euler, feedback))
}
hipError_t launch_adj_row_nl(const Dev& d, hipStream_t st, const AdjBufs& b, int Win, int euler) {
  QG_ADISPATCH(d.g.n2, row_nl(d, st, b, Win, euler))
}
hipError_t launch_adj_col_out(const Dev& d, hipStream_t st, const AdjBufs& b) {
  QG_ADISPATCH(d.g.n2, col_out(d, st, b))
}

}  // namespace qg
