// Transform-method kernels of the QG-YBJ+ time step for gfx950 (FP64).
//
// Data flow of one leapfrog step (spectral fields (iktx, ikty, n3) with row
// pitch sp; intermediates I0..I5 in the half-transformed (kx, y) layout):
//
//   ps_inv   : psi^n            -> I1 = Fy^-1[L psi], I3 = Fy^-1[L i ky psi],
//                                  I5 = Fy^-1[L (-ky^2) psi]
//   row_nl   : I0..I5           -> u, v, zeta, q, BR, BI rows (c2r in x),
//                                  the 8 products, r2c in x, pre-combined
//                                  flux tendencies -> I0..I5
//   col_upd  : I0..I5 + state   -> r2c in y, leapfrog + integrating factor +
//                                  Robert filter in place, then the inverse
//                                  y transforms of the new q, BR, BI (and of
//                                  i ky BR, i ky BI) -> I0 (q), I2,I3 (BR),
//                                  I4,I5 (BI)
//   row_qw   : I2..I5           -> J(BR,BI) and |B|^2 rows -> I3, I5
//   col_qw   : I3, I5, q^{n+1}  -> q* = q - q^w
//
// (then the vertical solves of qg_kernels.hip).  The reference computes the
// same quantities with 24 per-level 2D FFTW transforms and 8 MPI transposes
// per step (main_waqg.f90:326-467, derivatives.f90:607-1032).
//
// Every transform is a register-resident FFT (qg_regfft.hpp): a column
// workgroup holds TXC consecutive kx columns, T threads per column; a row
// workgroup holds RP rows, T threads per row.  Pointwise work between
// transforms (products, time update) runs in registers.
#include "qg_common.hpp"
#include "qg_regfft.hpp"
#include "qg_tx.hpp"

namespace qg {
namespace {

// ---------------------------------------------------------------------------
// ps_inv: inverse DFT in y of one spectral field into up to three
// intermediates with the multipliers 1, i ky, -ky^2 (u, v and zeta are formed
// from these in row_nl: compute_velo derivatives.f90:1302-1311,
// refraction_waqg :831-845), the mask L and the reference's realit fill of
// (0,-ky) after the multiplier (fft.f90:71-107).  Columns kx < W; rows of the
// dealiasing box only if box.  One launch runs up to four independent jobs
// (ColInvJobs): grid rows are the jobs' planes one after the other.
// ---------------------------------------------------------------------------
template <int N, int E>
__global__ __launch_bounds__((SH<N, E>::COLT)) void k_ps_inv(Geo g, const double2* __restrict__ tw, ColInvJobs jobs) {
  constexpr int T = SH<N, E>::T, TX = SH<N, E>::TX;
  __shared__ double2 lds[col_xlen<N, E, 1, TX>()];
  int z = blockIdx.y, ji = 0;
  while (ji + 1 < jobs.n && z >= jobs.j[ji].nz) z -= jobs.j[ji++].nz;
  const ColInvJob& J = jobs.j[ji];
  const int W = J.W;
  const int c = threadIdx.x % TX, t = threadIdx.x / TX;
  const int kx0 = blockIdx.x * TX, kx = kx0 + c;
  if (kx0 >= W) return;  // the grid is sized for the widest job
  const ColBuf<N, E> xb{lds + c};
  const TwSrc<N, E, (E <= 8)> tws(tw, t);  // before the data loads
  const double2* Fz = J.F + (size_t)z * g.ikty * g.sp;
  double2 v0[E];
#pragma unroll
  for (int m = 0; m < E; ++m) {
    const int iky = t + T * m;
    v0[m] = (kx < W && (!J.box || in_box_row(iky, g)) && (!J.mask || lmask(kx, iky, g)))
                ? Fz[(size_t)iky * g.sp + kx]
                : zero2();
  }
  const size_t pz = (size_t)z * g.n2 * g.sp;
#pragma unroll 1
  for (int o = 0; o < 3; ++o) {
    double2* out = J.o[o];
    if (!out) break;
    double2 v[E];
#pragma unroll
    for (int m = 0; m < E; ++m) {
      const double ky = (double)ky_of(t + T * m, g.n2);
      v[m] = o == 0 ? v0[m] : (o == 1 ? mul_ik(v0[m], ky) : cscale(v0[m], -(ky * ky)));
    }
    if (kx0 == 0) {
      rf_publish<N, E>(v, xb, t);
      if (c == 0) {
#pragma unroll
        for (int m = 0; m < E; ++m) {
          const int iky = t + T * m;
          if (iky > N / 2) v[m] = cconj(xb.pub(0, N - iky));
        }
      }
    }
    reg_fft<N, E, true>(v, xb, t, tws.get());
    if (kx < W) {
#pragma unroll
      for (int m = 0; m < E; ++m) st_nt(&out[pz + (size_t)(t + T * m) * g.sp + kx], v[m]);
    }
  }
}

// ---------------------------------------------------------------------------
// row_nl: per (z, y) row, c2r in x of u, v, zeta, q, BR, BI, the products of
// convol_waqg / refraction_waqg (derivatives.f90:661-674, 709-722, 756-769,
// 862-875), r2c in x and the pre-combination of the flux divergences:
//   I0 = i kx F(uq)                   I1 = F(vq)
//   I2 = i kx F(uBR) - F(zeta BI)/2   I3 = F(vBR)
//   I4 = i kx F(uBI) + F(zeta BR)/2   I5 = F(vBI)
// so that after the y transform  n_q = I0 + i ky I1,
// n_BR - r_BI/2 = I2 + i ky I3 and n_BI + r_BR/2 = I4 + i ky I5.
// CAP (kept tendencies) also writes F(zeta BR), F(zeta BI) to Z0, Z1, whose
// y transforms are r_BR, r_BI (refraction_waqg derivatives.f90:877-878).
// ---------------------------------------------------------------------------
template <int N, int E, int RP, int SB, bool CAP = false>
__global__ __launch_bounds__(RP*(N / E)) void k_row_nl(Dev d, int Win, double scale) {
  constexpr int T = N / E, RS = row_xstride<N, E, SB>();
  constexpr int MK0 = ((N / 3 + 2) + T - 1) / T;
  constexpr int MK = MK0 < E ? MK0 : E;  // positions t + T*m that can be < KX
  __shared__ double2 lds[RP * RS];
  const Geo& g = d.g;
  int t, p;
  row_lane<T, RP>(threadIdx.x, t, p);
  const RowBuf<N, E> xb{lds + p * RS};
  const size_t ro = ((size_t)blockIdx.x * RP + p) * g.sp;
  double2* const* I = d.I;
  const TwSrc<N, E, (E <= 8)> tws(d.tw, t);  // before the data loads
  // All loads are issued before any is used: positions with k >= Win read
  // entry 0 of the row (a valid address, same line as the others) and are
  // zeroed afterwards, so no branch splits the loads (a zero-initialised
  // branch makes the compiler wait on each load to copy it).
  double2 in[6][E];
#pragma unroll
  for (int m = 0; m < E; ++m) {
    const int pos = t + T * m;
    const int k = pos <= N / 2 ? pos : N - pos;
    const size_t a = ro + (k < Win ? k : 0);
#pragma unroll
    for (int f = 0; f < 6; ++f) in[f][m] = I[f][a];
  }
  double2 z[3][E];
#pragma unroll
  for (int m = 0; m < E; ++m) {
    const int pos = t + T * m;
    const int k = pos <= N / 2 ? pos : N - pos;
    const bool ok = k < Win;
    const double2 pq = ok ? in[0][m] : zero2(), ps = ok ? in[1][m] : zero2(), pbr = ok ? in[2][m] : zero2();
    const double2 py = ok ? in[3][m] : zero2(), pbi = ok ? in[4][m] : zero2(), pyy = ok ? in[5][m] : zero2();
    const double kk = (double)k, k2 = kk * kk;
    const double2 U = make_double2(-py.x, -py.y);
    const double2 V = mul_ik(ps, kk);
    const double2 Z = make_double2(-k2 * ps.x + pyy.x, -k2 * ps.y + pyy.y);
    z[0][m] = pack2<N>(U, V, pos);
    z[1][m] = pack2<N>(Z, pq, pos);
    z[2][m] = pack2<N>(pbr, pbi, pos);
  }
  reg_fft_multi<N, E, 3, SB, true>(z, xb, t, tws.get());
  // z0 = (u, v), z1 = (zeta, q), z2 = (BR, BI) at positions t + T m; the
  // product pairs, ordered so that each exchanged pair completes outputs:
  // (zeta BR, zeta BI), (u BR, v BR) | (u BI, v BI), (u q, v q)
  double2 pv[4][E];
#pragma unroll
  for (int m = 0; m < E; ++m) {
    const double u = z[0][m].x, v = z[0][m].y, zeta = z[1][m].x, q = z[1][m].y, br = z[2][m].x, bi = z[2][m].y;
    pv[0][m] = make_double2(zeta * br, zeta * bi);
    pv[1][m] = make_double2(u * br, v * br);
    pv[2][m] = make_double2(u * bi, v * bi);
    pv[3][m] = make_double2(u * q, v * q);
  }
  reg_fft_multi<N, E, 4, SB, false>(pv, xb, t, tws.get());
  const bool lin = (d.flags & F_LINEAR) != 0, pas = (d.flags & F_PASSIVE) != 0;
  const int KX = g.KX;
  double2 keepZR[MK];  // F(zeta BR)
  // first pair: I2 = i k F(u BR) - F(zeta BI)/2, I3 = F(v BR)
  rf_publish_range<N, E, 4, 0, 2>(pv, xb, t);
#pragma unroll
  for (int m = 0; m < MK; ++m) {
    const int k = t + T * m;
    keepZR[m] = zero2();
    if (k < KX) {
      const int km = (N - k) & (N - 1);
      double2 FzR, FzI, FuR, FvR;
      split2(pv[0][m], xb.pub(0, km), FzR, FzI);
      split2(pv[1][m], xb.pub(1, km), FuR, FvR);
      double2 w2 = mul_ik(FuR, (double)k), w3 = FvR;
      if (lin) { w2 = zero2(); w3 = zero2(); }
      if (!pas) w2 = make_double2(w2.x - 0.5 * FzI.x, w2.y - 0.5 * FzI.y);
      keepZR[m] = pas ? zero2() : FzR;
      st_nt(&I[2][ro + k], cscale(w2, scale));
      st_nt(&I[3][ro + k], cscale(w3, scale));
      if constexpr (CAP) {
        st_nt(&d.Z[0][ro + k], pas ? zero2() : cscale(FzR, scale));
        st_nt(&d.Z[1][ro + k], pas ? zero2() : cscale(FzI, scale));
      }
    }
  }
  // second pair: I4 = i k F(u BI) + F(zeta BR)/2, I5 = F(v BI),
  //              I0 = i k F(u q), I1 = F(v q)
  rf_publish_range<N, E, 4, 2, 2>(pv, xb, t);
#pragma unroll
  for (int m = 0; m < MK; ++m) {
    const int k = t + T * m;
    if (k < KX) {
      const int km = (N - k) & (N - 1);
      double2 FuI, FvI, Fuq, Fvq;
      split2(pv[2][m], xb.pub(0, km), FuI, FvI);
      split2(pv[3][m], xb.pub(1, km), Fuq, Fvq);
      const double kk = (double)k;
      double2 w0 = mul_ik(Fuq, kk), w1 = Fvq, w4 = mul_ik(FuI, kk), w5 = FvI;
      if (lin) { w0 = zero2(); w1 = zero2(); w4 = zero2(); w5 = zero2(); }
      w4 = make_double2(w4.x + 0.5 * keepZR[m].x, w4.y + 0.5 * keepZR[m].y);
      st_nt(&I[0][ro + k], cscale(w0, scale));
      st_nt(&I[1][ro + k], cscale(w1, scale));
      st_nt(&I[4][ro + k], cscale(w4, scale));
      st_nt(&I[5][ro + k], cscale(w5, scale));
    }
  }
}

// ---------------------------------------------------------------------------
// col_upd: r2c in y of the pre-combined tendencies of one field, the time
// update (Euler: main_waqg.f90:251-253; leapfrog + Robert-Asselin filter:
// :375-377, :400-402, :413-415) in place, then the inverse y transforms of
// the new field for the next step's products and for the feedback term
// (derivatives.f90:938-956).  blockIdx.z: 0 = q, 1 = BR, 2 = BI.
// ---------------------------------------------------------------------------
template <int N, int E, int SB>
__global__ __launch_bounds__((SH<N, E>::COLT)) void k_col_upd(Dev d, int euler, int has_dq) {
  constexpr int T = SH<N, E>::T, TX = SH<N, E>::TX;
  __shared__ double2 lds[col_xlen<N, E, SB, TX>()];
  const Geo& g = d.g;
  const int c = threadIdx.x % TX, t = threadIdx.x / TX;
  const int kx0 = blockIdx.x * TX, kx = kx0 + c, z = blockIdx.y, f = blockIdx.z;
  const ColBuf<N, E> xb{lds + c};
  double2* W1 = d.I[2 * f];
  double2* W2 = d.I[2 * f + 1];
  const size_t pz = (size_t)z * g.n2 * g.sp;
  const bool colok = kx < g.KX;
  const int toff = t * g.sp + kx;  // thread part of every row address
  const size_t mstride = (size_t)T * g.sp;
  double2 v[2][E];
#pragma unroll
  for (int m = 0; m < E; ++m) {
    v[0][m] = colok ? (W1 + pz + m * mstride)[toff] : zero2();
    v[1][m] = colok ? (W2 + pz + m * mstride)[toff] : zero2();
  }
  reg_fft_multi<N, E, 2, SB, false>(v, xb, t, d.tw);

  double2* X = f == 0 ? d.q : (f == 1 ? d.BR : d.BI);
  double2* Xo = f == 0 ? d.qo : (f == 1 ? d.BRo : d.BIo);
  const double2* Av = f == 1 ? d.AI : d.AR;
  const bool useA = f > 0 && !(d.flags & (F_NO_DISPERSION | F_PASSIVE));
  const double* e1t = f == 0 ? d.e1q : d.e1w;
  const double* e2t = f == 0 ? d.e2q : d.e2w;
  const double cd = 0.5 / (d.Bu * d.Ro);
  const double sgn = f == 1 ? 1.0 : -1.0;  // +c kh2 AI for BR, -c kh2 AR for BI
  const size_t sz = (size_t)z * g.ikty * g.sp;
  bool bad = false;
#pragma unroll
  for (int m = 0; m < E; ++m) {
    const int iky = t + T * m;
    double2 xnew = zero2();
    if (colok && in_box_row(iky, g)) {
      const double ky = (double)ky_of(iky, g.n2);
      const double2 F1 = v[0][m], F2 = v[1][m];
      const double2 tt = make_double2(F1.x - ky * F2.y, F1.y + ky * F2.x);
      const size_t idx = sz + (size_t)iky * g.sp + kx;
      double2 xonew = zero2();
      if (lmask(kx, iky, g)) {
        const size_t kidx = (size_t)iky * g.iktx + kx;  // (iky, kx) tables are unpadded
        const double2 x = ld_nt(&X[idx]);
        const double e1 = e1t[kidx];
        const double kh2 = (double)kx * (double)kx + ky * ky;
        double2 dqv = zero2();
        if (f == 0 && has_dq) dqv = d.dq[idx];
        double2 av = zero2();
        if (useA) av = ld_nt(&Av[idx]);
        if (euler) {
          if (f == 0) {
            xnew = make_double2(((x.x - d.dt * tt.x) + d.dt * dqv.x) * e1, ((x.y - d.dt * tt.y) + d.dt * dqv.y) * e1);
          } else {
            const double ck = (d.dt * cd) * kh2;
            xnew = make_double2(((x.x - d.dt * tt.x) - sgn * (ck * av.x)) * e1,
                                ((x.y - d.dt * tt.y) - sgn * (ck * av.y)) * e1);
          }
        } else {
          const double2 xo = ld_nt(&Xo[idx]);
          const double e2 = e2t[kidx];
          const double tdt = 2.0 * d.dt;
          if (f == 0) {
            xnew = make_double2((xo.x * e2 - (tdt * tt.x) * e1) + (tdt * dqv.x) * e2,
                                (xo.y * e2 - (tdt * tt.y) * e1) + (tdt * dqv.y) * e2);
          } else {
            const double ck = cd * kh2;
            xnew = make_double2(xo.x * e2 - (tdt * (tt.x + sgn * (ck * av.x))) * e1,
                                xo.y * e2 - (tdt * (tt.y + sgn * (ck * av.y))) * e1);
          }
          xonew = make_double2(x.x + d.gamma * ((xo.x - 2.0 * x.x) + xnew.x),
                               x.y + d.gamma * ((xo.y - 2.0 * x.y) + xnew.y));
        }
      }
      X[idx] = xnew;
      if (!euler) Xo[idx] = xonew;
      // the regular YBJ operator sums these tendencies over z (compute_sigma)
      if (f > 0 && d.tR) (f == 1 ? d.tR : d.tI)[idx] = tt;
      bad |= !(isfinite(xnew.x) && isfinite(xnew.y));
      v[0][m] = xnew;
      v[1][m] = mul_ik(xnew, ky);
    } else {
      v[0][m] = zero2();
      v[1][m] = zero2();
    }
  }
  if (bad) atomicOr(d.flag, 1);
  if (kx0 == 0) {
    // realit on the kx = 0 column of the new field and of i ky times it
    if constexpr (SB == 2) {
      rf_publish_range<N, E, 2, 0, 2>(v, xb, t);
      if (c == 0) {
#pragma unroll
        for (int m = 0; m < E; ++m) {
          const int iky = t + T * m;
          if (iky > N / 2) {
            v[0][m] = cconj(xb.pub(0, N - iky));
            v[1][m] = cconj(xb.pub(1, N - iky));
          }
        }
      }
    } else {
      rf_publish_range<N, E, 2, 0, 1>(v, xb, t);
      if (c == 0) {
#pragma unroll
        for (int m = 0; m < E; ++m) {
          const int iky = t + T * m;
          if (iky > N / 2) v[0][m] = cconj(xb.pub(0, N - iky));
        }
      }
      rf_publish_range<N, E, 2, 1, 1>(v, xb, t);
      if (c == 0) {
#pragma unroll
        for (int m = 0; m < E; ++m) {
          const int iky = t + T * m;
          if (iky > N / 2) v[1][m] = cconj(xb.pub(0, N - iky));
        }
      }
    }
  }
  reg_fft_multi<N, E, 2, SB, true>(v, xb, t, d.tw);
  if (colok) {
#pragma unroll
    for (int m = 0; m < E; ++m) st_nt(&(W1 + pz + m * mstride)[toff], v[0][m]);
    if (f > 0) {
#pragma unroll
      for (int m = 0; m < E; ++m) st_nt(&(W2 + pz + m * mstride)[toff], v[1][m]);
    }
  }
}

// ---------------------------------------------------------------------------
// row_qw: c2r in x of BR, BI, their x and y derivatives, the feedback
// products J = BRy BIx - BRx BIy and |B|^2 (derivatives.f90:967-1000), r2c in
// x; I3 = F(J) - kx^2/4 F(|B|^2), I5 = F(|B|^2)
// ---------------------------------------------------------------------------
template <int N, int E, int RP, int SB>
__global__ __launch_bounds__(RP*(N / E)) void k_row_qw(Dev d, double scale) {
  constexpr int T = N / E, RS = row_xstride<N, E, SB>();
  constexpr int MK0 = ((N / 3 + 2) + T - 1) / T;
  constexpr int MK = MK0 < E ? MK0 : E;
  __shared__ double2 lds[RP * RS];
  const Geo& g = d.g;
  int t, p;
  row_lane<T, RP>(threadIdx.x, t, p);
  const RowBuf<N, E> xb{lds + p * RS};
  const size_t ro = ((size_t)blockIdx.x * RP + p) * g.sp;
  double2* const* I = d.I;
  const int KX = g.KX;
  const TwSrc<N, E, (E <= 8)> tws(d.tw, t);  // before the data loads
  double2 z[3][E];
#pragma unroll
  for (int m = 0; m < E; ++m) {
    const int pos = t + T * m;
    const int k = pos <= N / 2 ? pos : N - pos;
    double2 br = zero2(), bry = zero2(), bi = zero2(), biy = zero2();
    if (k < KX) { br = I[2][ro + k]; bry = I[3][ro + k]; bi = I[4][ro + k]; biy = I[5][ro + k]; }
    const double kk = (double)k;
    z[0][m] = pack2<N>(br, bi, pos);
    z[1][m] = pack2<N>(mul_ik(br, kk), mul_ik(bi, kk), pos);
    z[2][m] = pack2<N>(bry, biy, pos);
  }
  reg_fft_multi<N, E, 3, SB, true>(z, xb, t, tws.get());
  // z0 = (BR, BI), z1 = (BRx, BIx), z2 = (BRy, BIy): (J, |B|^2)
  double2 pv[1][E];
#pragma unroll
  for (int m = 0; m < E; ++m)
    pv[0][m] = make_double2(z[2][m].x * z[1][m].y - z[1][m].x * z[2][m].y,
                            z[0][m].x * z[0][m].x + z[0][m].y * z[0][m].y);
  reg_fft_multi<N, E, 1, 1, false>(pv, xb, t, tws.get());
  rf_publish_range<N, E, 1, 0, 1>(pv, xb, t);
#pragma unroll
  for (int m = 0; m < MK; ++m) {
    const int k = t + T * m;
    if (k < KX) {
      double2 FJ, FM;
      split2(pv[0][m], xb.pub(0, (N - k) & (N - 1)), FJ, FM);
      const double c4 = 0.25 * (double)k * (double)k;
      st_nt(&I[3][ro + k], cscale(make_double2(FJ.x - c4 * FM.x, FJ.y - c4 * FM.y), scale));
      st_nt(&I[5][ro + k], cscale(FM, scale));
    }
  }
}

// ---------------------------------------------------------------------------
// col_qw: r2c in y of the feedback rows, q^w = Ro W2F L (F(J) - kh2/4 F(|B|^2))
// (derivatives.f90:1010-1027) and q* = q - q^w where L = 1
// (main_waqg.f90:430-439)
// ---------------------------------------------------------------------------
template <int N, int E, int SB>
__global__ __launch_bounds__((SH<N, E>::COLT)) void k_col_qw(Dev d) {
  constexpr int T = SH<N, E>::T, TX = SH<N, E>::TX;
  __shared__ double2 lds[col_xlen<N, E, SB, TX>()];
  const Geo& g = d.g;
  const int c = threadIdx.x % TX, t = threadIdx.x / TX;
  const int kx0 = blockIdx.x * TX, kx = kx0 + c, z = blockIdx.y;
  const ColBuf<N, E> xb{lds + c};
  const size_t pz = (size_t)z * g.n2 * g.sp;
  const bool colok = kx < g.KX;
  const int toff = t * g.sp + kx;
  const size_t mstride = (size_t)T * g.sp;
  double2 v[2][E];
#pragma unroll
  for (int m = 0; m < E; ++m) {
    v[0][m] = colok ? (d.I[3] + pz + m * mstride)[toff] : zero2();
    v[1][m] = colok ? (d.I[5] + pz + m * mstride)[toff] : zero2();
  }
  reg_fft_multi<N, E, 2, SB, false>(v, xb, t, d.tw);
  if (!colok) return;
  const size_t sz = (size_t)z * g.ikty * g.sp;
#pragma unroll
  for (int m = 0; m < E; ++m) {
    const int iky = t + T * m;
    if (in_box_row(iky, g)) {
      const size_t idx = sz + (size_t)iky * g.sp + kx;
      double2 out = zero2();
      if (lmask(kx, iky, g)) {
        const double ky = (double)ky_of(iky, g.n2);
        const double c4 = 0.25 * (ky * ky);
        const double2 qw = make_double2(((v[0][m].x - c4 * v[1][m].x) * d.Ro) * d.W2F,
                                        ((v[0][m].y - c4 * v[1][m].y) * d.Ro) * d.W2F);
        out = csub(d.q[idx], qw);
      }
      d.qs[idx] = out;
    }
  }
}

// Launchers: every kernel exists for E = 8 and E = 16 elements per thread;
// `e16` selects the variant (chosen per length by the tuning table).
template <int N>
struct LS {
  template <int E>
  static hipError_t ps_inv_e(const Dev& d, hipStream_t st, const ColInvJobs& jobs) {
    using S = SH<N, E>;
    int W = 0, nz = 0;
    for (int i = 0; i < jobs.n; ++i) {
      W = jobs.j[i].W > W ? jobs.j[i].W : W;
      nz += jobs.j[i].nz;
    }
    dim3 grid((W + S::TX - 1) / S::TX, nz);
    hipLaunchKernelGGL((k_ps_inv<N, E>), grid, dim3(S::COLT), 0, st, d.g, d.tw, jobs);
    return hipGetLastError();
  }
  template <int E, int WG, bool CAP = false>
  static hipError_t row_nl_e(const Dev& d, hipStream_t st, int Win) {
    constexpr int RP = SH<N, E>::template rp<WG>();
    const double scale = 1.0 / ((double)d.g.n1 * (double)d.g.n2);
    hipLaunchKernelGGL((k_row_nl<N, E, RP, 2, CAP>), dim3(d.g.n2 * d.g.nz / RP), dim3(RP * (N / E)), 0, st, d, Win,
                       scale);
    return hipGetLastError();
  }
  template <int E, int SB>
  static hipError_t col_upd_e(const Dev& d, hipStream_t st, int euler, int has_dq) {
    using S = SH<N, E>;
    dim3 grid((d.g.KX + S::TX - 1) / S::TX, d.g.nz, 3);
    hipLaunchKernelGGL((k_col_upd<N, E, SB>), grid, dim3(S::COLT), 0, st, d, euler, has_dq);
    return hipGetLastError();
  }
  template <int E, int WG>
  static hipError_t row_qw_e(const Dev& d, hipStream_t st) {
    constexpr int RP = SH<N, E>::template rp<WG>();
    const double scale = 1.0 / ((double)d.g.n1 * (double)d.g.n2);
    hipLaunchKernelGGL((k_row_qw<N, E, RP, 2>), dim3(d.g.n2 * d.g.nz / RP), dim3(RP * (N / E)), 0, st, d, scale);
    return hipGetLastError();
  }
  template <int E, int SB>
  static hipError_t col_qw_e(const Dev& d, hipStream_t st) {
    using S = SH<N, E>;
    dim3 grid((d.g.KX + S::TX - 1) / S::TX, d.g.nz);
    hipLaunchKernelGGL((k_col_qw<N, E, SB>), grid, dim3(S::COLT), 0, st, d);
    return hipGetLastError();
  }
  static hipError_t ps_inv(const Dev& d, hipStream_t st, const ColInvJobs& jobs) {
    return (d.tune.e16[0] & 1) ? ps_inv_e<16>(d, st, jobs) : ps_inv_e<8>(d, st, jobs);
  }
  // row variants: 0 = E8 in 64-thread blocks, 1 = E16/64, 2 = E8/128,
  // 3 = E16/32 (rows per block = threads / (N/E), at least one)
  static hipError_t row_nl(const Dev& d, hipStream_t st, int Win) {
    if (d.Z[0]) return row_nl_e<8, 64, true>(d, st, Win);  // kept tendencies: one shape
    switch (d.tune.e16[1]) {
      case 1: return row_nl_e<16, 64>(d, st, Win);
      case 2: return row_nl_e<8, 128>(d, st, Win);
      case 3: return row_nl_e<16, 32>(d, st, Win);
      default: return row_nl_e<8, 64>(d, st, Win);
    }
  }
  // column variants: 0 = E8, 1 = E16 (two exchange buffers), 2 = E8, 3 = E16
  // (one exchange buffer: half the LDS, twice the occupancy, more barriers)
  static hipError_t col_upd(const Dev& d, hipStream_t st, int euler, int has_dq) {
    switch (d.tune.e16[2]) {
      case 1: return col_upd_e<16, 2>(d, st, euler, has_dq);
      case 2: return col_upd_e<8, 1>(d, st, euler, has_dq);
      case 3: return col_upd_e<16, 1>(d, st, euler, has_dq);
      default: return col_upd_e<8, 2>(d, st, euler, has_dq);
    }
  }
  static hipError_t row_qw(const Dev& d, hipStream_t st) {
    switch (d.tune.e16[3]) {
      case 1: return row_qw_e<16, 64>(d, st);
      case 2: return row_qw_e<8, 128>(d, st);
      case 3: return row_qw_e<16, 32>(d, st);
      default: return row_qw_e<8, 64>(d, st);
    }
  }
  static hipError_t col_qw(const Dev& d, hipStream_t st) {
    switch (d.tune.e16[4]) {
      case 1: return col_qw_e<16, 2>(d, st);
      case 2: return col_qw_e<8, 1>(d, st);
      case 3: return col_qw_e<16, 1>(d, st);
      default: return col_qw_e<8, 2>(d, st);
    }
  }
};

}  // namespace

#define QG_SDISPATCH(n, call)                \
  switch (n) {                               \
    case 16: return LS<16>::call;            \
    case 32: return LS<32>::call;            \
    case 64: return LS<64>::call;            \
    case 128: return LS<128>::call;          \
    case 256: return LS<256>::call;          \
    case 512: return LS<512>::call;          \
    default: return hipErrorInvalidValue;    \
  }

hipError_t launch_col_inv(const Dev& d, hipStream_t st, const ColInvJobs& jobs) {
  if (jobs.n < 1 || jobs.n > ColInvJobs::MAXJ) return hipErrorInvalidValue;
  QG_SDISPATCH(d.g.n2, ps_inv(d, st, jobs))
}
hipError_t launch_col_inv(const Dev& d, hipStream_t st, const double2* F, double2* o0, double2* o1, double2* o2,
                          int W, int box, int mask, int nz) {
  ColInvJobs jobs{};
  jobs.j[0] = ColInvJob{F, {o0, o1, o2}, W, box, mask, nz};
  jobs.n = 1;
  return launch_col_inv(d, st, jobs);
}
hipError_t launch_row_nl(const Dev& d, hipStream_t st, int Win) { QG_SDISPATCH(d.g.n1, row_nl(d, st, Win)) }
hipError_t launch_col_upd(const Dev& d, hipStream_t st, int euler, int has_dq) {
  QG_SDISPATCH(d.g.n2, col_upd(d, st, euler, has_dq))
}
hipError_t launch_row_qw(const Dev& d, hipStream_t st) { QG_SDISPATCH(d.g.n1, row_qw(d, st)) }
hipError_t launch_col_qw(const Dev& d, hipStream_t st) { QG_SDISPATCH(d.g.n2, col_qw(d, st)) }

}  // namespace qg
