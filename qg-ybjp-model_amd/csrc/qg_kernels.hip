// HIP kernels around the QG-YBJ+ step for gfx950 (FP64 throughout): the
// vertical inversions (psi_solver, A_solver_ybj_plus: elliptic.f90:145-460),
// vertical diffusion, init_q, C = A_z, the FP64 diagnostics, mode clearing and
// the raw r2c / c2r transforms of fft.f90:55-107 used for initial conditions
// and tests.  The transform-method kernels of the step are in
// qg_step_kernels.hip.
#include "qg_common.hpp"
#include "qg_fft.hpp"

namespace qg {

static __device__ __forceinline__ double2 zc() { return make_double2(0.0, 0.0); }

// Two real half-spectra A, B (entry k of each) packed into one complex
// sequence Z whose inverse DFT is a + i b (a, b the real c2r of A and B).
// The imaginary parts of the k = 0 and k = N/2 entries are ignored, which is
// the c2r convention of FFTW and pocketfft.
template <int N>
__device__ __forceinline__ void put_pair(double2* lds, int S, int k, int s, double2 A, double2 B) {
  if (k == 0 || k == N / 2) {
    lds[k * S + s] = make_double2(A.x, B.x);
  } else {
    lds[k * S + s] = make_double2(A.x - B.y, A.y + B.x);
    lds[(N - k) * S + s] = make_double2(A.x + B.y, B.x - A.y);
  }
}

// Split the DFT Z of z = a + i b (a, b real) into A = DFT(a), B = DFT(b) at k.
template <int N>
__device__ __forceinline__ void split_pair(const double2* lds, int S, int k, int s, double2& A, double2& B) {
  const double2 Z = lds[k * S + s];
  const double2 Zm = cconj(lds[((N - k) & (N - 1)) * S + s]);
  A = make_double2((Z.x + Zm.x) * 0.5, (Z.y + Zm.y) * 0.5);
  const double2 dd = csub(Z, Zm);
  B = make_double2(dd.y * 0.5, -dd.x * 0.5);
}

__device__ __forceinline__ double2 mul_iky(double2 v, double ky) { return make_double2(-ky * v.y, ky * v.x); }

// ---------------------------------------------------------------------------
// configuration per transform length (columns per tile, rows per block)
// ---------------------------------------------------------------------------
template <int N>
struct Cfg;
template <> struct Cfg<16>  { static constexpr int TXI = 16, TXU = 16, PR = 16, PRC = 16; };
template <> struct Cfg<32>  { static constexpr int TXI = 16, TXU = 16, PR = 8,  PRC = 16; };
template <> struct Cfg<64>  { static constexpr int TXI = 16, TXU = 16, PR = 8,  PRC = 16; };
template <> struct Cfg<128> { static constexpr int TXI = 16, TXU = 8,  PR = 4,  PRC = 16; };
template <> struct Cfg<256> { static constexpr int TXI = 16, TXU = 8,  PR = 4,  PRC = 16; };
template <> struct Cfg<512> { static constexpr int TXI = 8,  TXU = 4,  PR = 2,  PRC = 16; };
constexpr int NT = 256;

// ---------------------------------------------------------------------------
// col_inv: inverse DFT in y of one spectral field, with an optional spectral
// multiplier, the mask L and the reference's realit fill of (0,-ky)
// (fft.f90:71-107).  Output: intermediate (kx, y) rows, columns kx < W.
// Used by the raw c2r entry point; the time step uses k_ps_inv.
// ---------------------------------------------------------------------------
template <int N, int TX>
__global__ __launch_bounds__(NT) void k_col_inv(Geo g, const double2* __restrict__ tw, const double2* __restrict__ F,
                                                 double2* __restrict__ out, int W, int box, int mul, int mask) {
  __shared__ double2 lds[N * TX];
  const int kx0 = blockIdx.x * TX, z = blockIdx.y;
  const double2* Fz = F + (size_t)z * g.ikty * g.sp;
  for (int e = threadIdx.x; e < N * TX; e += NT) {
    const int i = e / TX, c = e % TX, kx = kx0 + c;
    double2 v = zc();
    if (kx < W && (!box || in_box_row(i, g)) && (!mask || lmask(kx, i, g))) {
      v = Fz[(size_t)i * g.sp + kx];
      const double ky = (double)ky_of(i, g.n2);
      if (mul == 1) v = mul_iky(v, ky);
      else if (mul == 2) v = cscale(v, -(ky * ky));
    }
    lds[i * TX + c] = v;
  }
  __syncthreads();
  if (kx0 == 0) {
    for (int j = 1 + (int)threadIdx.x; j < N / 2; j += NT) lds[(N - j) * TX] = cconj(lds[j * TX]);
    __syncthreads();
  }
  fft_lds<N, true, TX, NT>(lds, tw);
  double2* oz = out + (size_t)z * g.n2 * g.sp;
  for (int e = threadIdx.x; e < N * TX; e += NT) {
    const int i = e / TX, c = e % TX, kx = kx0 + c;
    if (kx < W) oz[(size_t)i * g.sp + kx] = lds[i * TX + c];
  }
}

// ---------------------------------------------------------------------------
// solve: T(kh2) x = s * rhs for every box column, Re and Im parts in adjacent
// lanes (psi_solver elliptic.f90:145-247, A_solver_ybj_plus :346-460).
// LAPACK DGTSV elimination order without interchanges (the systems are
// strictly diagonally dominant); the eliminated diagonal is kept in scratch
// for the back substitution.  blockIdx.y: 0 = psi from q*, 1 = AR from BR,
// 2 = AI from BI.  Columns with kh2 = 0 or L = 0 get x = 0.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(NT) void k_solve(Dev d, int f_lo, const double2* __restrict__ rhs_psi) {
  const Geo& g = d.g;
  const int f = f_lo + (int)blockIdx.y / g.nm, mem = (int)blockIdx.y % g.nm;
  const int ncols = g.KX * g.nbox;
  const int gid = blockIdx.x * NT + threadIdx.x;
  const int col = gid >> 1, part = gid & 1;
  if (col >= ncols) return;
  const int kx = col % g.KX, iky = box_row(col / g.KX, g);
  const int ky = ky_of(iky, g.n2);
  const int kh2 = kx * kx + ky * ky;
  const double2* rhs = f == 0 ? rhs_psi : (f == 1 ? d.BR : d.BI);
  double2* out = f == 0 ? d.psi : (f == 1 ? d.AR : d.AI);
  const size_t plane2 = (size_t)2 * g.ikty * g.sp;
  const int n3 = g.n3;
  const double* rin = reinterpret_cast<const double*>(rhs) + part + (size_t)mem * n3 * plane2;
  double* xo = reinterpret_cast<double*>(out) + part + (size_t)mem * n3 * plane2;
  const size_t c2 = (size_t)2 * ((size_t)iky * g.sp + kx);
  if (kh2 == 0 || !lmask(kx, iky, g)) {
    for (int z = 0; z < n3; ++z) xo[z * plane2 + c2] = 0.0;
    return;
  }
  const double dz = d.dz;
  const double kt = f == 0 ? ((double)kh2 * dz) * dz : (((double)kh2 * dz) * dz) / 4.0;
  const double rs = f == 0 ? dz * dz : (dz * dz) * d.Bu;
  double* dsc = d.scratch + ((size_t)(f * g.nm + mem) * n3 * ncols) * 2 + (size_t)col * 2 + part;
  const size_t dst = (size_t)ncols * 2;
  double dp = -(d.dsum[0] + kt);
  double bp = rs * rin[c2];
  xo[c2] = bp;
  dsc[0] = dp;
  for (int z = 1; z < n3; ++z) {
    const double fact = d.dl[z - 1] / dp;
    const double dn = -(d.dsum[z] + kt) - fact * d.du[z - 1];
    const double bn = rs * rin[z * plane2 + c2] - fact * bp;
    xo[z * plane2 + c2] = bn;
    dsc[z * dst] = dn;
    dp = dn;
    bp = bn;
  }
  double x = bp / dp;
  xo[(size_t)(n3 - 1) * plane2 + c2] = x;
  for (int z = n3 - 2; z >= 0; --z) {
    x = (xo[z * plane2 + c2] - d.du[z] * x) / dsc[z * dst];
    xo[z * plane2 + c2] = x;
  }
}

// ---------------------------------------------------------------------------
// solve_c: the same systems with the z range of every column split into P
// chunks of M levels; LPC columns per block, one lane per complex column (the
// matrix is real, so Re and Im share one factorisation).  Each chunk
// eliminates its own block in registers (DGTSV order, no interchanges) for
// three right-hand sides: its data g and the two coupling spikes v (from the
// row above the chunk) and w (from the row below), so that
// x = g - x_above * v - x_below * w.  The 2P boundary unknowns of a column
// form a block-tridiagonal system with 2x2 blocks, solved by the lanes of
// chunk 0 in LDS; each chunk then reconstructs its levels.  Padding rows
// beyond n3 are identity rows decoupled from the system.
// ---------------------------------------------------------------------------
template <int M, int P, int LPC>
__global__ __launch_bounds__(LPC * P) void k_solve_c(Dev d, int f_lo, const double2* __restrict__ rhs_psi) {
  const Geo& g = d.g;
  const int lane = threadIdx.x % LPC, c = threadIdx.x / LPC;
  const int f = f_lo + (int)blockIdx.y / g.nm, mem = (int)blockIdx.y % g.nm;
  const int ncols = g.KX * g.nbox;
  const int col = blockIdx.x * LPC + lane;
  const bool valid = col < ncols;
  int kx = 0, iky = 0, kh2 = 0;
  if (valid) {
    kx = col % g.KX;
    iky = box_row(col / g.KX, g);
    const int ky = ky_of(iky, g.n2);
    kh2 = kx * kx + ky * ky;
  }
  const bool active = valid && kh2 != 0 && lmask(kx, iky, g);
  const size_t plane = (size_t)g.ikty * g.sp;
  const int n3 = g.n3;
  const size_t moff = (size_t)mem * n3 * plane;
  const double2* __restrict__ rhs = (f == 0 ? rhs_psi : (f == 1 ? d.BR : d.BI)) + moff;
  double2* __restrict__ out = (f == 0 ? d.psi : (f == 1 ? d.AR : d.AI)) + moff;
  const size_t ci = (size_t)iky * g.sp + kx;
  const double dz = d.dz;
  // inactive columns get a regular (kh2 = 1) system so that no lane divides
  // by zero; their output is forced to 0
  const double kh2d = active ? (double)kh2 : 1.0;
  const double kt = f == 0 ? (kh2d * dz) * dz : ((kh2d * dz) * dz) / 4.0;
  const double rs = f == 0 ? dz * dz : (dz * dz) * d.Bu;
  const int s = c * M;

  double2 gv[M];
  double rp[M], vv[M], wv[M];
#pragma unroll
  for (int i = 0; i < M; ++i) {
    const int z = s + i;
    double2 b = make_double2(0.0, 0.0);
    if (active && z < n3) b = rhs[z * plane + ci];
    gv[i] = make_double2(rs * b.x, rs * b.y);
  }
  const double ltop = (s > 0 && s < n3) ? d.dl[s - 1] : 0.0;  // coupling of row s to row s-1
  const int e = s + M - 1;
  const double ubot = (e < n3 - 1) ? d.du[e] : 0.0;          // coupling of row e to row e+1
  // forward elimination; rp holds the reciprocal pivots
#pragma unroll
  for (int i = 0; i < M; ++i) {
    const int z = s + i;
    const double dd = z < n3 ? -(d.dsum[z] + kt) : 1.0;
    if (i == 0) {
      rp[0] = 1.0 / dd;
      vv[0] = ltop;
    } else {
      const double lz = z < n3 ? d.dl[z - 1] : 0.0;
      const double uz = (z - 1 < n3 - 1) ? d.du[z - 1] : 0.0;
      const double fact = lz * rp[i - 1];
      rp[i] = 1.0 / (dd - fact * uz);
      gv[i].x -= fact * gv[i - 1].x;
      gv[i].y -= fact * gv[i - 1].y;
      vv[i] = -fact * vv[i - 1];
    }
    wv[i] = (i == M - 1) ? ubot : 0.0;
  }
  gv[M - 1].x *= rp[M - 1];
  gv[M - 1].y *= rp[M - 1];
  vv[M - 1] *= rp[M - 1];
  wv[M - 1] *= rp[M - 1];
#pragma unroll
  for (int i = M - 2; i >= 0; --i) {
    const int z = s + i;
    const double uz = z < n3 - 1 ? d.du[z] : 0.0;
    gv[i].x = (gv[i].x - uz * gv[i + 1].x) * rp[i];
    gv[i].y = (gv[i].y - uz * gv[i + 1].y) * rp[i];
    vv[i] = (vv[i] - uz * vv[i + 1]) * rp[i];
    wv[i] = (wv[i] - uz * wv[i + 1]) * rp[i];
  }
  // reduced system: y_c = (T_c, B_c) = (x_s, x_e);
  //   T_c + v_s B_{c-1} + w_s T_{c+1} = g_s ;  B_c + v_e B_{c-1} + w_e T_{c+1} = g_e
  // slots: 0,1 g_s (re, im); 2,3 g_e; 4 v_s; 5 v_e; 6 w_s; 7 w_e
  __shared__ double red[P][8][LPC];
  if constexpr (P > 1) {
    red[c][0][lane] = gv[0].x;
    red[c][1][lane] = gv[0].y;
    red[c][2][lane] = gv[M - 1].x;
    red[c][3][lane] = gv[M - 1].y;
    red[c][4][lane] = vv[0];
    red[c][5][lane] = vv[M - 1];
    red[c][6][lane] = wv[0];
    red[c][7][lane] = wv[M - 1];
    __syncthreads();
    if (c == 0) {
      // Block Thomas with A_b = [[0,vs],[0,ve]], C_b = [[ws,0],[we,0]].  Every
      // eliminated diagonal block has the form [[D0,0],[D2,1]]; (h0, h1,
      // 1/D0, D2) of block b overwrite (g_s, g_e, v_s, v_e) of chunk b.
      double rD0p = 1.0, D2p = 0.0;
      double2 h0p = make_double2(red[0][0][lane], red[0][1][lane]);
      double2 h1p = make_double2(red[0][2][lane], red[0][3][lane]);
      red[0][4][lane] = 1.0;
      red[0][5][lane] = 0.0;
#pragma unroll 1
      for (int b = 1; b < P; ++b) {
        const double i10 = -D2p * rD0p;  // second row of the inverse: (i10, 1)
        const double vs = red[b][4][lane], ve = red[b][5][lane];
        const double ws = red[b - 1][6][lane], we = red[b - 1][7][lane];
        const double m00 = vs * i10, m01 = vs, m10 = ve * i10, m11 = ve;
        const double rD0 = 1.0 / (1.0 - (m00 * ws + m01 * we));
        const double D2 = -(m10 * ws + m11 * we);
        double2 h0, h1;
        h0.x = red[b][0][lane] - (m00 * h0p.x + m01 * h1p.x);
        h0.y = red[b][1][lane] - (m00 * h0p.y + m01 * h1p.y);
        h1.x = red[b][2][lane] - (m10 * h0p.x + m11 * h1p.x);
        h1.y = red[b][3][lane] - (m10 * h0p.y + m11 * h1p.y);
        red[b][0][lane] = h0.x;
        red[b][1][lane] = h0.y;
        red[b][2][lane] = h1.x;
        red[b][3][lane] = h1.y;
        red[b][4][lane] = rD0;
        red[b][5][lane] = D2;
        rD0p = rD0; D2p = D2; h0p = h0; h1p = h1;
      }
      double2 yT = make_double2(0.0, 0.0);  // T of block b+1
#pragma unroll 1
      for (int b = P - 1; b >= 0; --b) {
        double2 t0 = make_double2(red[b][0][lane], red[b][1][lane]);
        double2 t1 = make_double2(red[b][2][lane], red[b][3][lane]);
        if (b < P - 1) {
          const double ws = red[b][6][lane], we = red[b][7][lane];
          t0.x -= ws * yT.x; t0.y -= ws * yT.y;
          t1.x -= we * yT.x; t1.y -= we * yT.y;
        }
        const double rD0 = red[b][4][lane], D2 = red[b][5][lane];
        const double2 T = make_double2(t0.x * rD0, t0.y * rD0);
        red[b][0][lane] = T.x;  // solution (T_b, B_b) into slots 0..3
        red[b][1][lane] = T.y;
        red[b][2][lane] = t1.x - D2 * T.x;
        red[b][3][lane] = t1.y - D2 * T.y;
        yT = T;
      }
    }
    __syncthreads();
  }
  if (!valid) return;
  double2 xl = make_double2(0.0, 0.0), xr = make_double2(0.0, 0.0);
  if constexpr (P > 1) {
    if (c > 0) xl = make_double2(red[c - 1][2][lane], red[c - 1][3][lane]);      // x at row s-1
    if (c < P - 1) xr = make_double2(red[c + 1][0][lane], red[c + 1][1][lane]);  // x at row e+1
  }
#pragma unroll
  for (int i = 0; i < M; ++i) {
    const int z = s + i;
    if (z < n3) {
      double2 x = make_double2(0.0, 0.0);
      if (active) {
        x.x = gv[i].x - xl.x * vv[i] - xr.x * wv[i];
        x.y = gv[i].y - xl.y * vv[i] - xr.y * wv[i];
      }
      out[z * plane + ci] = x;
    }
  }
}

// ---------------------------------------------------------------------------
// vertical diffusion of the old q (derivatives.f90:1877-1926), box modes
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(NT) void k_dq(Dev d) {
  const Geo& g = d.g;
  const int per = g.KX * g.nbox;
  const int e = blockIdx.x * NT + threadIdx.x;
  if (e >= per * g.nz) return;
  const int zt = e / per, z = zt % g.n3, r = e % per, kx = r % g.KX, iky = box_row(r / g.KX, g);
  const size_t plane = (size_t)g.ikty * g.sp;
  const size_t idx = zt * plane + (size_t)iky * g.sp + kx;
  double2 v = zc();
  if (lmask(kx, iky, g)) {
    const double dz2 = d.dz * d.dz;
    const double2 c = d.qo[idx];
    if (z == 0) {
      const double2 u = d.qo[idx + plane];
      v = make_double2(d.nuz * (u.x - c.x) / dz2, d.nuz * (u.y - c.y) / dz2);
    } else if (z == g.n3 - 1) {
      const double2 l = d.qo[idx - plane];
      v = make_double2(d.nuz * (l.x - c.x) / dz2, d.nuz * (l.y - c.y) / dz2);
    } else {
      const double2 u = d.qo[idx + plane], l = d.qo[idx - plane];
      v = make_double2(d.nuz * ((u.x - 2.0 * c.x) + l.x) / dz2, d.nuz * ((u.y - 2.0 * c.y) + l.y) / dz2);
    }
  }
  d.dq[idx] = v;
}

// zero every mode outside the box (all = 1: every mode)
__global__ __launch_bounds__(NT) void k_clear(Geo g, double2* f, int all) {
  const size_t plane = (size_t)g.ikty * g.sp;
  const size_t e = (size_t)blockIdx.x * NT + threadIdx.x;
  if (e >= plane * g.nz) return;
  const size_t r = e % plane;
  const int iky = (int)(r / g.sp), kx = (int)(r % g.sp);
  if (all || kx >= g.KX || !in_box_row(iky, g)) f[e] = zc();
}

// init_q (init.f90:591-652): q = T(kh2) psi / dz^2 at L = 1 modes
__global__ __launch_bounds__(NT) void k_init_q(Dev d, const double2* __restrict__ psi, double2* __restrict__ q) {
  const Geo& g = d.g;
  const size_t plane = (size_t)g.ikty * g.sp;
  const size_t e = (size_t)blockIdx.x * NT + threadIdx.x;
  if (e >= plane * g.nz) return;
  const int z = (int)(e / plane) % g.n3;
  const size_t r = e % plane;
  const int iky = (int)(r / g.sp), kx = (int)(r % g.sp);
  double2 v = zc();
  if (lmask(kx, iky, g)) {
    const int ky = ky_of(iky, g.n2);
    const double dz = d.dz;
    const double kt = ((double)(kx * kx + ky * ky) * dz) * dz;
    const double* a = d.a_ell;
    const int n3 = g.n3;
    const double2 c = psi[e];
    if (z == 0) {
      const double2 u = psi[e + plane];
      v = make_double2(a[0] * u.x - (a[0] + kt) * c.x, a[0] * u.y - (a[0] + kt) * c.y);
    } else if (z == n3 - 1) {
      const double2 l = psi[e - plane];
      v = make_double2(-(a[n3 - 2] + kt) * c.x + a[n3 - 2] * l.x, -(a[n3 - 2] + kt) * c.y + a[n3 - 2] * l.y);
    } else {
      const double2 u = psi[e + plane], l = psi[e - plane];
      const double dg = (a[z] + a[z - 1]) + kt;
      v = make_double2((a[z] * u.x - dg * c.x) + a[z - 1] * l.x, (a[z] * u.y - dg * c.y) + a[z - 1] * l.y);
    }
    v = make_double2(v.x / (dz * dz), v.y / (dz * dz));
  }
  q[e] = v;
}

// C = A_z: forward difference, zero at the top (elliptic.f90:437-450)
__global__ __launch_bounds__(NT) void k_cz(Dev d, const double2* __restrict__ A, double2* __restrict__ C) {
  const Geo& g = d.g;
  const size_t plane = (size_t)g.ikty * g.sp;
  const size_t e = (size_t)blockIdx.x * NT + threadIdx.x;
  if (e >= plane * g.nz) return;
  const int z = (int)(e / plane) % g.n3;
  double2 v = zc();
  if (z < g.n3 - 1) {
    const double2 a1 = A[e + plane], a0 = A[e];
    v = make_double2((a1.x - a0.x) / d.dz, (a1.y - a0.y) / d.dz);
  }
  C[e] = v;
}

// FP64 diagnostic sums (diagnostics.f90:127-158, :683-712; enstrophy as
// sum kh2^2 |psi|^2), one partial per block, 6 values:
// KE, PE, WKE, WPE, WCE, ENS (each still to be divided by n3)
__global__ __launch_bounds__(NT) void k_diag(Dev d) {
  const Geo& g = d.g;
  const int per = g.KX * g.nbox;
  const int e = blockIdx.x * NT + threadIdx.x;
  double acc[6] = {0, 0, 0, 0, 0, 0};
  if (e < per * g.n3) {
    const int z = e / per, r = e % per, kx = r % g.KX, iky = box_row(r / g.KX, g);
    if (lmask(kx, iky, g)) {
      const size_t plane = (size_t)g.ikty * g.sp;
      const size_t idx = ((size_t)blockIdx.y * g.n3 + z) * plane + (size_t)iky * g.sp + kx;
      const int ky = ky_of(iky, g.n2);
      const double kh2 = (double)(kx * kx + ky * ky);
      const double w = (kx == 0 && iky == 0) ? 0.5 : 1.0;
      const double2 ps = d.psi[idx];
      const double a2 = ps.x * ps.x + ps.y * ps.y;
      const double dz = d.dz;
      double2 b;
      if (z == 0) {
        const double2 u = d.psi[idx + plane];
        b = make_double2((u.x - ps.x) / (2.0 * dz), (u.y - ps.y) / (2.0 * dz));
      } else if (z == g.n3 - 1) {
        const double2 l = d.psi[idx - plane];
        b = make_double2((ps.x - l.x) / (2.0 * dz), (ps.y - l.y) / (2.0 * dz));
      } else {
        const double2 u = d.psi[idx + plane], l = d.psi[idx - plane];
        b = make_double2((u.x - l.x) / (2.0 * dz), (u.y - l.y) / (2.0 * dz));
      }
      const double Bu = d.Bu;
      const double2 br = d.BR[idx], bi = d.BI[idx], ar = d.AR[idx], ai = d.AI[idx];
      double2 cr = zc(), ci = zc();
      if (z < g.n3 - 1) {
        const double2 ar1 = d.AR[idx + plane], ai1 = d.AI[idx + plane];
        cr = make_double2((ar1.x - ar.x) / dz, (ar1.y - ar.y) / dz);
        ci = make_double2((ai1.x - ai.x) / dz, (ai1.y - ai.y) / dz);
      }
      acc[0] = w * kh2 * a2;
      acc[1] = w * (b.x * b.x + b.y * b.y) * (Bu / d.N2s[z]);
      acc[2] = w * (br.x * br.x + br.y * br.y + bi.x * bi.x + bi.y * bi.y);
      acc[3] = (0.5 / (d.N2u[z] * Bu)) * kh2 * (cr.x * cr.x + cr.y * cr.y + ci.x * ci.x + ci.y * ci.y);
      acc[4] = (1.0 / 8.0) * (1.0 / (Bu * Bu)) * kh2 * kh2 * (ar.x * ar.x + ar.y * ar.y + ai.x * ai.x + ai.y * ai.y);
      acc[5] = kh2 * kh2 * a2;
    }
  }
  __shared__ double red[6][NT];
#pragma unroll
  for (int j = 0; j < 6; ++j) red[j][threadIdx.x] = acc[j];
  __syncthreads();
  for (int s = NT / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
#pragma unroll
      for (int j = 0; j < 6; ++j) red[j][threadIdx.x] += red[j][threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x < 6) d.diag_part[((size_t)blockIdx.y * gridDim.x + blockIdx.x) * 6 + threadIdx.x] = red[threadIdx.x][0];
}

// Kept tendencies: with Z0, Z1 already y-transformed in place (r2c of
// zeta BR, zeta BI), the update's combined tendencies tR = nBR - rBI/2 and
// tI = nBI + rBR/2 are split into the reference's separate terms
// (convol_waqg derivatives.f90:783-784, refraction_waqg :877-878), all masked
// by L and zero elsewhere: tR <- nBR, tI <- nBI, Z0 <- rBR, Z1 <- rBI.
// Under F_LINEAR nB is set to zero as the reference sets it
// (main_waqg.f90:342-346), not left as the rounding residue of tR + rBI/2.
__global__ __launch_bounds__(NT) void k_tend_split(Dev d) {
  const Geo& g = d.g;
  const size_t n = (size_t)g.nz * g.ikty * g.iktx;
  const size_t e = (size_t)blockIdx.x * NT + threadIdx.x;
  if (e >= n) return;
  const int kx = (int)(e % g.iktx);
  const size_t row = e / g.iktx;  // z * ikty + iky
  const int iky = (int)(row % g.ikty);
  const size_t idx = row * g.sp + kx;
  double2 rr = zc(), ri = zc(), nr = zc(), ni = zc();
  if (lmask(kx, iky, g)) {
    rr = d.Z[0][idx];
    ri = d.Z[1][idx];
    if (!(d.flags & F_LINEAR)) {
      const double2 tr = d.tR[idx], ti = d.tI[idx];
      nr = make_double2(tr.x + 0.5 * ri.x, tr.y + 0.5 * ri.y);
      ni = make_double2(ti.x - 0.5 * rr.x, ti.y - 0.5 * rr.y);
    }
  }
  d.Z[0][idx] = rr;
  d.Z[1][idx] = ri;
  d.tR[idx] = nr;
  d.tI[idx] = ni;
}

// Wave energy conversion sums of we_conversion (diagnostics.f90:967-1076)
// and gamma_conversion (:1080-1256) by Parseval instead of physical-space
// products: with D = -kh2 A and every field an L-masked half spectrum,
//   sum_x f g = n1 n2 sum_k 2 Re(f_k conj(g_k))   (kh2 > 0 modes, which
// carry their kx = 0 mirror images), so per member, still to be divided by n3:
//   S0 = sum 2 Re(DR nBR* + DI nBI*)          (advection)
//   S1 = sum 2 Re(-DR rBI* + DI rBR*)         (refraction)
//   S2 = sum 2 Re(DR dR* + DI dI*), d = -(c/2) B with c = fw / dt  (dissipation)
// one partial per block in diag_part (slots 0..2 of 6).
__global__ __launch_bounds__(NT) void k_conv(Dev d) {
  const Geo& g = d.g;
  const int per = g.KX * g.nbox;
  const int e = blockIdx.x * NT + threadIdx.x;
  double acc[3] = {0, 0, 0};
  if (e < per * g.n3) {
    const int z = e / per, r = e % per, kx = r % g.KX, iky = box_row(r / g.KX, g);
    const int ky = ky_of(iky, g.n2);
    const double kh2 = (double)(kx * kx + ky * ky);
    if (lmask(kx, iky, g) && kh2 > 0.0) {
      const size_t idx = ((size_t)blockIdx.y * g.n3 + z) * g.ikty * g.sp + (size_t)iky * g.sp + kx;
      const double2 ar = d.AR[idx], ai = d.AI[idx];
      const double2 DR = make_double2(-kh2 * ar.x, -kh2 * ar.y), DI = make_double2(-kh2 * ai.x, -kh2 * ai.y);
      const double2 nR = d.tR[idx], nI = d.tI[idx], rR = d.Z[0][idx], rI = d.Z[1][idx];
      const double2 br = d.BR[idx], bi = d.BI[idx];
      const double hc = -0.5 * d.cw[(size_t)iky * g.iktx + kx];
      acc[0] = 2.0 * ((DR.x * nR.x + DR.y * nR.y) + (DI.x * nI.x + DI.y * nI.y));
      acc[1] = 2.0 * ((DI.x * rR.x + DI.y * rR.y) - (DR.x * rI.x + DR.y * rI.y));
      acc[2] = 2.0 * hc * ((DR.x * br.x + DR.y * br.y) + (DI.x * bi.x + DI.y * bi.y));
    }
  }
  __shared__ double red[3][NT];
#pragma unroll
  for (int j = 0; j < 3; ++j) red[j][threadIdx.x] = acc[j];
  __syncthreads();
  for (int s = NT / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
#pragma unroll
      for (int j = 0; j < 3; ++j) red[j][threadIdx.x] += red[j][threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x < 3) d.diag_part[((size_t)blockIdx.y * gridDim.x + blockIdx.x) * 6 + threadIdx.x] = red[threadIdx.x][0];
}

// ---------------------------------------------------------------------------
// Regular YBJ operator (ybj_plus = 0, main_waqg.f90:449-461), per (kx, ky)
// column of a member:
//   sumB (derivatives.f90:1072-1118): B -= mean_z B
//   compute_sigma (:1037-1067): sigma = Bu Ro sum_z (rBR + 2 nBI, rBI - 2 nBR)/kh2
//     = Bu Ro sum_z (2 tI, -2 tR)/kh2 with the update's combined tendencies
//     tR = nBR - rBI/2, tI = nBI + rBR/2
//   compute_A (:1120-1231): A~_0 = 0, A~_k = A~_{k-1} + (sum_{j<k} B_j) N^2_u(k-1) dz^2,
//     A = A~ + (sigma - sum_k A~_k)/n3
// at the L = 1 modes with kh2 > 0; A = 0 at the other box modes.
// A workgroup takes YC columns and splits each column's levels into YZ
// chunks, one thread per (column, chunk).  With the chunk's own running sum
// lb of B and the weights w_k = N^2_u(k-1) dz^2 (w_0 = 0),
//   A~_k = A~_in + sb_in W(k) + V(k),  W(k) = sum_{s<=j<=k} w_j,
//   V(k) = sum_{s<=j<=k} w_j lb(j-1),
// where A~_in, sb_in (the A~ and running B sum entering the chunk at level s)
// come from combining the chunks' totals in order in LDS.  The sums are
// associated differently from a single sweep, identically on every run.
// ---------------------------------------------------------------------------
constexpr int YC = 64, YZ = 4;

__global__ __launch_bounds__(YC* YZ) void k_ybj_regular(Dev d) {
  const Geo& g = d.g;
  const int c = threadIdx.x % YC, q = threadIdx.x / YC;
  const int ncols = g.KX * g.nbox;
  const int e = blockIdx.x * YC + c;
  const bool colok = e < ncols;
  const int ec = colok ? e : 0;
  const int kx = ec % g.KX, iky = box_row(ec / g.KX, g);
  const size_t plane = (size_t)g.ikty * g.sp;
  const size_t c0 = (size_t)blockIdx.y * g.n3 * plane + (size_t)iky * g.sp + kx;
  const int ky = ky_of(iky, g.n2);
  const int kh2i = kx * kx + ky * ky;
  const int n3 = g.n3, len = (n3 + YZ - 1) / YZ;
  const int z0 = q * len < n3 ? q * len : n3, z1 = z0 + len < n3 ? z0 + len : n3;
  const bool act = colok && lmask(kx, iky, g) && kh2i != 0;
  __shared__ double2 sh[8][YZ][YC];
  __shared__ double shr[2][YZ][YC];
  // pass 1: chunk sums of B and of the sigma terms
  double2 sR = zc(), sI = zc(), g1 = zc(), g2 = zc();
  const double kh2 = (double)(kh2i > 0 ? kh2i : 1);
  if (act) {
    for (int z = z0; z < z1; ++z) {
      const size_t i = c0 + z * plane;
      const double2 br = d.BR[i], bi = d.BI[i], tr = d.tR[i], ti = d.tI[i];
      sR = make_double2(sR.x + br.x, sR.y + br.y);
      sI = make_double2(sI.x + bi.x, sI.y + bi.y);
      g1 = make_double2(g1.x + (2.0 * ti.x) / kh2, g1.y + (2.0 * ti.y) / kh2);
      g2 = make_double2(g2.x + (-2.0 * tr.x) / kh2, g2.y + (-2.0 * tr.y) / kh2);
    }
  }
  sh[0][q][c] = sR;
  sh[1][q][c] = sI;
  sh[2][q][c] = g1;
  sh[3][q][c] = g2;
  __syncthreads();
  sR = zc(); sI = zc(); g1 = zc(); g2 = zc();
  for (int p = 0; p < YZ; ++p) {
    sR = make_double2(sR.x + sh[0][p][c].x, sR.y + sh[0][p][c].y);
    sI = make_double2(sI.x + sh[1][p][c].x, sI.y + sh[1][p][c].y);
    g1 = make_double2(g1.x + sh[2][p][c].x, g1.y + sh[2][p][c].y);
    g2 = make_double2(g2.x + sh[3][p][c].x, g2.y + sh[3][p][c].y);
  }
  __syncthreads();
  const double bro = d.Bu * d.Ro, dz2 = d.dz * d.dz;
  const double2 mR = make_double2(sR.x / n3, sR.y / n3), mI = make_double2(sI.x / n3, sI.y / n3);
  const double2 sgR = make_double2(g1.x * bro, g1.y * bro), sgI = make_double2(g2.x * bro, g2.y * bro);
  // pass 2: chunk totals L (sum b), W, V and the sums of W and V over the chunk
  double2 lbR = zc(), lbI = zc(), vR = zc(), vI = zc(), svR = zc(), svI = zc();
  double W = 0.0, sW = 0.0;
  if (act) {
    for (int z = z0; z < z1; ++z) {
      const size_t i = c0 + z * plane;
      const double2 br0 = d.BR[i], bi0 = d.BI[i];
      const double wz = z > 0 ? d.N2u[z - 1] * dz2 : 0.0;
      W += wz;
      vR = make_double2(vR.x + wz * lbR.x, vR.y + wz * lbR.y);
      vI = make_double2(vI.x + wz * lbI.x, vI.y + wz * lbI.y);
      sW += W;
      svR = make_double2(svR.x + vR.x, svR.y + vR.y);
      svI = make_double2(svI.x + vI.x, svI.y + vI.y);
      lbR = make_double2(lbR.x + (br0.x - mR.x), lbR.y + (br0.y - mR.y));
      lbI = make_double2(lbI.x + (bi0.x - mI.x), lbI.y + (bi0.y - mI.y));
    }
  }
  sh[0][q][c] = lbR;
  sh[1][q][c] = lbI;
  sh[2][q][c] = vR;
  sh[3][q][c] = vI;
  sh[4][q][c] = svR;
  sh[5][q][c] = svI;
  shr[0][q][c] = W;
  shr[1][q][c] = sW;
  __syncthreads();
  // in-order combine: A~ and the running B sum entering this chunk, sum of A~
  double2 aR = zc(), aI = zc(), sbR = zc(), sbI = zc(), saR = zc(), saI = zc();
  double2 inAR = zc(), inAI = zc(), inSR = zc(), inSI = zc();
  for (int p = 0; p < YZ; ++p) {
    if (p == q) {
      inAR = aR; inAI = aI; inSR = sbR; inSI = sbI;
    }
    const int n0 = p * len < n3 ? p * len : n3, n1 = n0 + len < n3 ? n0 + len : n3;
    const double np = (double)(n1 - n0), Wp = shr[0][p][c], sWp = shr[1][p][c];
    saR = make_double2(saR.x + (np * aR.x + sbR.x * sWp + sh[4][p][c].x), saR.y + (np * aR.y + sbR.y * sWp + sh[4][p][c].y));
    saI = make_double2(saI.x + (np * aI.x + sbI.x * sWp + sh[5][p][c].x), saI.y + (np * aI.y + sbI.y * sWp + sh[5][p][c].y));
    aR = make_double2(aR.x + sbR.x * Wp + sh[2][p][c].x, aR.y + sbR.y * Wp + sh[2][p][c].y);
    aI = make_double2(aI.x + sbI.x * Wp + sh[3][p][c].x, aI.y + sbI.y * Wp + sh[3][p][c].y);
    sbR = make_double2(sbR.x + sh[0][p][c].x, sbR.y + sh[0][p][c].y);
    sbI = make_double2(sbI.x + sh[1][p][c].x, sbI.y + sh[1][p][c].y);
  }
  if (!colok) return;
  if (!act) {
    for (int z = z0; z < z1; ++z) {
      d.AR[c0 + z * plane] = zc();
      d.AI[c0 + z * plane] = zc();
    }
    return;
  }
  const double2 cR = make_double2((sgR.x - saR.x) / n3, (sgR.y - saR.y) / n3);
  const double2 cI = make_double2((sgI.x - saI.x) / n3, (sgI.y - saI.y) / n3);
  // pass 3: A = A~ + (sigma - sum A~)/n3 and the mean-free B
  lbR = zc(); lbI = zc(); vR = zc(); vI = zc();
  W = 0.0;
  for (int z = z0; z < z1; ++z) {
    const size_t i = c0 + z * plane;
    const double2 br0 = d.BR[i], bi0 = d.BI[i];
    const double2 br = make_double2(br0.x - mR.x, br0.y - mR.y), bi = make_double2(bi0.x - mI.x, bi0.y - mI.y);
    const double wz = z > 0 ? d.N2u[z - 1] * dz2 : 0.0;
    W += wz;
    vR = make_double2(vR.x + wz * lbR.x, vR.y + wz * lbR.y);
    vI = make_double2(vI.x + wz * lbI.x, vI.y + wz * lbI.y);
    d.AR[i] = make_double2(inAR.x + inSR.x * W + vR.x + cR.x, inAR.y + inSR.y * W + vR.y + cR.y);
    d.AI[i] = make_double2(inAI.x + inSI.x * W + vI.x + cI.x, inAI.y + inSI.y * W + vI.y + cI.y);
    d.BR[i] = br;
    d.BI[i] = bi;
    lbR = make_double2(lbR.x + br.x, lbR.y + br.y);
    lbI = make_double2(lbI.x + bi.x, lbI.y + bi.y);
  }
}

hipError_t launch_ybj_regular(const Dev& d, hipStream_t st) {
  if (!d.tR || !d.tI) return hipErrorInvalidValue;
  const int nb = (d.g.KX * d.g.nbox + YC - 1) / YC;
  hipLaunchKernelGGL(k_ybj_regular, dim3(nb, d.g.nm), dim3(YC * YZ), 0, st, d);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Omega equation (derivatives.f90:1469-1586, elliptic.f90:463-567), a
// diagnostic: the right-hand side 2 J(psi_z, lap psi) at the unstaggered
// levels through the raw transforms, then per (kx, ky) column the Dirichlet
// problem w_{k-1} - (2 + dz^2 kh2 N^2_u(k)/Bu) w_k + w_{k+1} = dz^2 rhs_k on
// the n3 - 1 interior interfaces (w = 0 at the top).
// ---------------------------------------------------------------------------
// bx, by = i k psi_z, xx, xy = -i k kh2 psibar at L = 1 (0 elsewhere), with
// psi_z = (psi_{k+1} - psi_k)/dz and psibar = (psi_{k+1} + psi_k)/2 at
// interface k, psi_z = 0 and psibar = psi at the top level of each member
__global__ __launch_bounds__(NT) void k_omega_prep(Dev d, double2* __restrict__ bx, double2* __restrict__ by,
                                                   double2* __restrict__ xx, double2* __restrict__ xy) {
  const Geo& g = d.g;
  const size_t plane = (size_t)g.ikty * g.sp;
  const size_t e = (size_t)blockIdx.x * NT + threadIdx.x;
  if (e >= plane * g.nz) return;
  const int z = (int)(e / plane), r = (int)(e % plane);
  const int iky = r / g.sp, kx = r % g.sp;
  double2 vbx = zc(), vby = zc(), vxx = zc(), vxy = zc();
  if (kx < g.iktx && lmask(kx, iky, g)) {
    const bool top = (z % g.n3) == g.n3 - 1;
    const double2 p0 = d.psi[e];
    double2 pz = zc(), pb = p0;
    if (!top) {
      const double2 p1 = d.psi[e + plane];
      pz = make_double2((p1.x - p0.x) / d.dz, (p1.y - p0.y) / d.dz);
      pb = make_double2(0.5 * (p1.x + p0.x), 0.5 * (p1.y + p0.y));
    }
    const double k_x = (double)kx, k_y = (double)ky_of(iky, g.n2), kh2 = k_x * k_x + k_y * k_y;
    vbx = make_double2(-k_x * pz.y, k_x * pz.x);
    vby = make_double2(-k_y * pz.y, k_y * pz.x);
    vxx = make_double2(k_x * kh2 * pb.y, -k_x * kh2 * pb.x);
    vxy = make_double2(k_y * kh2 * pb.y, -k_y * kh2 * pb.x);
  }
  bx[e] = vbx;
  by[e] = vby;
  xx[e] = vxx;
  xy[e] = vxy;
}

__global__ __launch_bounds__(NT) void k_omega_prod(const double* __restrict__ bx, const double* __restrict__ by,
                                                   const double* __restrict__ xx, const double* __restrict__ xy,
                                                   double* __restrict__ out, size_t n) {
  const size_t e = (size_t)blockIdx.x * NT + threadIdx.x;
  if (e < n) out[e] = 2.0 * (bx[e] * xy[e] - by[e] * xx[e]);
}

// Thomas sweep per box column of a member (no pivoting: the diagonal
// dominates, |d| > 2 = |dl| + |du|); c' in the solver scratch
__global__ __launch_bounds__(NT) void k_omega_solve(Dev d, const double2* __restrict__ rhs, double2* __restrict__ w) {
  const Geo& g = d.g;
  const int ncols = g.KX * g.nbox;
  const int e = blockIdx.x * NT + threadIdx.x;
  if (e >= ncols) return;
  const int kx = e % g.KX, iky = box_row(e / g.KX, g);
  const size_t plane = (size_t)g.ikty * g.sp;
  const size_t c0 = (size_t)blockIdx.y * g.n3 * plane + (size_t)iky * g.sp + kx;
  const int ky = ky_of(iky, g.n2);
  const int kh2i = kx * kx + ky * ky;
  const int m = g.n3 - 1;
  if (!lmask(kx, iky, g) || kh2i == 0 || m < 1) {
    for (int z = 0; z < g.n3; ++z) w[c0 + z * plane] = zc();
    return;
  }
  double* cp = d.scratch + (size_t)blockIdx.y * g.n3 * ncols + e;
  const double dz2 = d.dz * d.dz, kh2 = (double)kh2i;
  double2 xp = zc();
  double cprev = 0.0;
  for (int k = 0; k < m; ++k) {
    const double dk = -(2.0 + dz2 * kh2 / d.a_ell[k]);
    const double den = k == 0 ? dk : dk - cprev;
    const double2 b = rhs[c0 + k * plane];
    const double2 bk = make_double2(dz2 * b.x, dz2 * b.y);
    xp = k == 0 ? make_double2(bk.x / den, bk.y / den) : make_double2((bk.x - xp.x) / den, (bk.y - xp.y) / den);
    cprev = 1.0 / den;
    cp[(size_t)k * ncols] = cprev;
    w[c0 + k * plane] = xp;
  }
  w[c0 + (size_t)m * plane] = zc();
  double2 nx = xp;
  for (int k = m - 2; k >= 0; --k) {
    const double2 x = w[c0 + k * plane];
    const double c = cp[(size_t)k * ncols];
    nx = make_double2(x.x - c * nx.x, x.y - c * nx.y);
    w[c0 + k * plane] = nx;
  }
}

hipError_t launch_omega(const Dev& d, hipStream_t st, double2* const sp[4], double* const ph[4], double2* w) {
  const Geo& g = d.g;
  const size_t nsp = (size_t)g.ikty * g.sp * g.nz;
  hipLaunchKernelGGL(k_omega_prep, dim3((unsigned)((nsp + NT - 1) / NT)), dim3(NT), 0, st, d, sp[0], sp[1], sp[2],
                     sp[3]);
  hipError_t e = hipGetLastError();
  for (int i = 0; i < 4 && e == hipSuccess; ++i) e = launch_c2r(d, st, sp[i], ph[i], g.n1, g.nz);
  if (e != hipSuccess) return e;
  const size_t nph = (size_t)g.n1 * g.n2 * g.nz;
  // 2 (bx xy - by xx), in the buffer of bx
  hipLaunchKernelGGL(k_omega_prod, dim3((unsigned)((nph + NT - 1) / NT)), dim3(NT), 0, st, ph[0], ph[1], ph[2], ph[3],
                     ph[0], nph);
  e = hipGetLastError();
  if (e == hipSuccess) e = launch_r2c(d, st, ph[0], g.n1, sp[0], g.nz);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_omega_solve, dim3((g.KX * g.nbox + NT - 1) / NT, g.nm), dim3(NT), 0, st, d, sp[0], w);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// hspec_waves (diagnostics.f90:547-638): horizontal spectra of the waves at
// one level.  The L = 1 modes fall into shells mode = floor(|k| + 1/2); per
// shell
//   KE = 1/2 sum |BR + i BI|^2
//   PE = 1/2 sum 1/4 kh^2 / Bu (|C_{k-1}|^2 / N^2_{k-1} + |C_k|^2 / N^2_k)
//        with C = A_z at the interfaces below and above the level (only the
//        one above at the bottom level; C = 0 at the top)
//   modes = 2 per mode.
// One workgroup per (level, member, chunk of HCH of the box): each round
// stages one mode per thread in LDS; thread t then sums the staged modes of
// shells t and t + NSP in staging order into the chunk's partial spectrum,
// and k_hspec_sum adds the chunks' partials in chunk order, so the sums do
// not depend on scheduling.
// ---------------------------------------------------------------------------
constexpr int NSP = 256, HCH = 32;

__global__ __launch_bounds__(NSP) void k_hspec_waves(Dev d, SpecLevels L, int nbin, double* __restrict__ part) {
  const Geo& g = d.g;
  const int li = blockIdx.x, m = blockIdx.y, ch = blockIdx.z, t = threadIdx.x;
  const int z = L.lev[li] - 1;
  const size_t plane = (size_t)g.ikty * g.sp;
  const size_t mz = (size_t)m * g.n3;
  const int per = g.KX * g.nbox;
  const double Bu = d.Bu, dz = d.dz;
  __shared__ int cb[NSP];
  __shared__ double cke[NSP], cpe[NSP];
  double ke[2] = {0.0, 0.0}, pe[2] = {0.0, 0.0}, nm[2] = {0.0, 0.0};
  const int len = (per + HCH - 1) / HCH, beg = ch * len, end = min(per, beg + len);
  for (int base = beg; base < end; base += NSP) {
    const int r = base + t;
    int bin = -1;
    double k1 = 0.0, p1 = 0.0;
    if (r < end) {
      const int kx = r % g.KX, iky = box_row(r / g.KX, g);
      if (lmask(kx, iky, g)) {
        const int ky = ky_of(iky, g.n2);
        const double kh2 = (double)(kx * kx + ky * ky);
        bin = (int)floor(sqrt(kh2) + 0.5);
        const size_t col = (size_t)iky * g.sp + kx;
        const double2 br = d.BR[(mz + z) * plane + col], bi = d.BI[(mz + z) * plane + col];
        const double lr = br.x - bi.y, lim = br.y + bi.x;  // BR + i BI
        k1 = lr * lr + lim * lim;
        // |C_k|^2 kh^2 / (Bu N^2_k), C_k = (A_{k+1} - A_k)/dz, 0 at the top interface
        double c[2] = {0.0, 0.0};
#pragma unroll
        for (int s = 0; s < 2; ++s) {
          const int k = z - 1 + s;
          if (k >= 0 && k < g.n3 - 1) {
            const size_t i0 = (mz + k) * plane + col, i1 = i0 + plane;
            const double2 a0 = d.AR[i0], a1 = d.AR[i1], b0 = d.AI[i0], b1 = d.AI[i1];
            const double2 cr = make_double2((a1.x - a0.x) / dz, (a1.y - a0.y) / dz);
            const double2 ci = make_double2((b1.x - b0.x) / dz, (b1.y - b0.y) / dz);
            c[s] = (1.0 / (Bu * d.N2u[k])) * kh2 * (cr.x * cr.x + cr.y * cr.y + ci.x * ci.x + ci.y * ci.y);
          }
        }
        p1 = 0.25 * (z == 0 ? c[1] : c[0] + c[1]);
      }
    }
    cb[t] = bin;
    cke[t] = k1;
    cpe[t] = p1;
    __syncthreads();
    for (int j = 0; j < NSP; ++j) {
      const int b = cb[j] - t;
      if (b == 0 || b == NSP) {
        const int h = b == 0 ? 0 : 1;
        ke[h] += cke[j];
        pe[h] += cpe[j];
        nm[h] += 2.0;
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const int b = t + h * NSP;
    if (b < nbin) {
      double* o = part + ((((size_t)m * L.n + li) * HCH + ch) * nbin + b) * 3;
      o[0] = ke[h];
      o[1] = pe[h];
      o[2] = nm[h];
    }
  }
}

// out[((m * nlev + li) * nbin + b) * 3 + j]: the chunks' partials summed in
// chunk order, KE and PE halved (diagnostics.f90:627)
__global__ __launch_bounds__(NT) void k_hspec_sum(int n, int nbin, const double* __restrict__ part,
                                                   double* __restrict__ out) {
  const int e = blockIdx.x * NT + threadIdx.x;  // (m * nlev + li) * nbin + b
  if (e >= n) return;
  const int b = e % nbin, ml = e / nbin;
  double s[3] = {0.0, 0.0, 0.0};
  for (int ch = 0; ch < HCH; ++ch) {
    const double* p = part + (((size_t)ml * HCH + ch) * nbin + b) * 3;
    s[0] += p[0];
    s[1] += p[1];
    s[2] += p[2];
  }
  out[(size_t)e * 3] = 0.5 * s[0];
  out[(size_t)e * 3 + 1] = 0.5 * s[1];
  out[(size_t)e * 3 + 2] = s[2];
}

// we_vave (diagnostics.f90:750-813): the vertically averaged wave energy
// density 0.5 (BR^2 + BI^2) / n3 per physical point (x, y) of each member,
// from the c2r fields br, bi ((z, y, x) order, member-major); out is
// (member, y, x).
__global__ __launch_bounds__(NT) void k_vave(Geo g, const double* __restrict__ br, const double* __restrict__ bi,
                                             double* __restrict__ out) {
  const size_t nxy = (size_t)g.n1 * g.n2;
  const size_t e = (size_t)blockIdx.x * NT + threadIdx.x;  // m * nxy + y * n1 + x
  if (e >= nxy * g.nm) return;
  const size_t m = e / nxy, xy = e % nxy;
  double s = 0.0;
  for (int z = 0; z < g.n3; ++z) {
    const size_t i = ((size_t)m * g.n3 + z) * nxy + xy;
    s += 0.5 * (br[i] * br[i] + bi[i] * bi[i]);
  }
  out[e] = s / g.n3;
}

// ---------------------------------------------------------------------------
// raw transforms (fft.f90:55-83) for initial conditions and tests
// ---------------------------------------------------------------------------
// r2c in x of pairs of real rows -> half rows (all kx), scaled
template <int N, int P>
__global__ __launch_bounds__(NT) void k_row_r2c(Geo g, const double2* __restrict__ tw, const double* __restrict__ r,
                                                 int ld, double2* __restrict__ out, double scale) {
  constexpr int S = P / 2, H = N / 2 + 1;
  __shared__ double2 lds[N * S];
  const size_t row0 = (size_t)blockIdx.x * P;
  for (int e = threadIdx.x; e < S * N; e += NT) {
    const int s = e / N, x = e % N;
    lds[x * S + s] = make_double2(r[(row0 + 2 * s) * ld + x], r[(row0 + 2 * s + 1) * ld + x]);
  }
  __syncthreads();
  fft_lds<N, false, S, NT>(lds, tw);
  for (int e = threadIdx.x; e < S * H; e += NT) {
    const int s = e / H, k = e % H;
    double2 A, B;
    split_pair<N>(lds, S, k, s, A, B);
    out[(row0 + 2 * s) * g.sp + k] = cscale(A, scale);
    out[(row0 + 2 * s + 1) * g.sp + k] = cscale(B, scale);
  }
}

// forward DFT in y of half rows -> spectral field (all rows, all kx)
// (in may equal F: every block reads its columns into LDS before it writes)
template <int N, int TX>
__global__ __launch_bounds__(NT) void k_col_fwd(Geo g, const double2* __restrict__ tw, const double2* in, double2* F) {
  __shared__ double2 lds[N * TX];
  const int kx0 = blockIdx.x * TX, z = blockIdx.y;
  const double2* iz = in + (size_t)z * g.n2 * g.sp;
  for (int e = threadIdx.x; e < N * TX; e += NT) {
    const int i = e / TX, c = e % TX, kx = kx0 + c;
    lds[i * TX + c] = kx < g.iktx ? iz[(size_t)i * g.sp + kx] : zc();
  }
  __syncthreads();
  fft_lds<N, false, TX, NT>(lds, tw);
  double2* Fz = F + (size_t)z * g.ikty * g.sp;
  for (int e = threadIdx.x; e < N * TX; e += NT) {
    const int i = e / TX, c = e % TX, kx = kx0 + c;
    if (kx < g.iktx) Fz[(size_t)i * g.sp + kx] = lds[i * TX + c];
  }
}

// c2r in x of pairs of half rows (all kx) -> real rows
template <int N, int P>
__global__ __launch_bounds__(NT) void k_row_c2r(Geo g, const double2* __restrict__ tw, const double2* __restrict__ in,
                                                 double* __restrict__ r, int ld) {
  constexpr int S = P / 2, H = N / 2 + 1;
  __shared__ double2 lds[N * S];
  const size_t row0 = (size_t)blockIdx.x * P;
  for (int e = threadIdx.x; e < S * H; e += NT) {
    const int s = e / H, k = e % H;
    put_pair<N>(lds, S, k, s, in[(row0 + 2 * s) * g.sp + k], in[(row0 + 2 * s + 1) * g.sp + k]);
  }
  __syncthreads();
  fft_lds<N, true, S, NT>(lds, tw);
  for (int e = threadIdx.x; e < S * N; e += NT) {
    const int s = e / N, x = e % N;
    const double2 v = lds[x * S + s];
    r[(row0 + 2 * s) * ld + x] = v.x;
    r[(row0 + 2 * s + 1) * ld + x] = v.y;
  }
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
template <int N>
struct L {
  using C = Cfg<N>;
  static hipError_t col_inv(const Dev& d, hipStream_t st, const double2* F, double2* out, int W, int box, int mul,
                            int mask, int nz) {
    dim3 grid((W + C::TXI - 1) / C::TXI, nz);
    hipLaunchKernelGGL((k_col_inv<N, C::TXI>), grid, dim3(NT), 0, st, d.g, d.tw, F, out, W, box, mul, mask);
    return hipGetLastError();
  }
  static hipError_t r2c(const Dev& d, hipStream_t st, const double* r, int ld, double2* k, int nz) {
    const double scale = 1.0 / ((double)d.g.n1 * (double)d.g.n2);
    hipLaunchKernelGGL((k_row_r2c<N, C::PRC>), dim3(d.g.n2 * nz / C::PRC), dim3(NT), 0, st, d.g, d.tw, r, ld,
                       d.qs, scale);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL((k_col_fwd<N, C::TXI>), dim3((d.g.iktx + C::TXI - 1) / C::TXI, nz), dim3(NT), 0, st, d.g,
                       d.tw, d.qs, k);
    return hipGetLastError();
  }
  // forward y transform of nz (kx, y) planes in place, all kx < iktx
  static hipError_t yfwd(const Dev& d, hipStream_t st, double2* P) {
    hipLaunchKernelGGL((k_col_fwd<N, C::TXI>), dim3((d.g.iktx + C::TXI - 1) / C::TXI, d.g.nz), dim3(NT), 0, st, d.g,
                       d.tw, P, P);
    return hipGetLastError();
  }
  static hipError_t c2r(const Dev& d, hipStream_t st, const double2* k, double* r, int ld, int nz) {
    hipError_t e = col_inv(d, st, k, d.qs, d.g.iktx, 0, 0, 0, nz);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL((k_row_c2r<N, C::PRC>), dim3(d.g.n2 * nz / C::PRC), dim3(NT), 0, st, d.g, d.tw, d.qs, r, ld);
    return hipGetLastError();
  }
};

#define QG_DISPATCH(n, call)                 \
  switch (n) {                               \
    case 16: return L<16>::call;             \
    case 32: return L<32>::call;             \
    case 64: return L<64>::call;             \
    case 128: return L<128>::call;           \
    case 256: return L<256>::call;           \
    case 512: return L<512>::call;           \
    default: return hipErrorInvalidValue;    \
  }

bool supported_n(int n) { return n == 16 || n == 32 || n == 64 || n == 128 || n == 256 || n == 512; }

hipError_t launch_vave(const Dev& d, hipStream_t st, double* br, double* bi, double* out) {
  hipError_t e = launch_c2r(d, st, d.BR, br, d.g.n1, d.g.nz);
  if (e == hipSuccess) e = launch_c2r(d, st, d.BI, bi, d.g.n1, d.g.nz);
  if (e != hipSuccess) return e;
  const size_t n = (size_t)d.g.n1 * d.g.n2 * d.g.nm;
  hipLaunchKernelGGL(k_vave, dim3((unsigned)((n + NT - 1) / NT)), dim3(NT), 0, st, d.g, br, bi, out);
  return hipGetLastError();
}
hipError_t launch_r2c(const Dev& d, hipStream_t st, const double* r, int ld, double2* k, int nz) {
  QG_DISPATCH(d.g.n1, r2c(d, st, r, ld, k, nz))
}
hipError_t launch_c2r(const Dev& d, hipStream_t st, const double2* k, double* r, int ld, int nz) {
  QG_DISPATCH(d.g.n1, c2r(d, st, k, r, ld, nz))
}

template <int M, int P, int LPC>
static hipError_t solve_c(const Dev& d, hipStream_t st, int f_lo, int nf, const double2* rhs_psi) {
  const int ncols = d.g.KX * d.g.nbox;
  dim3 grid((ncols + LPC - 1) / LPC, nf * d.g.nm);
  hipLaunchKernelGGL((k_solve_c<M, P, LPC>), grid, dim3(LPC * P), 0, st, d, f_lo, rhs_psi);
  return hipGetLastError();
}

template <int M, int LPC>
static hipError_t solve_m(const Dev& d, hipStream_t st, int f_lo, int nf, const double2* rhs_psi, int P) {
  switch (P) {
    case 1: return solve_c<M, 1, LPC>(d, st, f_lo, nf, rhs_psi);
    case 2: return solve_c<M, 2, LPC>(d, st, f_lo, nf, rhs_psi);
    case 4: return solve_c<M, 4, LPC>(d, st, f_lo, nf, rhs_psi);
    case 8: return solve_c<M, 8, LPC>(d, st, f_lo, nf, rhs_psi);
    case 16: return solve_c<M, 16, LPC>(d, st, f_lo, nf, rhs_psi);
    default: return solve_c<M, 32, LPC>(d, st, f_lo, nf, rhs_psi);
  }
}

// Shape of the partitioned solve for a solve variant v: M levels per chunk
// and LPC columns per block; P = the smallest power of two with M * P >= n3.
void solve_shape(int n3, int v, int* M, int* P, int* LPC) {
  static const int tm[4] = {8, 8, 16, 16}, tl[4] = {16, 32, 16, 32};
  v = (v < 0 || v > 3) ? 0 : v;
  int m = tm[v];
  int p = 1;
  while (m * p < n3) p *= 2;
  // at most 512 threads and 32 KiB of LDS per block when the levels allow
  while (m < 16 && (p > 32 || p * tl[v] > 512)) {
    m *= 2;
    p = 1;
    while (m * p < n3) p *= 2;
  }
  *M = m;
  *P = p;
  *LPC = p * tl[v] > 512 ? 16 : tl[v];
}

hipError_t launch_solve(const Dev& d, hipStream_t st, int f_lo, int f_hi, const double2* rhs_psi) {
  if (f_hi < f_lo) return hipSuccess;
  const int nf = f_hi - f_lo + 1;
  int M, P, LPC;
  solve_shape(d.g.n3, d.tune.solve, &M, &P, &LPC);
  if (P <= 32) {
    if (M == 8) return LPC == 16 ? solve_m<8, 16>(d, st, f_lo, nf, rhs_psi, P)
                                 : solve_m<8, 32>(d, st, f_lo, nf, rhs_psi, P);
    return LPC == 16 ? solve_m<16, 16>(d, st, f_lo, nf, rhs_psi, P) : solve_m<16, 32>(d, st, f_lo, nf, rhs_psi, P);
  }
  // more than 512 levels: one sequential Thomas sweep per column
  const int ncols = d.g.KX * d.g.nbox;
  dim3 grid((2 * ncols + NT - 1) / NT, nf * d.g.nm);
  hipLaunchKernelGGL(k_solve, grid, dim3(NT), 0, st, d, f_lo, rhs_psi);
  return hipGetLastError();
}
hipError_t launch_dq(const Dev& d, hipStream_t st) {
  const int n = d.g.KX * d.g.nbox * d.g.nz;
  hipLaunchKernelGGL(k_dq, dim3((n + NT - 1) / NT), dim3(NT), 0, st, d);
  return hipGetLastError();
}
hipError_t launch_clear_outside(const Dev& d, hipStream_t st, double2* f, int all) {
  const size_t n = (size_t)d.g.ikty * d.g.sp * d.g.nz;
  hipLaunchKernelGGL(k_clear, dim3((unsigned)((n + NT - 1) / NT)), dim3(NT), 0, st, d.g, f, all);
  return hipGetLastError();
}
hipError_t launch_init_q(const Dev& d, hipStream_t st, const double2* psi, double2* q) {
  const size_t n = (size_t)d.g.ikty * d.g.sp * d.g.nz;
  hipLaunchKernelGGL(k_init_q, dim3((unsigned)((n + NT - 1) / NT)), dim3(NT), 0, st, d, psi, q);
  return hipGetLastError();
}
hipError_t launch_cz(const Dev& d, hipStream_t st, const double2* A, double2* C) {
  const size_t n = (size_t)d.g.ikty * d.g.sp * d.g.nz;
  hipLaunchKernelGGL(k_cz, dim3((unsigned)((n + NT - 1) / NT)), dim3(NT), 0, st, d, A, C);
  return hipGetLastError();
}
hipError_t launch_tend_split(const Dev& d, hipStream_t st) {
  if (!d.Z[0] || !d.Z[1] || !d.tR || !d.tI) return hipErrorInvalidValue;
  for (int i = 0; i < 2; ++i) {
    hipError_t e = hipSuccess;
    switch (d.g.n2) {
      case 16: e = L<16>::yfwd(d, st, d.Z[i]); break;
      case 32: e = L<32>::yfwd(d, st, d.Z[i]); break;
      case 64: e = L<64>::yfwd(d, st, d.Z[i]); break;
      case 128: e = L<128>::yfwd(d, st, d.Z[i]); break;
      case 256: e = L<256>::yfwd(d, st, d.Z[i]); break;
      case 512: e = L<512>::yfwd(d, st, d.Z[i]); break;
      default: return hipErrorInvalidValue;
    }
    if (e != hipSuccess) return e;
  }
  const size_t n = (size_t)d.g.nz * d.g.ikty * d.g.iktx;
  hipLaunchKernelGGL(k_tend_split, dim3((unsigned)((n + NT - 1) / NT)), dim3(NT), 0, st, d);
  return hipGetLastError();
}
int diag_blocks(const Geo& g) { return (g.KX * g.nbox * g.n3 + NT - 1) / NT; }
hipError_t launch_conv(const Dev& d, hipStream_t st, int* nblocks) {
  if (!d.Z[0] || !d.Z[1] || !d.tR || !d.tI || !d.cw) return hipErrorInvalidValue;
  const int nb = diag_blocks(d.g);
  *nblocks = nb;
  hipLaunchKernelGGL(k_conv, dim3(nb, d.g.nm), dim3(NT), 0, st, d);
  return hipGetLastError();
}
hipError_t launch_diag(const Dev& d, hipStream_t st, int* nblocks) {
  const int nb = diag_blocks(d.g);
  *nblocks = nb;
  hipLaunchKernelGGL(k_diag, dim3(nb, d.g.nm), dim3(NT), 0, st, d);
  return hipGetLastError();
}

int hspec_chunks() { return HCH; }
hipError_t launch_hspec_waves(const Dev& d, hipStream_t st, const SpecLevels& L, int nbin, double* out,
                              double* part) {
  if (L.n < 1 || L.n > SpecLevels::MAXL || nbin < 1 || nbin > 2 * NSP) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_hspec_waves, dim3(L.n, d.g.nm, HCH), dim3(NSP), 0, st, d, L, nbin, part);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  const int n = d.g.nm * L.n * nbin;
  hipLaunchKernelGGL(k_hspec_sum, dim3((n + NT - 1) / NT), dim3(NT), 0, st, n, nbin, part, out);
  return hipGetLastError();
}

const char* kernel_name(int id) {
  static const char* names[K_NKERNELS] = {"col_inv", "row_nl", "col_upd", "row_qw", "col_qw", "solve", "solve_a", "aux"};
  return (id >= 0 && id < K_NKERNELS) ? names[id] : "?";
}

}  // namespace qg
