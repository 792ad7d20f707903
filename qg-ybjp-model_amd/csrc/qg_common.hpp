// Internal declarations shared by the kernels (qg_kernels.hip) and the C-ABI
// layer (qg_api.cpp).  Not part of the public interface (include/qgybj.h).
#pragma once
#include <hip/hip_runtime.h>
#include <cstddef>
#include <cstdint>

namespace qg {

// Spectral-grid geometry of one member.  Spectral fields use the reference's
// Fortran layout (iktx, ikty, n3), kx fastest (main_waqg.f90:15-59).  The hot
// loop only touches the dealiasing box: kx < KX and rows iky < nlo or
// iky >= n2 - nhi, which contains every mode with L = 1 (init.f90:44-80).
struct Geo {
  int n1, n2, n3;
  int nm;      // members stacked in one context (same grid and parameters)
  int nz;      // nm * n3: levels of every device field, member-major
               // (member m, level k at plane m * n3 + k); the horizontal
               // kernels treat all nz planes alike, the vertical ones work
               // per member
  int iktx, ikty;
  int sp;      // device row pitch in complex elements: iktx rounded up to a
               // multiple of 8 so that every 8-column tile is one 128-byte line
  int kmax;    // largest |k| component with L = 1
  int KX;      // kmax + 1 columns kept
  int nlo;     // rows 0..kmax     (ky = 0..kmax)
  int nhi;     // rows n2-kmax..n2-1 (ky = -kmax..-1)
  int nbox;    // nlo + nhi
  int kh2max;  // L = 1 requires kx^2 + ky^2 <= kh2max
};

// ky label of row iky: 0..n2/2-1, the Nyquist row relabelled 0, then
// -n2/2+1..-1 (init.f90:35-41)
__host__ __device__ inline int ky_of(int iky, int n2) {
  return iky < n2 / 2 ? iky : (iky == n2 / 2 ? 0 : iky - n2);
}

// Dealiasing mask L (init.f90:44-80): zero on the Nyquist row, on kx = 0 with
// ky < 0, and outside the radius int(n1/3+0.5)-0.5 (an integer bound on kh2).
__host__ __device__ inline bool lmask(int kx, int iky, const Geo& g) {
  if (iky == g.n2 / 2) return false;
  const int ky = ky_of(iky, g.n2);
  if (kx == 0 && ky < 0) return false;
  return kx * kx + ky * ky <= g.kh2max;
}

__host__ __device__ inline bool in_box_row(int iky, const Geo& g) {
  return iky < g.nlo || iky >= g.n2 - g.nhi;
}
__host__ __device__ inline int box_row(int ii, const Geo& g) {
  return ii < g.nlo ? ii : ii + (g.n2 - g.nlo - g.nhi);
}

Geo make_geo(int n1, int n2, int n3, int nm = 1);

// flag bits, same values as include/qgybj.h
enum : int {
  F_FIXED_FLOW = 1, F_NO_FEEDBACK = 2, F_PASSIVE = 4, F_LINEAR = 8, F_NO_DISPERSION = 16, F_INVISCID = 32,
  F_YBJ_REGULAR = 64, F_ALL = 127
};

// Kernel variant choice per hot kernel (ps_inv, row_nl, col_upd, row_qw,
// col_qw).  Column kernels: 1 = 16 elements per thread, 0 = 8.  Row kernels:
// 0 = 8 elements / 256 threads, 1 = 16 / 256, 2 = 8 / 64, 3 = 16 / 64.
// solve: shape of the partitioned vertical solve (solve_shape).  a_main: 1 =
// the A inversions run with the psi inversion in one launch on the member's
// stream, 0 = on a side stream overlapped with the q^w work.
struct Tune {
  int e16[5];
  int solve;
  int a_main;
};

// Device-side view of one member: state, work buffers and tables.
struct Dev {
  Geo g;
  Tune tune;
  int flags;
  double dt, gamma, Ro, Bu, W2F, nuz, dz;
  // state ((iktx, ikty, n3) with row pitch sp)
  double2 *psi, *q, *qo, *BR, *BI, *BRo, *BIo, *AR, *AI;
  // work: q* (and temporaries), dq, intermediates I0..I5 (pitch sp, n2 rows/level)
  double2 *qs, *dq;
  double2* I[6];
  // regular YBJ only: the update's combined B tendencies nBR - rBI/2 and
  // nBI + rBR/2 (box), summed over z for sigma; null otherwise
  double2 *tR, *tI;
  // kept B tendencies (qg_keep_tendencies), null otherwise: row_nl writes
  // F(zeta BR), F(zeta BI) as (kx, y) planes into Z, which the tendency split
  // turns into rBR, rBI in place; tR, tI then become nBR, nBI
  double2* Z[2];
  double* scratch;  // Thomas elimination diagonals, 3 * n3 * ncols * 2
  // tables
  const double2* tw;       // TWN twiddles
  const double *e1q, *e2q, *e1w, *e2w;  // exp(-f), exp(-2f) per (iky, kx)
  const double* cw;                     // wave dissipation coefficient fw / dt per (iky, kx)
  const double *dl, *du, *dsum;         // vertical operator, n3 each
  const double *a_ell, *N2u, *N2s;      // n3 each
  double* diag_part;                    // per-block partial sums
  int* flag;                            // non-finite detector
};

// kernel ids for per-kernel timing
enum KernelId : int {
  K_COL_INV = 0, K_ROW_NL, K_COL_UPD, K_ROW_QW, K_COL_QW, K_SOLVE, K_SOLVE_A, K_AUX, K_NKERNELS
};
const char* kernel_name(int id);

// ---- launchers (qg_kernels.hip); all enqueue on `st` and return hipError_t
// inverse y transform of one spectral field into up to three intermediates
// with the multipliers 1 (o0), i*ky (o1), -ky^2 (o2); o1/o2 may be null;
// W columns; box rows only if box; mask applies L
hipError_t launch_col_inv(const Dev& d, hipStream_t st, const double2* F, double2* o0, double2* o1, double2* o2,
                          int W, int box, int mask, int nz);
// several independent inverses in one launch (one job per input field; a
// job's nz planes are consecutive in the input and in each output)
struct ColInvJob {
  const double2* F;
  double2* o[3];
  int W, box, mask, nz;
};
struct ColInvJobs {
  static constexpr int MAXJ = 4;
  ColInvJob j[MAXJ];
  int n;
};
hipError_t launch_col_inv(const Dev& d, hipStream_t st, const ColInvJobs& jobs);
hipError_t launch_row_nl(const Dev& d, hipStream_t st, int Win);
hipError_t launch_col_upd(const Dev& d, hipStream_t st, int euler, int has_dq);
hipError_t launch_row_qw(const Dev& d, hipStream_t st);
hipError_t launch_col_qw(const Dev& d, hipStream_t st);
hipError_t launch_solve(const Dev& d, hipStream_t st, int f_lo, int f_hi, const double2* rhs_psi);
void solve_shape(int n3, int v, int* M, int* P, int* LPC);
hipError_t launch_dq(const Dev& d, hipStream_t st);
hipError_t launch_clear_outside(const Dev& d, hipStream_t st, double2* f, int all);
hipError_t launch_init_q(const Dev& d, hipStream_t st, const double2* psi, double2* q);
hipError_t launch_cz(const Dev& d, hipStream_t st, const double2* A, double2* C);
hipError_t launch_diag(const Dev& d, hipStream_t st, int* nblocks);
// omega equation: w (spectral, all levels) from psi, through four spectral
// and four physical work fields (sp: ikty * sp * nz, ph: n1 * n2 * nz each)
hipError_t launch_omega(const Dev& d, hipStream_t st, double2* const sp[4], double* const ph[4], double2* w);
// regular YBJ: sumB, sigma and compute_A in one pass per (kx, ky) column
hipError_t launch_ybj_regular(const Dev& d, hipStream_t st);
// kept tendencies: y transform of the Z planes and the split of the update's
// combined tendencies into nBR, nBI (in tR, tI) and rBR, rBI (in Z), masked
hipError_t launch_tend_split(const Dev& d, hipStream_t st);
// wave energy conversion sums (one partial per block and member in
// diag_part, 3 values, see k_conv)
hipError_t launch_conv(const Dev& d, hipStream_t st, int* nblocks);
// we_vave: c2r of BR, BI into br, bi (n1 * n2 * nz each), then the vertical
// average of 0.5 (BR^2 + BI^2) per (member, y, x) into out
hipError_t launch_vave(const Dev& d, hipStream_t st, double* br, double* bi, double* out);
// levels (1-based, as the reference's height array) of one spectra launch
struct SpecLevels {
  static constexpr int MAXL = 32;
  int n;
  int lev[MAXL];
};
// out[((member * L.n + level) * nbin + shell) * 3 + {KE, PE, modes}]
// chunks of the box per level: part holds nm * L.n * hspec_chunks() * nbin * 3
int hspec_chunks();
hipError_t launch_hspec_waves(const Dev& d, hipStream_t st, const SpecLevels& L, int nbin, double* out,
                              double* part);
hipError_t launch_r2c(const Dev& d, hipStream_t st, const double* r, int ld, double2* k, int nz);
hipError_t launch_c2r(const Dev& d, hipStream_t st, const double2* k, double* r, int ld, int nz);
int diag_blocks(const Geo& g);
bool supported_n(int n);

}  // namespace qg
