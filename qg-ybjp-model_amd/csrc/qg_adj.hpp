// Pointwise kernels of the tangent-linear and adjoint steps (qg_adj_kernels.hip).
// Internal; not part of the public interface (include/qgybj.h).
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>

#include "qg_common.hpp"

namespace qg {

enum : int { TAB_NONE = 0, TAB_E1Q, TAB_E2Q, TAB_E1W, TAB_E2W };

// one term of a spectral combination: coefficient
//   a * (c0 + ckx kx + cky ky + ckh kh2) * (i if imag) * W^wp * table(kx, ky)
struct SMulTerm {
  const double2* in = nullptr;
  double a = 1.0, c0 = 1.0, ckx = 0.0, cky = 0.0, ckh = 0.0;
  int imag = 0, wp = 0, tab = TAB_NONE;
};
constexpr int SMUL_TERMS = 4;
// out = (accumulate ? out : 0) + sum of the terms, at L = 1 modes of levels
// [zlo, zhi) of every member; elsewhere out = 0 unless accumulate (then
// unchanged, or 0 with clear_off)
struct SMulArgs {
  SMulTerm t[SMUL_TERMS];
  int zlo = 0, zhi = 1 << 30;
  int accumulate = 0, clear_off = 0;
};

constexpr int PFMA_TERMS = 4;
// out = (accumulate ? out : 0) + sum_j s_j * x_j * (y_j or 1)
struct PFmaArgs {
  const double* x[PFMA_TERMS] = {nullptr, nullptr, nullptr, nullptr};
  const double* y[PFMA_TERMS] = {nullptr, nullptr, nullptr, nullptr};
  double s[PFMA_TERMS] = {1.0, 1.0, 1.0, 1.0};
  int accumulate = 0;
};

// buffers of the fused adjoint step (qg_adj_step.hip): forward and adjoint
// half-transformed (kx, y) planes, the state adjoint and the inversion outputs
struct AdjBufs {
  double2* F[6];
  double2* A[6];
  double2 *psi, *q, *qo, *BR, *BI, *BRo, *BIo, *AR, *AI;
  const double2* qstar;  // S_psi(psibar')
  const double2* sR;     // S_A(ARbar')
  const double2* sI;     // S_A(AIbar')
};
hipError_t launch_adj_col_qw(const Dev& d, hipStream_t st, const AdjBufs& b);
hipError_t launch_adj_row_qw(const Dev& d, hipStream_t st, const AdjBufs& b);
hipError_t launch_adj_col_upd(const Dev& d, hipStream_t st, const AdjBufs& b, int euler, int feedback);
hipError_t launch_adj_row_nl(const Dev& d, hipStream_t st, const AdjBufs& b, int Win, int euler);
hipError_t launch_adj_col_out(const Dev& d, hipStream_t st, const AdjBufs& b);

struct CopyBoxArgs {
  const double2* src[4];
  double2* dst[4];
  int n;
};
hipError_t launch_copy_box(const Geo& g, hipStream_t st, const CopyBoxArgs& a);

hipError_t launch_smul(const Dev& d, hipStream_t st, double2* out, const SMulArgs& a);
hipError_t launch_pfma(hipStream_t st, double* out, size_t n, const PFmaArgs& a);
int obj_blocks(const Geo& g);
hipError_t launch_obj(const Dev& d, hipStream_t st, const double2* AR, const double2* AI, int zlo, int zhi,
                      double* part);

}  // namespace qg
