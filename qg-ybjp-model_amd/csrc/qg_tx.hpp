// Helpers shared by the register-FFT row and column kernels of the step
// (qg_step_kernels.hip) and of its adjoint (qg_adj_step.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "qg_fft.hpp"

namespace qg {

__device__ __forceinline__ double2 zero2() { return make_double2(0.0, 0.0); }

// Streaming store of an intermediate (kx, y) plane, which only the next
// transform kernel reads: the non-temporal hint writes it through instead of
// leaving dirty lines to be written back while the next kernel streams.  A
// memory-only kernel with the row kernels' access pattern goes from 4.3 to
// 5.9 TB/s with it; on the whole step (every plane store non-temporal, the
// state fields and solve outputs not, which measured slower) 0.495 ->
// 0.479 ms at 256^2 x 128.
__device__ __forceinline__ void st_nt(double2* p, double2 v) {
  __builtin_nontemporal_store(v.x, &p->x);
  __builtin_nontemporal_store(v.y, &p->y);
}
// Non-temporal load of a state entry that this kernel reads once and that is
// next read a whole step later (the update's q, B, A fields): 165 -> 160 us
// for the update kernel at 256^2 x 128.  The intermediate planes, which the
// next kernel reads, are loaded normally (non-temporal loads of those
// measured slower).
__device__ __forceinline__ double2 ld_nt(const double2* p) {
  return make_double2(__builtin_nontemporal_load(&p->x), __builtin_nontemporal_load(&p->y));
}
__device__ __forceinline__ double2 mul_ik(double2 v, double k) { return make_double2(-k * v.y, k * v.x); }

// Pack entry k of two real half-spectra A, B into position pos of a complex
// sequence whose inverse DFT is a + i b: Z_pos = A_k + i B_k for pos <= N/2,
// conj(A_k) + i conj(B_k) for pos > N/2 (k = N - pos); the imaginary parts of
// the k = 0 and k = N/2 entries are dropped (the c2r convention of FFTW and
// pocketfft).
template <int N>
__device__ __forceinline__ double2 pack2(double2 A, double2 B, int pos) {
  if (pos == 0 || pos == N / 2) return make_double2(A.x, B.x);
  if (pos < N / 2) return make_double2(A.x - B.y, A.y + B.x);
  return make_double2(A.x + B.y, B.x - A.y);
}

// Split Z = DFT(a + i b) at k into A = DFT(a), B = DFT(b), given Z_k and
// Z_{N-k}.
__device__ __forceinline__ void split2(double2 Z, double2 Zr, double2& A, double2& B) {
  const double2 Zm = cconj(Zr);
  A = make_double2((Z.x + Zm.x) * 0.5, (Z.y + Zm.y) * 0.5);
  B = make_double2((Z.y - Zm.y) * 0.5, -(Z.x - Zm.x) * 0.5);
}

// Workgroup shapes for a transform length N with E elements per thread:
// T threads per sequence; column kernels take TX consecutive kx columns
// (at least 8, i.e. one 128-byte line per row); row kernels take RP rows
// with RP * T threads, RP = WG / T for a target of WG threads (at least one
// row, at most N rows so that RP divides n2).
template <int N, int E>
struct SH {
  static constexpr int T = N / E;
  static constexpr int TX = (64 / T) > 8 ? 64 / T : 8;
  static constexpr int COLT = TX * T;  // threads of a column workgroup
  template <int WG>
  static constexpr int rp() { return WG / T < 1 ? 1 : (WG / T < N ? WG / T : N); }
};

}  // namespace qg
