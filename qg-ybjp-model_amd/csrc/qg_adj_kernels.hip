// Generic pointwise kernels of the tangent-linear and adjoint steps.
//
// The linearised step is assembled on the host (qg_adj.cpp) from these
// kernels, the raw transforms (launch_r2c / launch_c2r), the vertical solves
// (launch_solve) and the init_q operator (launch_init_q):
//  - k_smul: spectral linear combination of up to four fields with per-mode
//    coefficients a * (c0 + ckx kx + cky ky + ckh kh2) * (i if imag) *
//    W^wp * table(kx, ky), restricted to L = 1 modes (and optionally to a
//    range of levels); W is the Hermitian weight of a half-spectrum mode
//    (2, and 1 at kx = ky = 0)
//  - k_pfma: physical multiply-add of up to four products
//  - k_obj: partial sums of sum W kh2 (|AR|^2 + |AI|^2) over a level range
//  - k_copy_box: copy of the dealiasing box of up to four fields
#include <hip/hip_runtime.h>

#include "qg_adj.hpp"
#include "qg_common.hpp"
#include "qg_fft.hpp"

namespace qg {

namespace {

constexpr int NTA = 256;

__device__ __forceinline__ double2 zc() { return make_double2(0.0, 0.0); }

__device__ __forceinline__ const double* table_ptr(const Dev& d, int t) {
  switch (t) {
    case TAB_E1Q: return d.e1q;
    case TAB_E2Q: return d.e2q;
    case TAB_E1W: return d.e1w;
    case TAB_E2W: return d.e2w;
    default: return nullptr;
  }
}

__global__ __launch_bounds__(NTA) void k_smul(Dev d, double2* __restrict__ out, SMulArgs a) {
  const Geo& g = d.g;
  const size_t plane = (size_t)g.ikty * g.sp;
  const size_t e = (size_t)blockIdx.x * NTA + threadIdx.x;
  if (e >= plane * g.nz) return;
  const size_t r = e % plane;
  const int iky = (int)(r / g.sp), kx = (int)(r % g.sp);
  if (kx >= g.iktx) return;
  const int z = (int)(e / plane) % g.n3;
  const bool on = lmask(kx, iky, g) && z >= a.zlo && z < a.zhi;
  double2 acc = a.accumulate ? out[e] : zc();
  if (on) {
    const int ky = ky_of(iky, g.n2);
    const double kh2 = (double)(kx * kx + ky * ky);
    const double w = (kx == 0 && ky == 0) ? 1.0 : 2.0;
#pragma unroll
    for (int j = 0; j < SMUL_TERMS; ++j) {
      const SMulTerm& t = a.t[j];
      if (!t.in) continue;
      double c = t.a * (((t.c0 + t.ckx * kx) + t.cky * ky) + t.ckh * kh2);
      if (t.wp > 0) c *= w;
      if (t.wp < 0) c /= w;
      const double* tab = table_ptr(d, t.tab);
      if (tab) c *= tab[(size_t)iky * g.iktx + kx];
      const double2 v = t.in[e];
      if (t.imag) {
        acc.x -= c * v.y;
        acc.y += c * v.x;
      } else {
        acc.x += c * v.x;
        acc.y += c * v.y;
      }
    }
  } else if (!a.accumulate || a.clear_off) {
    acc = zc();
  }
  out[e] = acc;
}

__global__ __launch_bounds__(NTA) void k_pfma(double* __restrict__ out, size_t n, PFmaArgs a) {
  const size_t i = (size_t)blockIdx.x * NTA + threadIdx.x;
  if (i >= n) return;
  double acc = a.accumulate ? out[i] : 0.0;
#pragma unroll
  for (int j = 0; j < PFMA_TERMS; ++j) {
    if (!a.x[j]) continue;
    const double v = a.y[j] ? a.x[j][i] * a.y[j][i] : a.x[j][i];
    acc += a.s[j] * v;
  }
  out[i] = acc;
}

// sum over L modes of levels [zlo, zhi) of every member of W kh2 (|AR|^2 + |AI|^2)
__global__ __launch_bounds__(NTA) void k_obj(Dev d, const double2* __restrict__ AR, const double2* __restrict__ AI,
                                             int zlo, int zhi, double* __restrict__ part) {
  const Geo& g = d.g;
  const size_t plane = (size_t)g.ikty * g.sp;
  const size_t e = (size_t)blockIdx.x * NTA + threadIdx.x;
  double acc = 0.0;
  if (e < plane * g.nz) {
    const size_t r = e % plane;
    const int iky = (int)(r / g.sp), kx = (int)(r % g.sp);
    const int z = (int)(e / plane) % g.n3;
    if (kx < g.iktx && z >= zlo && z < zhi && lmask(kx, iky, g)) {
      const int ky = ky_of(iky, g.n2);
      const double kh2 = (double)(kx * kx + ky * ky);
      const double w = (kx == 0 && ky == 0) ? 1.0 : 2.0;
      const double2 ar = AR[e], ai = AI[e];
      acc = w * kh2 * ((ar.x * ar.x + ar.y * ar.y) + (ai.x * ai.x + ai.y * ai.y));
    }
  }
  __shared__ double red[NTA];
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int s = NTA / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// copy the dealiasing box of up to four fields (every member level)
__global__ __launch_bounds__(NTA) void k_copy_box(Geo g, CopyBoxArgs a) {
  const size_t per = (size_t)g.KX * g.nbox;
  const size_t e = (size_t)blockIdx.x * NTA + threadIdx.x;
  if (e >= per * g.nz) return;
  const size_t z = e / per, r = e % per;
  const int kx = (int)(r % g.KX), iky = box_row((int)(r / g.KX), g);
  const size_t idx = z * g.ikty * g.sp + (size_t)iky * g.sp + kx;
#pragma unroll
  for (int j = 0; j < 4; ++j)
    if (j < a.n) a.dst[j][idx] = a.src[j][idx];
}

}  // namespace

hipError_t launch_copy_box(const Geo& g, hipStream_t st, const CopyBoxArgs& a) {
  const size_t n = (size_t)g.KX * g.nbox * g.nz;
  hipLaunchKernelGGL(k_copy_box, dim3((unsigned)((n + NTA - 1) / NTA)), dim3(NTA), 0, st, g, a);
  return hipGetLastError();
}

hipError_t launch_smul(const Dev& d, hipStream_t st, double2* out, const SMulArgs& a) {
  const size_t n = (size_t)d.g.ikty * d.g.sp * d.g.nz;
  hipLaunchKernelGGL(k_smul, dim3((unsigned)((n + NTA - 1) / NTA)), dim3(NTA), 0, st, d, out, a);
  return hipGetLastError();
}

hipError_t launch_pfma(hipStream_t st, double* out, size_t n, const PFmaArgs& a) {
  hipLaunchKernelGGL(k_pfma, dim3((unsigned)((n + NTA - 1) / NTA)), dim3(NTA), 0, st, out, n, a);
  return hipGetLastError();
}

int obj_blocks(const Geo& g) { return (int)(((size_t)g.ikty * g.sp * g.nz + NTA - 1) / NTA); }

hipError_t launch_obj(const Dev& d, hipStream_t st, const double2* AR, const double2* AI, int zlo, int zhi,
                      double* part) {
  hipLaunchKernelGGL(k_obj, dim3((unsigned)obj_blocks(d.g)), dim3(NTA), 0, st, d, AR, AI, zlo, zhi, part);
  return hipGetLastError();
}

}  // namespace qg
