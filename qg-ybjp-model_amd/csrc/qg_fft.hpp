// Batched FP64 complex FFTs staged in LDS for gfx950.
//
// One workgroup transforms S independent length-N sequences that sit in LDS
// interleaved as lds[i*S + s] (element i of sequence s), so that consecutive
// threads touch consecutive 16-byte words.  Each pass is a Stockham
// autosort radix-R step (natural order in, natural order out, no bit
// reversal) following the GPU formulation of Govindaraju et al., "High
// Performance Discrete Fourier Transforms on Graphics Processors" (SC'08):
// a thread loads R elements at stride N/R, applies the twiddles
// W_{Ns*R}^{(j mod Ns) r}, does an R-point DFT in registers and stores at
// stride Ns.  All reads of a pass complete (barrier) before its writes, so a
// single LDS buffer is used in place.
//
// Twiddles come from one table tw[m] = exp(-2*pi*i*m/TWN) computed on the
// host in extended precision (correctly rounded to FP64); the forward
// transform uses exp(-i...), the inverse its conjugate and no scaling.
#pragma once
#include <hip/hip_runtime.h>

namespace qg {

constexpr int TWN = 1024;  // twiddle table length (max supported N)

__device__ __forceinline__ double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 cconj(double2 a) { return make_double2(a.x, -a.y); }
__device__ __forceinline__ double2 cscale(double2 a, double s) { return make_double2(a.x * s, a.y * s); }
// i*a and -i*a
__device__ __forceinline__ double2 cmul_i(double2 a) { return make_double2(-a.y, a.x); }
__device__ __forceinline__ double2 cmul_mi(double2 a) { return make_double2(a.y, -a.x); }

// multiply by W4 = exp(-+ i pi/2): -i for the forward, +i for the inverse
template <bool INV>
__device__ __forceinline__ double2 w4(double2 a) { return INV ? cmul_i(a) : cmul_mi(a); }

// cos(pi/8), sin(pi/8), sqrt(1/2) rounded to FP64
constexpr double C8 = 0.92387953251128675613;
constexpr double S8 = 0.38268343236508977173;
constexpr double R2 = 0.70710678118654752440;

// multiply by W_M^m = exp(-2 pi i m/M) (forward) or its conjugate (inverse),
// for the constant angles used inside the radix-8/16 kernels
template <bool INV>
__device__ __forceinline__ double2 wconst(double2 a, double c, double s) {
  // forward: (c - i s); inverse: (c + i s)
  return INV ? make_double2(a.x * c - a.y * s, a.x * s + a.y * c)
             : make_double2(a.x * c + a.y * s, a.y * c - a.x * s);
}

template <bool INV>
__device__ __forceinline__ void dft2(double2& a0, double2& a1) {
  double2 t = a0;
  a0 = cadd(t, a1);
  a1 = csub(t, a1);
}

template <bool INV>
__device__ __forceinline__ void dft4(double2& a0, double2& a1, double2& a2, double2& a3) {
  double2 t0 = cadd(a0, a2), t1 = csub(a0, a2);
  double2 t2 = cadd(a1, a3), t3 = w4<INV>(csub(a1, a3));
  a0 = cadd(t0, t2);
  a2 = csub(t0, t2);
  a1 = cadd(t1, t3);
  a3 = csub(t1, t3);
}

template <bool INV>
__device__ __forceinline__ void dft8(double2* v) {
  // even / odd halves
  double2 e0 = v[0], e1 = v[2], e2 = v[4], e3 = v[6];
  double2 o0 = v[1], o1 = v[3], o2 = v[5], o3 = v[7];
  dft4<INV>(e0, e1, e2, e3);
  dft4<INV>(o0, o1, o2, o3);
  o1 = wconst<INV>(o1, R2, R2);   // W8^1
  o2 = w4<INV>(o2);               // W8^2
  o3 = wconst<INV>(o3, -R2, R2);  // W8^3 = cos(3pi/4) - i sin(3pi/4)
  v[0] = cadd(e0, o0); v[4] = csub(e0, o0);
  v[1] = cadd(e1, o1); v[5] = csub(e1, o1);
  v[2] = cadd(e2, o2); v[6] = csub(e2, o2);
  v[3] = cadd(e3, o3); v[7] = csub(e3, o3);
}

template <bool INV>
__device__ __forceinline__ void dft16(double2* v) {
  // 4 x 4: inner DFT4 over j of x[c+4j], twiddle W16^(c r), outer DFT4 over c
  double2 y[4][4];
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    y[c][0] = v[c]; y[c][1] = v[c + 4]; y[c][2] = v[c + 8]; y[c][3] = v[c + 12];
    dft4<INV>(y[c][0], y[c][1], y[c][2], y[c][3]);
  }
  // c = 1: W16^1, W16^2, W16^3
  y[1][1] = wconst<INV>(y[1][1], C8, S8);
  y[1][2] = wconst<INV>(y[1][2], R2, R2);
  y[1][3] = wconst<INV>(y[1][3], S8, C8);
  // c = 2: W16^2, W16^4, W16^6
  y[2][1] = wconst<INV>(y[2][1], R2, R2);
  y[2][2] = w4<INV>(y[2][2]);
  y[2][3] = wconst<INV>(y[2][3], -R2, R2);
  // c = 3: W16^3, W16^6, W16^9
  y[3][1] = wconst<INV>(y[3][1], S8, C8);
  y[3][2] = wconst<INV>(y[3][2], -R2, R2);
  y[3][3] = wconst<INV>(y[3][3], -C8, -S8);
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    double2 a0 = y[0][r], a1 = y[1][r], a2 = y[2][r], a3 = y[3][r];
    dft4<INV>(a0, a1, a2, a3);
    v[r] = a0; v[r + 4] = a1; v[r + 8] = a2; v[r + 12] = a3;
  }
}

template <int R, bool INV>
__device__ __forceinline__ void dft(double2* v) {
  if constexpr (R == 2) dft2<INV>(v[0], v[1]);
  else if constexpr (R == 4) dft4<INV>(v[0], v[1], v[2], v[3]);
  else if constexpr (R == 8) dft8<INV>(v);
  else if constexpr (R == 16) dft16<INV>(v);
  else static_assert(R == 2, "unsupported radix");
}

// One Stockham radix-R pass over S interleaved sequences of length N.
// Ns = product of the radices of the previous passes.
template <int N, int R, int Ns, bool INV, int S, int NT>
__device__ __forceinline__ void st_pass(double2* __restrict__ lds, const double2* __restrict__ tw) {
  constexpr int NB = N / R;
  constexpr int TOT = NB * S;
  constexpr int PER = (TOT + NT - 1) / NT;
  constexpr int TWS = TWN / N;
  double2 v[PER][R];
#pragma unroll
  for (int q = 0; q < PER; ++q) {
    const int b = (int)threadIdx.x + q * NT;
    if ((TOT % NT) == 0 || b < TOT) {
      const int s = b % S, j = b / S;
#pragma unroll
      for (int r = 0; r < R; ++r) v[q][r] = lds[(j + r * NB) * S + s];
      if constexpr (Ns > 1) {
        const int k = j % Ns;
#pragma unroll
        for (int r = 1; r < R; ++r) {
          double2 w = tw[(k * r * (N / (Ns * R))) * TWS];
          if (INV) w.y = -w.y;
          v[q][r] = cmul(v[q][r], w);
        }
      }
      dft<R, INV>(v[q]);
    }
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < PER; ++q) {
    const int b = (int)threadIdx.x + q * NT;
    if ((TOT % NT) == 0 || b < TOT) {
      const int s = b % S, j = b / S;
      const int base = (j / Ns) * Ns * R + (j % Ns);
#pragma unroll
      for (int r = 0; r < R; ++r) lds[(base + r * Ns) * S + s] = v[q][r];
    }
  }
  __syncthreads();
}

// Full length-N transform of S interleaved sequences in LDS.  The caller must
// have synchronised after filling LDS; the function ends synchronised.
template <int N, bool INV, int S, int NT>
__device__ __forceinline__ void fft_lds(double2* __restrict__ lds, const double2* __restrict__ tw) {
  if constexpr (N == 16) {
    st_pass<16, 16, 1, INV, S, NT>(lds, tw);
  } else if constexpr (N == 32) {
    st_pass<32, 8, 1, INV, S, NT>(lds, tw);
    st_pass<32, 4, 8, INV, S, NT>(lds, tw);
  } else if constexpr (N == 64) {
    st_pass<64, 8, 1, INV, S, NT>(lds, tw);
    st_pass<64, 8, 8, INV, S, NT>(lds, tw);
  } else if constexpr (N == 128) {
    st_pass<128, 16, 1, INV, S, NT>(lds, tw);
    st_pass<128, 8, 16, INV, S, NT>(lds, tw);
  } else if constexpr (N == 256) {
    st_pass<256, 16, 1, INV, S, NT>(lds, tw);
    st_pass<256, 16, 16, INV, S, NT>(lds, tw);
  } else if constexpr (N == 512) {
    st_pass<512, 8, 1, INV, S, NT>(lds, tw);
    st_pass<512, 8, 8, INV, S, NT>(lds, tw);
    st_pass<512, 8, 64, INV, S, NT>(lds, tw);
  } else if constexpr (N == 1024) {
    st_pass<1024, 16, 1, INV, S, NT>(lds, tw);
    st_pass<1024, 8, 16, INV, S, NT>(lds, tw);
    st_pass<1024, 8, 128, INV, S, NT>(lds, tw);
  } else {
    static_assert(N == 16, "unsupported FFT length");
  }
}

}  // namespace qg
