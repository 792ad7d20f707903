// Register-resident FP64 FFT of one length-N sequence spread over T = N/E
// threads: thread t holds element t + T*m in v[m] (m < E), before and after
// the transform ("strided distribution").
//
// The transform is the Stockham radix sequence of RPlan<N, E>.  The first
// pass has radix E and works on the registers directly (one butterfly per
// thread, no twiddles); every later pass reads its inputs from the
// sequence's LDS exchange buffer, and the last pass writes its outputs back
// into registers at exactly the strided positions, so that a pointwise
// operation (products, the time update) can run between two transforms with
// no data movement.  A 2-pass transform costs one LDS write and one LDS read
// per element.  LDS addresses carry one pad element per E elements so that
// the radix-E strided writes of the first pass are free of bank conflicts.
//
// All threads of the workgroup must call reg_fft together (it contains
// __syncthreads); every thread owns one sequence position set.
#pragma once
#include "qg_fft.hpp"

namespace qg {

template <int N, int E>
struct RPlan;
template <> struct RPlan<16, 16>  { static constexpr int P = 1, R1 = 16, R2 = 1, R3 = 1; };
template <> struct RPlan<16, 8>   { static constexpr int P = 2, R1 = 8,  R2 = 2, R3 = 1; };
template <> struct RPlan<32, 16>  { static constexpr int P = 2, R1 = 16, R2 = 2, R3 = 1; };
template <> struct RPlan<32, 8>   { static constexpr int P = 2, R1 = 8,  R2 = 4, R3 = 1; };
template <> struct RPlan<64, 16>  { static constexpr int P = 2, R1 = 16, R2 = 4, R3 = 1; };
template <> struct RPlan<64, 8>   { static constexpr int P = 2, R1 = 8,  R2 = 8, R3 = 1; };
template <> struct RPlan<128, 16> { static constexpr int P = 2, R1 = 16, R2 = 8, R3 = 1; };
template <> struct RPlan<128, 8>  { static constexpr int P = 3, R1 = 8,  R2 = 8, R3 = 2; };
template <> struct RPlan<256, 16> { static constexpr int P = 2, R1 = 16, R2 = 16, R3 = 1; };
template <> struct RPlan<256, 8>  { static constexpr int P = 3, R1 = 8,  R2 = 8, R3 = 4; };
template <> struct RPlan<512, 16> { static constexpr int P = 3, R1 = 16, R2 = 8, R3 = 4; };
template <> struct RPlan<512, 8>  { static constexpr int P = 3, R1 = 8,  R2 = 8, R3 = 8; };

// padded LDS index and the size of one sequence's exchange buffer; the +1
// keeps consecutive buffers (one per column) on different banks
template <int E>
__device__ __forceinline__ int xpad(int i) { return i + i / E; }
template <int N, int E>
constexpr int xbuf_len() { return N + N / E + 1; }

template <bool INV>
__device__ __forceinline__ double2 twid(const double2* __restrict__ tw, int idx) {
  double2 w = tw[idx];
  if (INV) w.y = -w.y;
  return w;
}

// Twiddles w^r, r = 1..R-1, of one butterfly from a single table load
// w = tw[k * TWN / (Ns R)]: w^(2a) = (w^a)^2 and w^(p+a) = w^p w^a for the
// largest power of two p <= r, so every power is at most log2(R) products
// from the table value (one table load per butterfly instead of R-1).
template <int R, bool INV>
__device__ __forceinline__ void tw_powers(const double2* __restrict__ tw, int idx, double2 (&wr)[R]) {
  wr[0] = make_double2(1.0, 0.0);
  if constexpr (R > 1) wr[1] = twid<INV>(tw, idx);
#pragma unroll
  for (int r = 2; r < R; ++r) {
    int p = 1;
    while (2 * p <= r) p *= 2;
    wr[r] = (r == p) ? cmul(wr[p / 2], wr[p / 2]) : cmul(wr[p], wr[r - p]);
  }
}

template <int N, int E, bool INV>
__device__ __forceinline__ void rf_first(double2 (&v)[E], double2* xb, int t) {
  dft<E, INV>(v);
  __syncthreads();  // the buffer may still be read by a previous transform
#pragma unroll
  for (int r = 0; r < E; ++r) xb[xpad<E>(t * E + r)] = v[r];
  __syncthreads();
}

template <int N, int E, int R, int Ns, bool INV>
__device__ __forceinline__ void rf_mid(double2* xb, int t, const double2* __restrict__ tw) {
  constexpr int T = N / E, NB = N / R, BPT = E / R, TWS = TWN / N;
  double2 w[BPT][R];
#pragma unroll
  for (int b = 0; b < BPT; ++b) {
    const int j = t + T * b;
#pragma unroll
    for (int r = 0; r < R; ++r) w[b][r] = xb[xpad<E>(j + r * NB)];
    const int k = j % Ns;
    double2 wr[R];
    tw_powers<R, INV>(tw, (k * (N / (Ns * R))) * TWS, wr);
#pragma unroll
    for (int r = 1; r < R; ++r) w[b][r] = cmul(w[b][r], wr[r]);
    dft<R, INV>(w[b]);
  }
  __syncthreads();
#pragma unroll
  for (int b = 0; b < BPT; ++b) {
    const int j = t + T * b;
    const int base = (j / Ns) * Ns * R + (j % Ns);
#pragma unroll
    for (int r = 0; r < R; ++r) xb[xpad<E>(base + r * Ns)] = w[b][r];
  }
  __syncthreads();
}

template <int N, int E, int R, int Ns, bool INV>
__device__ __forceinline__ void rf_last(double2 (&v)[E], const double2* xb, int t, const double2* __restrict__ tw) {
  constexpr int T = N / E, NB = N / R, BPT = E / R, TWS = TWN / N;
  static_assert(Ns * R == N, "last pass");
#pragma unroll
  for (int b = 0; b < BPT; ++b) {
    const int j = t + T * b;
    double2 w[R];
#pragma unroll
    for (int r = 0; r < R; ++r) w[r] = xb[xpad<E>(j + r * NB)];
    const int k = j % Ns;
        double2 wr[R];
        tw_powers<R, INV>(tw, (k * (N / (Ns * R))) * TWS, wr);
#pragma unroll
        for (int r = 1; r < R; ++r) w[r] = cmul(w[r], wr[r]);
    dft<R, INV>(w);
#pragma unroll
    for (int r = 0; r < R; ++r) v[b + r * BPT] = w[r];
  }
}

// In-place transform of the strided distribution v (forward exp(-i...),
// inverse exp(+i...) unnormalised).  xb: this sequence's exchange buffer.
template <int N, int E, bool INV>
__device__ __forceinline__ void reg_fft(double2 (&v)[E], double2* xb, int t, const double2* __restrict__ tw) {
  using Pl = RPlan<N, E>;
  if constexpr (Pl::P == 1) {
    dft<E, INV>(v);
  } else if constexpr (Pl::P == 2) {
    rf_first<N, E, INV>(v, xb, t);
    rf_last<N, E, Pl::R2, Pl::R1, INV>(v, xb, t, tw);
  } else {
    rf_first<N, E, INV>(v, xb, t);
    rf_mid<N, E, Pl::R2, Pl::R1, INV>(xb, t, tw);
    rf_last<N, E, Pl::R3, Pl::R1 * Pl::R2, INV>(v, xb, t, tw);
  }
}

// S independent sequences transformed in groups of SB: every pass of a
// group runs between the same two barriers, so the LDS latency of one
// sequence overlaps the arithmetic of the others.  xb holds SB consecutive
// exchange buffers of xbuf_len<N, E>() elements.
template <int N, int E, int S, int SB, int S0, bool INV>
__device__ __forceinline__ void rf_group(double2 (&v)[S][E], double2* xb, int t, const double2* __restrict__ tw) {
  using Pl = RPlan<N, E>;
  constexpr int XL = xbuf_len<N, E>(), T = N / E, TWS = TWN / N;
  constexpr int G = (S - S0) < SB ? (S - S0) : SB;
#pragma unroll
  for (int s = S0; s < S0 + G; ++s) dft<E, INV>(v[s]);
  if constexpr (Pl::P > 1) {
    __syncthreads();
#pragma unroll
    for (int s = 0; s < G; ++s)
#pragma unroll
      for (int r = 0; r < E; ++r) xb[s * XL + xpad<E>(t * E + r)] = v[S0 + s][r];
    __syncthreads();
    if constexpr (Pl::P == 3) {
      constexpr int R = Pl::R2, Ns = Pl::R1, NB = N / R, BPT = E / R;
#pragma unroll
      for (int s = 0; s < G; ++s)
#pragma unroll
        for (int b = 0; b < BPT; ++b) {
          const int j = t + T * b;
          double2* w = &v[S0 + s][b * R];
#pragma unroll
          for (int r = 0; r < R; ++r) w[r] = xb[s * XL + xpad<E>(j + r * NB)];
          const int k = j % Ns;
        double2 wr[R];
        tw_powers<R, INV>(tw, (k * (N / (Ns * R))) * TWS, wr);
#pragma unroll
        for (int r = 1; r < R; ++r) w[r] = cmul(w[r], wr[r]);
          dft<R, INV>(w);
        }
      __syncthreads();
#pragma unroll
      for (int s = 0; s < G; ++s)
#pragma unroll
        for (int b = 0; b < BPT; ++b) {
          const int j = t + T * b;
          const int base = (j / Ns) * Ns * R + (j % Ns);
#pragma unroll
          for (int r = 0; r < R; ++r) xb[s * XL + xpad<E>(base + r * Ns)] = v[S0 + s][b * R + r];
        }
      __syncthreads();
    }
    constexpr int R = (Pl::P == 2) ? Pl::R2 : Pl::R3;
    constexpr int Ns = N / R, NB = N / R, BPT = E / R;
#pragma unroll
    for (int s = 0; s < G; ++s)
#pragma unroll
      for (int b = 0; b < BPT; ++b) {
        const int j = t + T * b;
        double2 w[R];
#pragma unroll
        for (int r = 0; r < R; ++r) w[r] = xb[s * XL + xpad<E>(j + r * NB)];
        const int k = j % Ns;
        double2 wr[R];
        tw_powers<R, INV>(tw, (k * (N / (Ns * R))) * TWS, wr);
#pragma unroll
        for (int r = 1; r < R; ++r) w[r] = cmul(w[r], wr[r]);
        dft<R, INV>(w);
#pragma unroll
        for (int r = 0; r < R; ++r) v[S0 + s][b + r * BPT] = w[r];
      }
  }
  if constexpr (S0 + G < S) rf_group<N, E, S, SB, S0 + G, INV>(v, xb, t, tw);
}

template <int N, int E, int S, int SB, bool INV>
__device__ __forceinline__ void reg_fft_multi(double2 (&v)[S][E], double2* xb, int t,
                                              const double2* __restrict__ tw) {
  rf_group<N, E, S, SB, 0, INV>(v, xb, t, tw);
}

// Publish sequences S0 .. S0+G-1 of v into buffers 0 .. G-1 so that any
// position can be read (the k <-> N-k pairs of two-for-one real transforms,
// the realit fill); the caller synchronises before the buffers are written
// again.
template <int N, int E, int S, int S0, int G>
__device__ __forceinline__ void rf_publish_range(const double2 (&v)[S][E], double2* xb, int t) {
  constexpr int T = N / E, XL = xbuf_len<N, E>();
  __syncthreads();
#pragma unroll
  for (int s = 0; s < G; ++s)
#pragma unroll
    for (int m = 0; m < E; ++m) xb[s * XL + xpad<E>(t + T * m)] = v[S0 + s][m];
  __syncthreads();
}

// Exchange the strided distribution through the buffer so that every thread
// can read arbitrary positions of the sequence (used for the k <-> N-k pairs
// of two-for-one real transforms and for the realit fill).  After the call,
// xb[xpad<E>(i)] holds element i; the caller synchronises before the buffer
// is written again.
template <int N, int E>
__device__ __forceinline__ void rf_publish(const double2 (&v)[E], double2* xb, int t) {
  constexpr int T = N / E;
  __syncthreads();
#pragma unroll
  for (int m = 0; m < E; ++m) xb[xpad<E>(t + T * m)] = v[m];
  __syncthreads();
}

}  // namespace qg
