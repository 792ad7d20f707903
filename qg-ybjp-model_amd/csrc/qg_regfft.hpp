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

// Twiddles w^r, r = 1..R-1, of one butterfly from its base w (a table value,
// conjugated for the inverse): w^(2a) = (w^a)^2 and w^(p+a) = w^p w^a for
// the largest power of two p <= r, so every power is at most log2(R)
// products away from the table value.
template <int R, bool INV>
__device__ __forceinline__ void tw_pow(double2 w1, double2 (&wr)[R]) {
  if (INV) w1.y = -w1.y;
  wr[0] = make_double2(1.0, 0.0);
  if constexpr (R > 1) wr[1] = w1;
#pragma unroll
  for (int r = 2; r < R; ++r) {
    int p = 1;
    while (2 * p <= r) p *= 2;
    wr[r] = (r == p) ? cmul(wr[p / 2], wr[p / 2]) : cmul(wr[p], wr[r - p]);
  }
}

// Twiddle bases of thread t for the passes after the first: one table value
// per butterfly of the middle pass (m) and of the last pass (l).  Loaded
// once per kernel (or per transform) ahead of the first pass, so that no
// pass waits on a table read.
template <int N, int E>
struct TwB {
  using Pl = RPlan<N, E>;
  static constexpr int RM = Pl::P == 3 ? Pl::R2 : 1;
  static constexpr int RL = Pl::P == 2 ? Pl::R2 : (Pl::P == 3 ? Pl::R3 : 1);
  static constexpr int BM = Pl::P == 3 ? E / RM : 0;
  static constexpr int BL = Pl::P >= 2 ? E / RL : 0;
  double2 m[BM > 0 ? BM : 1];
  double2 l[BL > 0 ? BL : 1];
  __device__ __forceinline__ void load(const double2* __restrict__ tw, int t) {
    constexpr int T = N / E, TWS = TWN / N;
    if constexpr (BM > 0) {
      constexpr int Ns = Pl::R1;
#pragma unroll
      for (int b = 0; b < BM; ++b) m[b] = tw[(((t + T * b) % Ns) * (N / (Ns * RM))) * TWS];
    }
    if constexpr (BL > 0) {
      constexpr int Ns = N / RL;
#pragma unroll
      for (int b = 0; b < BL; ++b) l[b] = tw[(((t + T * b) % Ns) * (N / (Ns * RL))) * TWS];
    }
  }
};

template <int N, int E, bool INV>
__device__ __forceinline__ void rf_first(double2 (&v)[E], double2* xb, int t) {
  dft<E, INV>(v);
  __syncthreads();  // the buffer may still be read by a previous transform
#pragma unroll
  for (int r = 0; r < E; ++r) xb[xpad<E>(t * E + r)] = v[r];
  __syncthreads();
}

template <int N, int E, int R, int Ns, bool INV>
__device__ __forceinline__ void rf_mid(double2* xb, int t, const TwB<N, E>& tb) {
  constexpr int T = N / E, NB = N / R, BPT = E / R;
  double2 w[BPT][R];
#pragma unroll
  for (int b = 0; b < BPT; ++b) {
    const int j = t + T * b;
#pragma unroll
    for (int r = 0; r < R; ++r) w[b][r] = xb[xpad<E>(j + r * NB)];
    double2 wr[R];
    tw_pow<R, INV>(tb.m[b], wr);
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
__device__ __forceinline__ void rf_last(double2 (&v)[E], const double2* xb, int t, const TwB<N, E>& tb) {
  constexpr int T = N / E, NB = N / R, BPT = E / R;
  static_assert(Ns * R == N, "last pass");
#pragma unroll
  for (int b = 0; b < BPT; ++b) {
    const int j = t + T * b;
    double2 w[R];
#pragma unroll
    for (int r = 0; r < R; ++r) w[r] = xb[xpad<E>(j + r * NB)];
    double2 wr[R];
    tw_pow<R, INV>(tb.l[b], wr);
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
__device__ __forceinline__ void reg_fft(double2 (&v)[E], double2* xb, int t, const TwB<N, E>& tb) {
  using Pl = RPlan<N, E>;
  if constexpr (Pl::P == 1) {
    dft<E, INV>(v);
  } else if constexpr (Pl::P == 2) {
    rf_first<N, E, INV>(v, xb, t);
    rf_last<N, E, Pl::R2, Pl::R1, INV>(v, xb, t, tb);
  } else {
    rf_first<N, E, INV>(v, xb, t);
    rf_mid<N, E, Pl::R2, Pl::R1, INV>(xb, t, tb);
    rf_last<N, E, Pl::R3, Pl::R1 * Pl::R2, INV>(v, xb, t, tb);
  }
}
template <int N, int E, bool INV>
__device__ __forceinline__ void reg_fft(double2 (&v)[E], double2* xb, int t, const double2* __restrict__ tw) {
  TwB<N, E> tb;
  tb.load(tw, t);
  reg_fft<N, E, INV>(v, xb, t, tb);
}

// S independent sequences transformed in groups of SB: every pass of a
// group runs between the same two barriers, so the LDS latency of one
// sequence overlaps the arithmetic of the others.  xb holds SB consecutive
// exchange buffers of xbuf_len<N, E>() elements.
template <int N, int E, int S, int SB, int S0, bool INV>
__device__ __forceinline__ void rf_group(double2 (&v)[S][E], double2* xb, int t, const TwB<N, E>& tb) {
  using Pl = RPlan<N, E>;
  constexpr int XL = xbuf_len<N, E>(), T = N / E;
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
          double2 wr[R];
          tw_pow<R, INV>(tb.m[b], wr);
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
    constexpr int NB = N / R, BPT = E / R;
#pragma unroll
    for (int s = 0; s < G; ++s)
#pragma unroll
      for (int b = 0; b < BPT; ++b) {
        const int j = t + T * b;
        double2 w[R];
#pragma unroll
        for (int r = 0; r < R; ++r) w[r] = xb[s * XL + xpad<E>(j + r * NB)];
        double2 wr[R];
        tw_pow<R, INV>(tb.l[b], wr);
#pragma unroll
        for (int r = 1; r < R; ++r) w[r] = cmul(w[r], wr[r]);
        dft<R, INV>(w);
#pragma unroll
        for (int r = 0; r < R; ++r) v[S0 + s][b + r * BPT] = w[r];
      }
  }
  if constexpr (S0 + G < S) rf_group<N, E, S, SB, S0 + G, INV>(v, xb, t, tb);
}

template <int N, int E, int S, int SB, bool INV>
__device__ __forceinline__ void reg_fft_multi(double2 (&v)[S][E], double2* xb, int t, const TwB<N, E>& tb) {
  rf_group<N, E, S, SB, 0, INV>(v, xb, t, tb);
}
template <int N, int E, int S, int SB, bool INV>
__device__ __forceinline__ void reg_fft_multi(double2 (&v)[S][E], double2* xb, int t,
                                              const double2* __restrict__ tw) {
  TwB<N, E> tb;
  tb.load(tw, t);
  rf_group<N, E, S, SB, 0, INV>(v, xb, t, tb);
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
