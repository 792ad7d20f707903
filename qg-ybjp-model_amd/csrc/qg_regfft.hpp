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
// per element; the buffer layout (XBuf) keeps every exchange free of LDS
// bank conflicts.
//
// All threads of the workgroup must call reg_fft together (it contains
// __syncthreads); every thread owns one sequence position set.
#pragma once
#include "qg_fft.hpp"

namespace qg {

// Radix sequences: pass p has radix R<p+1>, P passes; the first radix is E.
template <int P_, int A, int B = 1, int C = 1, int D = 1, int F = 1>
struct RP {
  static constexpr int P = P_, R1 = A, R2 = B, R3 = C, R4 = D, R5 = F;
};
template <int N, int E>
struct RPlan;
template <> struct RPlan<16, 16> : RP<1, 16> {};
template <> struct RPlan<16, 8> : RP<2, 8, 2> {};
template <> struct RPlan<16, 4> : RP<2, 4, 4> {};
template <> struct RPlan<32, 16> : RP<2, 16, 2> {};
template <> struct RPlan<32, 8> : RP<2, 8, 4> {};
template <> struct RPlan<32, 4> : RP<3, 4, 4, 2> {};
template <> struct RPlan<64, 16> : RP<2, 16, 4> {};
template <> struct RPlan<64, 8> : RP<2, 8, 8> {};
template <> struct RPlan<64, 4> : RP<3, 4, 4, 4> {};
template <> struct RPlan<128, 16> : RP<2, 16, 8> {};
template <> struct RPlan<128, 8> : RP<3, 8, 8, 2> {};
template <> struct RPlan<128, 4> : RP<4, 4, 4, 4, 2> {};
template <> struct RPlan<256, 16> : RP<2, 16, 16> {};
template <> struct RPlan<256, 8> : RP<3, 8, 8, 4> {};
template <> struct RPlan<256, 4> : RP<4, 4, 4, 4, 4> {};
template <> struct RPlan<512, 16> : RP<3, 16, 8, 4> {};
template <> struct RPlan<512, 8> : RP<3, 8, 8, 8> {};
template <> struct RPlan<512, 4> : RP<5, 4, 4, 4, 4, 2> {};

// radix of pass p and the span Ns of the sub-transforms it combines
template <int N, int E>
constexpr int rp_radix(int p) {
  using Pl = RPlan<N, E>;
  return p == 0 ? Pl::R1 : p == 1 ? Pl::R2 : p == 2 ? Pl::R3 : p == 3 ? Pl::R4 : Pl::R5;
}
template <int N, int E>
constexpr int rp_ns(int p) {
  int ns = 1;
  for (int q = 0; q < p; ++q) ns *= rp_radix<N, E>(q);
  return ns;
}

// LDS exchange buffers as seen by one thread.  gfx950 serves a b128 LDS read
// in 4 cycles of 16 lanes (lane l in cycle parity(l>>2 ^ l>>3 ^ l>>4) +
// 2*(l>>5)) over 64 four-byte banks, i.e. 16 distinct 16-byte slots, and a
// b128 write in 8 cycles of 8 consecutive lanes over 8 slots; two lanes of
// one cycle on the same slot at different addresses cost an extra cycle.
//  - ST == 1 (row kernels): the buffers of one row are contiguous, S buffers
//    of SL elements with one pad element per E for the transform passes, so
//    that the stride-E writes of the first pass fall on distinct slots; the
//    rows of a wave are row_xstride() apart (8 slots modulo 16) and lanes
//    are assigned to (row, position) by row_lane, which together make every
//    transform pass conflict-free.  The publish/read-back of whole sequences
//    (pub) uses the same buffers without padding.
//  - ST > 1 (column kernels, ST = TX columns): the TX columns of a workgroup
//    are interleaved element by element with no padding: the 8 lanes of a
//    write cycle are 8 columns (distinct slots) and a read cycle holds 8
//    columns at two positions of opposite parity (8 slots apart).
// Offsets known at compile time stay additive, so LDS instructions keep
// immediate offsets.
template <int N, int E, int ST>
struct XBuf {
  static constexpr int SL = ST == 1 ? N + N / E : N;
  double2* p;
  __device__ __forceinline__ double2& operator()(int s, int i) const {
    if constexpr (ST == 1) return p[s * SL + i + i / E];
    else return p[(s * SL + i) * ST];
  }
  __device__ __forceinline__ double2& pub(int s, int i) const {
    if constexpr (ST == 1) return p[s * SL + i];
    else return p[(s * SL + i) * ST];
  }
};
template <int N, int E>
using RowBuf = XBuf<N, E, 1>;
template <int N, int E>
using ColBuf = XBuf<N, E, ((64 / (N / E)) > 8 ? 64 / (N / E) : 8)>;

// Elements of LDS a row workgroup needs per row for SB buffers, i.e. the
// spacing of consecutive rows (8 slots modulo 16).
template <int N, int E, int SB>
constexpr int row_xstride() {
  constexpr int L = SB * RowBuf<N, E>::SL;
  return L + ((8 - L) % 16 + 16) % 16;
}
// Elements of LDS a column workgroup needs for SB buffers per column.
template <int N, int E, int SB, int TX>
constexpr int col_xlen() { return TX * SB * N; }

// Lane -> (position t, row p) of a row workgroup of RP rows with T threads
// per row.  Every quad of lanes holds 4 consecutive positions of one row (64
// contiguous bytes of a global row); for T = 8, 16, 32 and whole waves the
// 16 quads of a wave take the (row, position-quad) cells of a table chosen so
// that, with the row spacing above, no LDS cycle of the transform passes
// (first-pass writes, middle and last-pass reads, middle-pass writes) or of
// the unpadded publish/read-back has two lanes on one slot.  The tables
// are packed 6 bits per quad: row << 3 | position-quad.  Other shapes keep
// the plain mapping.
template <int T>
struct RowQuads {
  static constexpr unsigned long long lo = 0, hi = 0;
};
template <> struct RowQuads<8> { static constexpr unsigned long long lo = 0x670621c68829ull, hi = 0x211009e01439ull; };
template <> struct RowQuads<16> { static constexpr unsigned long long lo = 0x2422910c801bull, hi = 0x60b4814d065aull; };
template <> struct RowQuads<32> { static constexpr unsigned long long lo = 0x1ca2ce0c63c2ull, hi = 0x30900d205101ull; };

template <int T, int RP>
__device__ __forceinline__ void row_lane(int tid, int& t, int& p) {
  if constexpr ((T == 8 || T == 16 || T == 32) && (RP * T) % 64 == 0) {
    const int w = tid >> 6, l = tid & 63, q = l >> 2;
    const unsigned long long k = q < 8 ? RowQuads<T>::lo : RowQuads<T>::hi;
    const int e = (int)((k >> (6 * (q & 7))) & 63);
    t = 4 * (e & 7) + (l & 3);
    p = (64 / T) * w + (e >> 3);
  } else {
    t = tid % T;
    p = tid / T;
  }
}

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
// per butterfly of each pass.  Loaded once per kernel (or per transform)
// ahead of the first pass, so that no pass waits on a table read.
template <int N, int E>
struct TwB {
  using Pl = RPlan<N, E>;
  static constexpr int NP = Pl::P > 1 ? Pl::P - 1 : 1;
  double2 w[NP][E];  // w[p - 1][b]: pass p, butterfly b < E / radix
  __device__ __forceinline__ void load(const double2* __restrict__ tw, int t) {
    constexpr int T = N / E, TWS = TWN / N;
#pragma unroll
    for (int p = 1; p < Pl::P; ++p) {
      const int R = rp_radix<N, E>(p), Ns = rp_ns<N, E>(p);
#pragma unroll
      for (int b = 0; b < E / R; ++b) w[p - 1][b] = tw[(((t + T * b) % Ns) * (N / (Ns * R))) * TWS];
    }
  }
};

// Twiddle source of a transform kernel.  PRE: the bases are loaded once at
// kernel entry, before the data loads, and shared by every transform of the
// kernel (otherwise each transform loads them after its data has arrived, one
// more cache round trip per transform); costs the bases' registers for the
// kernel's lifetime.  !PRE: each transform loads its own.
template <int N, int E, bool PRE>
struct TwSrc;
template <int N, int E>
struct TwSrc<N, E, true> {
  TwB<N, E> tb;
  __device__ __forceinline__ TwSrc(const double2* __restrict__ tw, int t) { tb.load(tw, t); }
  __device__ __forceinline__ const TwB<N, E>& get() const { return tb; }
};
template <int N, int E>
struct TwSrc<N, E, false> {
  const double2* tw;
  __device__ __forceinline__ TwSrc(const double2* __restrict__ tw_, int) : tw(tw_) {}
  __device__ __forceinline__ const double2* get() const { return tw; }
};

// Passes PI .. P-2 (the middle passes) of the group of G sequences S0 ..
// S0+G-1: read from LDS, twiddle, DFT, write back in Stockham order.
template <int N, int E, int S, int G, int S0, bool INV, int ST, int PI>
__device__ __forceinline__ void rf_mid_passes(double2 (&v)[S][E], const XBuf<N, E, ST>& xb, int t,
                                              const TwB<N, E>& tb) {
  if constexpr (PI < RPlan<N, E>::P - 1) {
    constexpr int R = rp_radix<N, E>(PI), Ns = rp_ns<N, E>(PI), NB = N / R, BPT = E / R, T = N / E;
#pragma unroll
    for (int s = 0; s < G; ++s)
#pragma unroll
      for (int b = 0; b < BPT; ++b) {
        const int j = t + T * b;
        double2* w = &v[S0 + s][b * R];
#pragma unroll
        for (int r = 0; r < R; ++r) w[r] = xb(s, j + r * NB);
        double2 wr[R];
        tw_pow<R, INV>(tb.w[PI - 1][b], wr);
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
        for (int r = 0; r < R; ++r) xb(s, base + r * Ns) = v[S0 + s][b * R + r];
      }
    __syncthreads();
    rf_mid_passes<N, E, S, G, S0, INV, ST, PI + 1>(v, xb, t, tb);
  }
}

// S independent sequences transformed in groups of SB: every pass of a
// group runs between the same two barriers, so the LDS latency of one
// sequence overlaps the arithmetic of the others.  xb holds SB exchange
// buffers.
template <int N, int E, int S, int SB, int S0, bool INV, int ST>
__device__ __forceinline__ void rf_group(double2 (&v)[S][E], const XBuf<N, E, ST>& xb, int t,
                                         const TwB<N, E>& tb) {
  using Pl = RPlan<N, E>;
  constexpr int T = N / E;
  constexpr int G = (S - S0) < SB ? (S - S0) : SB;
#pragma unroll
  for (int s = S0; s < S0 + G; ++s) dft<E, INV>(v[s]);
  if constexpr (Pl::P > 1) {
    __syncthreads();  // the buffers may still be read by a previous transform
#pragma unroll
    for (int s = 0; s < G; ++s)
#pragma unroll
      for (int r = 0; r < E; ++r) xb(s, t * E + r) = v[S0 + s][r];
    __syncthreads();
    rf_mid_passes<N, E, S, G, S0, INV, ST, 1>(v, xb, t, tb);
    constexpr int PL = Pl::P - 1, R = rp_radix<N, E>(PL), NB = N / R, BPT = E / R;
#pragma unroll
    for (int s = 0; s < G; ++s)
#pragma unroll
      for (int b = 0; b < BPT; ++b) {
        const int j = t + T * b;
        double2 w[R];
#pragma unroll
        for (int r = 0; r < R; ++r) w[r] = xb(s, j + r * NB);
        double2 wr[R];
        tw_pow<R, INV>(tb.w[PL - 1][b], wr);
#pragma unroll
        for (int r = 1; r < R; ++r) w[r] = cmul(w[r], wr[r]);
        dft<R, INV>(w);
#pragma unroll
        for (int r = 0; r < R; ++r) v[S0 + s][b + r * BPT] = w[r];
      }
  }
  if constexpr (S0 + G < S) rf_group<N, E, S, SB, S0 + G, INV>(v, xb, t, tb);
}

// In-place transform of the strided distribution v (forward exp(-i...),
// inverse exp(+i...) unnormalised).  xb: this sequence's exchange buffer.
template <int N, int E, bool INV, int ST>
__device__ __forceinline__ void reg_fft(double2 (&v)[E], const XBuf<N, E, ST>& xb, int t, const TwB<N, E>& tb) {
  double2(&v1)[1][E] = *reinterpret_cast<double2(*)[1][E]>(&v);
  rf_group<N, E, 1, 1, 0, INV>(v1, xb, t, tb);
}
template <int N, int E, bool INV, int ST>
__device__ __forceinline__ void reg_fft(double2 (&v)[E], const XBuf<N, E, ST>& xb, int t,
                                        const double2* __restrict__ tw) {
  TwB<N, E> tb;
  tb.load(tw, t);
  reg_fft<N, E, INV>(v, xb, t, tb);
}

template <int N, int E, int S, int SB, bool INV, int ST>
__device__ __forceinline__ void reg_fft_multi(double2 (&v)[S][E], const XBuf<N, E, ST>& xb, int t,
                                              const TwB<N, E>& tb) {
  rf_group<N, E, S, SB, 0, INV>(v, xb, t, tb);
}
template <int N, int E, int S, int SB, bool INV, int ST>
__device__ __forceinline__ void reg_fft_multi(double2 (&v)[S][E], const XBuf<N, E, ST>& xb, int t,
                                              const double2* __restrict__ tw) {
  TwB<N, E> tb;
  tb.load(tw, t);
  rf_group<N, E, S, SB, 0, INV>(v, xb, t, tb);
}

// Publish sequences S0 .. S0+G-1 of v into buffers 0 .. G-1 so that any
// position can be read (the k <-> N-k pairs of two-for-one real transforms,
// the realit fill); the caller synchronises before the buffers are written
// again.
template <int N, int E, int S, int S0, int G, int ST>
__device__ __forceinline__ void rf_publish_range(const double2 (&v)[S][E], const XBuf<N, E, ST>& xb, int t) {
  constexpr int T = N / E;
  __syncthreads();
#pragma unroll
  for (int s = 0; s < G; ++s)
#pragma unroll
    for (int m = 0; m < E; ++m) xb.pub(s, t + T * m) = v[S0 + s][m];
  __syncthreads();
}

// Exchange the strided distribution through the buffer so that every thread
// can read arbitrary positions of the sequence (used for the k <-> N-k pairs
// of two-for-one real transforms and for the realit fill).  After the call,
// xb.pub(0, i) holds element i; the caller synchronises before the buffer
// is written again.
template <int N, int E, int ST>
__device__ __forceinline__ void rf_publish(const double2 (&v)[E], const XBuf<N, E, ST>& xb, int t) {
  constexpr int T = N / E;
  __syncthreads();
#pragma unroll
  for (int m = 0; m < E; ++m) xb.pub(0, t + T * m) = v[m];
  __syncthreads();
}

}  // namespace qg
