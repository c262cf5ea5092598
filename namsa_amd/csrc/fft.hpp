// Register-resident Stockham FFT engine for gfx950.
//
// A length-N transform is computed by T = N/E cooperating threads, each holding
// E elements in registers in the "natural distributed" layout: thread t holds
// element t + T*m in register m.  A stage of radix R (R | E) performs E/R
// radix-R butterflies per thread on registers {q + r*E/R}; its outputs go to
// LDS at Stockham autosort positions and are read back in the natural layout
// for the next stage.  The last stage writes its outputs back to the very
// registers it read, so a transform starts and ends in the natural layout and
// a pointwise multiply can sit between an inverse and a forward transform with
// no data movement.  Twiddles depend only on (stage, t) and are computed once
// per thread in double precision (sincospi) and kept in registers.
//
// Sign convention: forward = exp(-2 pi i jk/N), inverse = exp(+2 pi i jk/N),
// both unnormalised (the cuFFT C2C convention the reference calls into,
// namsa/msa.py:962,976,981,987).
#pragma once
#include "common.hpp"

namespace namsa {

// ------------------------------------------------------------------ radix-R DFTs
// Butterfly constants are carried as float pairs hi + lo (the double value to
// ~1e-15) and applied with one fused multiply-add: x * hi + (x * lo).  The
// nearest floats alone have magnitude errors of 3e-8..6e-8 (2 fl(sqrt(1/2))^2
// = 1 - 3.4e-8, fl(cos pi/8)^2 + fl(sin pi/8)^2 = 1 - 5.7e-8) that shrink every
// rotated element the same way and compound coherently over hundreds of
// slices (-2.4e-7 of the intensity per slice at 2048^2); with the pairs only
// the unbiased rounding of each product remains.
struct KPair {
  float hi, lo;
};
constexpr KPair kR2{0.70710676908493042f, 1.2101617485882343e-08f};   // sqrt(1/2)
constexpr KPair kC8{0.92387950420379639f, 2.830748968563057e-08f};    // cos(pi/8)
constexpr KPair kS8{0.38268342614173889f, 6.2233507236442165e-09f};   // sin(pi/8)
constexpr KPair kR3{0.86602538824081421f, 1.5543625053737742e-08f};   // sqrt(3)/2

__device__ __forceinline__ float kmul(float x, KPair k) { return fmaf(x, k.hi, x * k.lo); }

// (x + i y) * (c + i s) with c, s given as pairs
__device__ __forceinline__ float2 kcmul(float2 v, KPair c, KPair s, bool neg_s) {
  const float sh = neg_s ? -s.hi : s.hi, sl = neg_s ? -s.lo : s.lo;
  const float re = fmaf(v.x, c.hi, fmaf(-v.y, sh, fmaf(v.x, c.lo, -v.y * sl)));
  const float im = fmaf(v.x, sh, fmaf(v.y, c.hi, fmaf(v.x, sl, v.y * c.lo)));
  return make_float2(re, im);
}

template <int R, bool INV>
struct Dft;

template <bool INV>
struct Dft<2, INV> {
  static __device__ __forceinline__ void run(float2* a) {
    float2 t = a[0];
    a[0] = cadd(t, a[1]);
    a[1] = csub(t, a[1]);
  }
};

template <bool INV>
struct Dft<3, INV> {
  static __device__ __forceinline__ void run(float2* a) {
    float2 t1 = cadd(a[1], a[2]);
    const float2 d = csub(a[1], a[2]);
    float2 t2 = make_float2(kmul(d.x, kR3), kmul(d.y, kR3));
    float2 m = make_float2(a[0].x - 0.5f * t1.x, a[0].y - 0.5f * t1.y);
    a[0] = cadd(a[0], t1);
    float2 r = mul_mi<INV>(t2);
    a[1] = cadd(m, r);
    a[2] = csub(m, r);
  }
};

template <bool INV>
struct Dft<4, INV> {
  static __device__ __forceinline__ void run(float2* a) {
    float2 t0 = cadd(a[0], a[2]);
    float2 t1 = csub(a[0], a[2]);
    float2 t2 = cadd(a[1], a[3]);
    float2 t3 = mul_mi<INV>(csub(a[1], a[3]));
    a[0] = cadd(t0, t2);
    a[2] = csub(t0, t2);
    a[1] = cadd(t1, t3);
    a[3] = csub(t1, t3);
  }
};

template <bool INV>
struct Dft<8, INV> {
  static __device__ __forceinline__ void run(float2* a) {
    float2 e[4] = {a[0], a[2], a[4], a[6]};
    float2 o[4] = {a[1], a[3], a[5], a[7]};
    Dft<4, INV>::run(e);
    Dft<4, INV>::run(o);
    // o[k] *= exp(-+ 2 pi i k / 8)
    if (INV) {
      o[1] = make_float2(kmul(o[1].x - o[1].y, kR2), kmul(o[1].x + o[1].y, kR2));
      o[3] = make_float2(-kmul(o[3].x + o[3].y, kR2), kmul(o[3].x - o[3].y, kR2));
    } else {
      o[1] = make_float2(kmul(o[1].x + o[1].y, kR2), kmul(o[1].y - o[1].x, kR2));
      o[3] = make_float2(kmul(o[3].y - o[3].x, kR2), -kmul(o[3].x + o[3].y, kR2));
    }
    o[2] = mul_mi<INV>(o[2]);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      a[k] = cadd(e[k], o[k]);
      a[k + 4] = csub(e[k], o[k]);
    }
  }
};

template <bool INV>
struct Dft<16, INV> {
  static __device__ __forceinline__ void run(float2* a) {
    // cos / sin of k*pi/8, k = 0..7, as constant pairs (k = 0, 4 are exact)
    constexpr KPair Z{0.0f, 0.0f};
    const KPair C[8] = {Z, kC8, kR2, kS8, Z, {-kS8.hi, -kS8.lo}, {-kR2.hi, -kR2.lo}, {-kC8.hi, -kC8.lo}};
    const KPair S[8] = {Z, kS8, kR2, kC8, Z, kC8, kR2, kS8};
    float2 e[8], o[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      e[k] = a[2 * k];
      o[k] = a[2 * k + 1];
    }
    Dft<8, INV>::run(e);
    Dft<8, INV>::run(o);
#pragma unroll
    for (int k = 1; k < 8; ++k) {
      if (k == 4)
        o[k] = mul_mi<INV>(o[k]);
      else
        o[k] = kcmul(o[k], C[k], S[k], !INV);
    }
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      a[k] = cadd(e[k], o[k]);
      a[k + 8] = csub(e[k], o[k]);
    }
  }
};

// Radix 6 and 12 as prime-factor (Good-Thomas) transforms: for R = R1 R2 with
// gcd(R1, R2) = 1, input n = (R2 n1 + R1 n2) mod R and output
// k = (R2 (R2^-1 mod R1) k1 + R1 (R1^-1 mod R2) k2) mod R turn the length-R
// DFT into R2 DFTs of length R1 followed by R1 of length R2 with no twiddles
// between them (W_R^(nk) = W_R1^(n1 k1) W_R2^(n2 k2)), so the only constants
// are the radix-3 pair's.
//   R = 6 = 2 x 3:  n = (3 n1 + 2 n2) mod 6,  k = (3 k1 + 4 k2) mod 6
//   R = 12 = 3 x 4: n = (4 n1 + 3 n2) mod 12, k = (4 k1 + 9 k2) mod 12
template <int R1, int R2, int M1, int M2, bool INV>
__device__ __forceinline__ void pfa(float2* a) {
  constexpr int R = R1 * R2;
  float2 y[R2][R1];
#pragma unroll
  for (int n2 = 0; n2 < R2; ++n2) {
#pragma unroll
    for (int n1 = 0; n1 < R1; ++n1) y[n2][n1] = a[(R2 * n1 + R1 * n2) % R];
    Dft<R1, INV>::run(y[n2]);
  }
#pragma unroll
  for (int k1 = 0; k1 < R1; ++k1) {
    float2 z[R2];
#pragma unroll
    for (int n2 = 0; n2 < R2; ++n2) z[n2] = y[n2][k1];
    Dft<R2, INV>::run(z);
#pragma unroll
    for (int k2 = 0; k2 < R2; ++k2) a[(M1 * k1 + M2 * k2) % R] = z[k2];
  }
}

template <bool INV>
struct Dft<6, INV> {
  static __device__ __forceinline__ void run(float2* a) { pfa<2, 3, 3, 4, INV>(a); }
};

template <bool INV>
struct Dft<12, INV> {
  static __device__ __forceinline__ void run(float2* a) { pfa<3, 4, 4, 9, INV>(a); }
};

// ------------------------------------------------------------------ schedules
template <int... Rs>
struct Radices {};

template <int... Rs>
struct Prod;
template <>
struct Prod<> {
  static constexpr int value = 1;
};
template <int R, int... Rs>
struct Prod<R, Rs...> {
  static constexpr int value = R * Prod<Rs...>::value;
};

// Twiddle registers a thread needs for stages with NS > 1.
template <int E, int NS, int... Rs>
struct TwCount;
template <int E, int NS>
struct TwCount<E, NS> {
  static constexpr int value = 0;
};
template <int E, int NS, int R, int... Rs>
struct TwCount<E, NS, R, Rs...> {
  static constexpr int value = (NS > 1 ? (E / R) * (R - 1) : 0) + TwCount<E, NS * R, Rs...>::value;
};
// Compact twiddles: only w per butterfly is kept; w^2 .. w^(R-1) are formed
// by complex multiplication when the stage runs (fewer registers, a few more
// multiplies; the powers carry a few float32 ulp of rounding).
template <int E, int NS, int... Rs>
struct TwCountC;
template <int E, int NS>
struct TwCountC<E, NS> {
  static constexpr int value = 0;
};
// registers per butterfly in compact mode: for R > 4 the powers of two w,
// w^2, .. up to w^(R-1) (w, w^2, w^4 for R = 6, 8; and w^8 for R = 12, 16),
// the others products of at most three of them; for R <= 4 only w (w^2, w^3
// by a two-multiply chain).  A forward and an inverse transform multiply to
// |w^r|^2 per twiddle, so the magnitude error of the formed powers is what
// accumulates over the slices: with at most two products per power it stays
// ~1.5x that of directly rounded twiddles (a 1-pair power chain is ~3x and
// grows to 1.6e-4 of the CBED after 1000 slices).
constexpr int compact_per_bfly(int R) { return R > 8 ? 4 : (R > 4 ? 3 : 1); }
template <int E, int NS, int R, int... Rs>
struct TwCountC<E, NS, R, Rs...> {
  static constexpr int value =
      (NS > 1 ? (E / R) * compact_per_bfly(R) : 0) + TwCountC<E, NS * R, Rs...>::value;
};

// Synchronisation between the LDS write and read phases of one exchange.
struct WaveSync {  // all threads of one transform live in one wavefront
  static __device__ __forceinline__ void sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  }
};
struct BlockSync {
  static __device__ __forceinline__ void sync() { __syncthreads(); }
};

// LDS accessors.  RowLds: one transform per contiguous (padded) region.
// ColLds: C interleaved transforms, element (i, c) at i*C + c (padded).
// An exchange is: write, barrier, read, then either a second barrier (single
// buffer, so the next write cannot overtake a slow reader) or a switch to the
// other of two buffers (DOUBLE): a later write to this buffer comes after the
// next exchange's barrier, which every reader of this one has passed.
// SPLIT: stages may address it as at_split (see ColLds below), the padding
// period being 16.
template <bool SPLIT = false>
struct RowLdsT {
  static constexpr bool DOUBLE = false;
  static constexpr int P = SPLIT ? 16 : 0;  // 0: stages use at() only
  float2* base;
  __device__ __forceinline__ float2& at(int i) const { return base[i + (i >> 4)]; }
  __device__ __forceinline__ float2& at_split(int i0, int k) const { return base[i0 + i0 / 16 + (k + k / 16)]; }
  __device__ __forceinline__ void flip() {}
  static constexpr int padded(int n) { return n + n / 16; }
};
using RowLds = RowLdsT<false>;
// ColLds::at_split(i0, k) is at(i0 + k) for a k that is a multiple of the
// padding period P or an i0 that is: the padding then splits into a lane part
// and a constant part, so an unrolled loop over k addresses LDS as one lane
// base plus immediate offsets (with at(), the compiler evaluates every padded
// index separately and keeps them live across a persistent loop).
template <int C, bool DB = false>
struct ColLds {
  static_assert(32 % C == 0, "column tile must divide the 32-float2 padding period");
  static constexpr bool DOUBLE = DB;
  static constexpr int P = 32 / C;  // one pad every P indices
  float2* base;
  int c;
  int other;                       // DOUBLE: offset of the other buffer (+size or -size)
  const float2* roots = nullptr;   // LR transforms: exp(-2 pi i j / N), j < N/4 (Fft)
  __device__ __forceinline__ float2& at(int i) const { return base[i * C + c + ((i * C) >> 5)]; }
  __device__ __forceinline__ float2& at_split(int i0, int k) const {
    return base[i0 * C + c + i0 / P + (k * C + k / P)];
  }
  __device__ __forceinline__ void flip() {
    if constexpr (DB) {
      base += other;
      other = -other;
    }
  }
  static constexpr int padded(int n) { return n * C + (n * C) / 32; }
};

// ---------------------------------------------------------------- twiddles
// Twiddles are fixed by (length, schedule, thread), so each schedule's set is
// computed at compile time into a table in the code object; a kernel loads a
// thread's twiddles once (coalesced: entry i of thread t at i * T + t).
// Computing them at run time (FP64 sincospi per twiddle in every workgroup)
// cost about 3 us per launch of the slice loop's passes.
namespace twc {

struct Pair {
  float c, s;
};

constexpr double kTwoPi = 6.283185307179586476925286766559005768;

// sin x and cos x for |x| <= pi/4 from their Taylor series through x^29 and
// x^28 (truncation below 1e-22), summed from the smallest term.
constexpr void sincos_small(double x, double& s, double& c) {
  constexpr int K = 15;
  const double x2 = x * x;
  double ts[K] = {}, tc[K] = {};
  ts[0] = x;
  tc[0] = 1.0;
  for (int n = 1; n < K; ++n) {
    ts[n] = ts[n - 1] * (-x2 / double((2 * n) * (2 * n + 1)));
    tc[n] = tc[n - 1] * (-x2 / double((2 * n - 1) * (2 * n)));
  }
  s = 0.0;
  c = 0.0;
  for (int n = K - 1; n >= 0; --n) {
    s += ts[n];
    c += tc[n];
  }
}

constexpr double abs_d(double x) { return x < 0 ? -x : x; }

// nextafterf(x, up ? +inf : -inf)
constexpr float next_float(float x, bool up) {
  if (x == 0.0f) return up ? 1.40129846e-45f : -1.40129846e-45f;
  unsigned b = __builtin_bit_cast(unsigned, x);
  b = ((x > 0.0f) == up) ? b + 1u : b - 1u;
  return __builtin_bit_cast(float, b);
}

// exp(i theta) as a float pair whose magnitude is as close to 1 as float
// allows: of the two float neighbours of each of cos and sin, the pair with
// the smallest |c^2 + s^2 - 1|.  Rounding cos and sin independently leaves a
// magnitude error of a few 1e-8 that does not average out over a transform's
// few distinct twiddles, and it compounds over hundreds of slices (and over
// the powers of the compact-twiddle mode).
constexpr Pair unit_pair(double c, double s) {
  const float c0 = float(c), s0 = float(s);
  const float c1 = double(c0) > c ? next_float(c0, false) : next_float(c0, true);
  const float s1 = double(s0) > s ? next_float(s0, false) : next_float(s0, true);
  const float cs[2] = {c0, c1}, ss[2] = {s0, s1};
  Pair best{c0, s0};
  double err = abs_d(double(c0) * c0 + double(s0) * s0 - 1.0);
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 2; ++j) {
      const double e = abs_d(double(cs[i]) * cs[i] + double(ss[j]) * ss[j] - 1.0);
      if (e < err) {
        err = e;
        best = Pair{cs[i], ss[j]};
      }
    }
  return best;
}

// exp(2 pi i m / d): the angle is split exactly into the nearest quarter
// turn q and a remainder of at most 1/8 turn, so the series only sees
// |x| <= pi/4 and the quarter-turn rotation is exact.
constexpr Pair unit_turns(long long m, long long d) {
  m %= d;
  if (m < 0) m += d;
  const long long q = (4 * m + d / 2) / d;
  const double x = kTwoPi * double(4 * m - q * d) / double(4 * d);
  double s = 0.0, c = 0.0;
  sincos_small(x, s, c);
  switch (q & 3) {
    case 0: return unit_pair(c, s);
    case 1: return unit_pair(-s, c);
    case 2: return unit_pair(-c, -s);
    default: return unit_pair(s, -c);
  }
}

}  // namespace twc

// LR: the full twiddle set formed at each use from a quarter table of the
// N-th roots of unity in LDS (ColLds::roots, filled by the kernel from
// kRootsQ<N>): w = exp(-2 pi i e / N) is roots[e mod N/4] turned by the
// exact quarter turn (-i)^(e div N/4).  unit_turns gives the same float pair
// for an angle whatever the order it is written in (power-of-two scalings of
// its integers are exact), and the quarter turns of the table reproduce it
// (checked for every exponent at N = 2048), so LR transforms are bit-identical
// to those holding the table's twiddles in registers, at no register cost.
template <int N, int E, class RL, bool CT = false, bool LR = false>
struct Fft;

// A schedule's twiddle table (Fft::make_table), evaluated by the compiler and
// stored in the code object's read-only data.
template <class F>
__device__ constexpr typename F::Table kTwTable = F::make_table();

template <int N>
struct RootsQ {
  twc::Pair v[N / 4];
  static constexpr RootsQ make() {
    RootsQ r{};
    for (int j = 0; j < N / 4; ++j) r.v[j] = twc::unit_turns(-static_cast<long long>(j), N);
    return r;
  }
};
template <int N>
__device__ constexpr RootsQ<N> kRootsQ = RootsQ<N>::make();

template <int N, int E, bool CT, bool LR, int... Rs>
struct Fft<N, E, Radices<Rs...>, CT, LR> {
  static_assert(Prod<Rs...>::value == N, "radix schedule must multiply to N");
  static_assert(!(CT && LR), "LDS roots give the full twiddle set");
  static_assert(!LR || (N & (N - 1)) == 0, "LDS roots need a power-of-two length");
  static constexpr int T = N / E;
  static constexpr int NTW = CT ? TwCountC<E, 1, Rs...>::value : TwCount<E, 1, Rs...>::value;
  // registers per butterfly of a twiddled stage
  template <int R>
  static constexpr int per_bfly() { return CT ? compact_per_bfly(R) : R - 1; }
  // exponent of the r-th stored twiddle of a butterfly
  template <int R>
  static constexpr int stored_power(int r) { return CT && R > 4 ? 1 << (r - 1) : r; }
  static constexpr int NTW_TAB = NTW > 0 ? NTW : 1;
  static constexpr int NTW_ALLOC = LR ? 1 : NTW_TAB;  // twiddle registers a caller holds
  static constexpr int NSTAGES = sizeof...(Rs);

  // ---------------- twiddles (forward sign; inverse uses the conjugate)
  // Entry OFF + q * P + (r - 1) of thread t is w^p with w = exp(-2 pi i k /
  // (NS R)), k = (t + T q) mod NS, p = stored_power(r): the r-th stored power
  // for butterfly q of the stage whose sub-transforms have length NS.
  struct Table {
    twc::Pair v[NTW_TAB * T];
  };

  template <int NS, int OFF>
  static constexpr void fill_stages(Table&) {}

  template <int NS, int OFF, int R, int... Rest>
  static constexpr void fill_stages(Table& tab) {
    constexpr int Q = E / R;
    constexpr int P = per_bfly<R>();
    if constexpr (NS > 1) {
      for (int t = 0; t < T; ++t)
        for (int q = 0; q < Q; ++q) {
          const int k = (t + T * q) % NS;
          for (int r = 1; r <= P; ++r)
            tab.v[(OFF + q * P + (r - 1)) * T + t] =
                twc::unit_turns(-static_cast<long long>(stored_power<R>(r)) * k, static_cast<long long>(NS) * R);
        }
      fill_stages<NS * R, OFF + Q * P, Rest...>(tab);
    } else {
      fill_stages<NS * R, OFF, Rest...>(tab);
    }
  }

  static constexpr Table make_table() {
    Table tab{};
    fill_stages<1, 0, Rs...>(tab);
    return tab;
  }

  static __device__ __forceinline__ void twiddles(int t, float2 (&tw)[NTW_ALLOC]) {
    if constexpr (NTW > 0 && !LR) {
      const twc::Pair* p = kTwTable<Fft>.v + t;
#pragma unroll
      for (int i = 0; i < NTW; ++i) tw[i] = make_float2(p[i * T].c, p[i * T].s);
    }
  }

  // ---------------- one stage
  template <bool INV, int R, int NS, int OFF, bool LAST, class Lds, class Sync, bool LAUNDERED>
  static __device__ __forceinline__ void stage(float2 (&v)[E], Lds& lds, int t,
                                               const float2 (&tw)[NTW_ALLOC]) {
    constexpr int Q = E / R;
    static_assert(E % R == 0, "radix must divide the per-thread element count");
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      float2 a[R];
#pragma unroll
      for (int r = 0; r < R; ++r) a[r] = v[q + r * Q];
      if constexpr (NS > 1) {
        if constexpr (CT && R > 4) {
          // stored w^(2^j); w^r = w^h w^(r-h), h the top bit of r
          constexpr int P = compact_per_bfly(R);
          float2 pw[R];
#pragma unroll
          for (int j = 0; j < P; ++j) {
            float2 x = tw[OFF + q * P + j];
            // opaque to the optimiser (no instruction is emitted): keeps the
            // products from being hoisted out of the persistent loop, which
            // would bring back one register pair per power (run() may have
            // done this once for all stored powers instead)
            if constexpr (!LAUNDERED) asm volatile("" : "+v"(x.x), "+v"(x.y));
            pw[1 << j] = x;
          }
#pragma unroll
          for (int r = 3; r < R; ++r) {
            const int h = r >= 8 ? 8 : (r >= 4 ? 4 : 2);
            if (r != h) pw[r] = cmul(pw[h], pw[r - h]);
          }
#pragma unroll
          for (int r = 1; r < R; ++r) a[r] = INV ? cmulc(a[r], pw[r]) : cmul(a[r], pw[r]);
        } else if constexpr (CT) {
          float2 w1 = tw[OFF + q];
          if constexpr (!LAUNDERED) asm volatile("" : "+v"(w1.x), "+v"(w1.y));  // see above
          float2 w = w1;
#pragma unroll
          for (int r = 1; r < R; ++r) {
            a[r] = INV ? cmulc(a[r], w) : cmul(a[r], w);
            if (r + 1 < R) w = cmul(w, w1);
          }
        } else if constexpr (LR) {
          // w^r = exp(-2 pi i e / N), e = r k N / (NS R) mod N, k = (t + T q) mod NS; t is made
          // opaque so that the exponents are formed here, not hoisted into registers
          int tt = t;
          asm volatile("" : "+v"(tt));
          constexpr int SCALE = N / (NS * R), QN = N / 4;
          const int step = ((tt + T * q) & (NS - 1)) * SCALE;
          int e = 0;
#pragma unroll
          for (int r = 1; r < R; ++r) {
            e = (e + step) & (N - 1);
            const float2 u = lds.roots[e & (QN - 1)];
            const int quarter = e / QN;
            float2 w = (quarter & 1) ? make_float2(u.y, -u.x) : u;
            if (quarter & 2) w = make_float2(-w.x, -w.y);
            a[r] = INV ? cmulc(a[r], w) : cmul(a[r], w);
          }
        } else {
#pragma unroll
          for (int r = 1; r < R; ++r) {
            const float2 w = tw[OFF + q * (R - 1) + (r - 1)];
            a[r] = INV ? cmulc(a[r], w) : cmul(a[r], w);
          }
        }
      }
      Dft<R, INV>::run(a);
      if constexpr (LAST) {
#pragma unroll
        for (int r = 0; r < R; ++r) v[q + r * Q] = a[r];
      } else {
        const int j = t + T * q;
        const int k = j % NS;
        const int dst = (j / NS) * NS * R + k;
        // split form: the stride NS is a multiple of the padding period, or
        // NS = 1 and dst = j R is
        constexpr int P = Lds::P;
        constexpr bool SPLIT = P > 0 && (NS % P == 0 || (NS == 1 && R % P == 0));
        int d = dst;
        // Otherwise, where the layout has a split form, the R padded
        // addresses are formed here each time (d is made opaque) rather than
        // hoisted out of the caller's persistent loop into R live registers.
        if constexpr (P > 0 && !SPLIT) asm volatile("" : "+v"(d));
#pragma unroll
        for (int r = 0; r < R; ++r) {
          if constexpr (SPLIT)
            lds.at_split(d, r * NS) = a[r];
          else
            lds.at(d + r * NS) = a[r];
        }
      }
    }
    if constexpr (!LAST) {
      Sync::sync();
      constexpr bool SPLIT = Lds::P > 0 && T % (Lds::P > 0 ? Lds::P : 1) == 0;
#pragma unroll
      for (int m = 0; m < E; ++m) {
        if constexpr (SPLIT)
          v[m] = lds.at_split(t, T * m);
        else
          v[m] = lds.at(t + T * m);
      }
      if constexpr (Lds::DOUBLE)
        lds.flip();
      else
        Sync::sync();
    }
  }

  template <bool INV, int NS, int OFF, class Lds, class Sync, bool LAUNDERED>
  static __device__ __forceinline__ void stages(float2 (&)[E], Lds&, int,
                                                const float2 (&)[NTW_ALLOC]) {}

  template <bool INV, int NS, int OFF, class Lds, class Sync, bool LAUNDERED, int R, int... Rest>
  static __device__ __forceinline__ void stages(float2 (&v)[E], Lds& lds, int t,
                                                const float2 (&tw)[NTW_ALLOC]) {
    stage<INV, R, NS, OFF, sizeof...(Rest) == 0, Lds, Sync, LAUNDERED>(v, lds, t, tw);
    constexpr int NEXT_OFF = OFF + (NS > 1 ? (E / R) * per_bfly<R>() : 0);
    stages<INV, NS * R, NEXT_OFF, Lds, Sync, LAUNDERED, Rest...>(v, lds, t, tw);
  }

  // Full transform, natural layout in and out.  LAUNDER (compact twiddles
  // only): pass the stored powers through the empty asm once here, rewriting
  // the caller's twiddle registers in place, instead of copying each power
  // before each use in the stages.
  template <bool INV, class Sync, bool LAUNDER = false, class Lds>
  static __device__ __forceinline__ void run(float2 (&v)[E], Lds& lds, int t, float2 (&tw)[NTW_ALLOC]) {
    if constexpr (CT && LAUNDER) {
#pragma unroll
      for (int i = 0; i < NTW; ++i) asm volatile("" : "+v"(tw[i].x), "+v"(tw[i].y));
    }
    stages<INV, 1, 0, Lds, Sync, CT && LAUNDER, Rs...>(v, lds, t, tw);
  }
};

}  // namespace namsa
