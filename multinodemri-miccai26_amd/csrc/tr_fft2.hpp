// Two-level register FFT for line lengths L = N1 * N2 with N1, N2 <= 16.
//
// Four-step decomposition (n = j + N2 r, k = k1 + N1 k2):
//   X[k1 + N1 k2] = sum_j w_N2^{j k2} [ w_L^{j k1} sum_r x[j + N2 r] w_N1^{r k1} ].
// Pass 1: worker j < N2 holds x[j + N2 r] (r < N1) in registers, runs an
// N1-point DFT there, applies w_L^{j k1} and writes row k1 of an N1 x N2
// LDS tile.  Pass 2 (after one barrier): worker k1 < N1 reads its row, runs
// an N2-point DFT in registers and holds X[k1 + N1 k2] (k2 < N2).  Compared
// with one LDS round trip per radix stage this moves the line through LDS
// once.  The register DFTs are mixed-radix (2, 3, 4, 5, 8) with twiddles
// folded to constants at compile time.
#pragma once

#include "tr_fft.hpp"

namespace tr {

// ---- compile-time sin / cos of 2 pi a / N (Taylor series in double) --------
namespace ctm {
constexpr double kPi = 3.14159265358979323846264338327950288;

constexpr double taylor_sin(double x) {
  double term = x, sum = x;
  for (int n = 1; n < 24; ++n) {
    term *= -x * x / ((2.0 * n) * (2.0 * n + 1.0));
    sum += term;
  }
  return sum;
}

constexpr double taylor_cos(double x) {
  double term = 1.0, sum = 1.0;
  for (int n = 1; n < 24; ++n) {
    term *= -x * x / ((2.0 * n - 1.0) * (2.0 * n));
    sum += term;
  }
  return sum;
}

// angle 2 pi a / N reduced to [-pi, pi]
constexpr double angle(int a, int N) {
  a %= N;
  if (2 * a > N) a -= N;
  return 2.0 * kPi * (double)a / (double)N;
}
}  // namespace ctm

// w_N^a = exp(-2 pi i a / N) for a in [0, N)
template <int N>
struct TwConst {
  float re[N];
  float im[N];
  constexpr TwConst() : re(), im() {
    for (int a = 0; a < N; ++a) {
      re[a] = (float)ctm::taylor_cos(ctm::angle(a, N));
      im[a] = (float)-ctm::taylor_sin(ctm::angle(a, N));
    }
  }
};

constexpr bool reg_base(int N) { return N == 2 || N == 3 || N == 4 || N == 5 || N == 8; }

constexpr int reg_radix(int N) {
  return N % 4 == 0 ? 4 : N % 2 == 0 ? 2 : N % 3 == 0 ? 3 : N % 5 == 0 ? 5 : 1;
}

constexpr bool reg_supported(int N) {
  if (N == 1 || reg_base(N)) return true;
  const int R = reg_radix(N);
  return R > 1 && reg_supported(N / R);
}

// In-place forward DFT of v[0..N) held in registers (natural order in and out).
template <int N>
struct RegFft {
  __host__ __device__ static __forceinline__ void run(float2 (&v)[N]) {
    if constexpr (N == 1) {
      return;
    } else if constexpr (reg_base(N)) {
      Dft<N>::run(v, nullptr, 0);
    } else {
      constexpr int R = reg_radix(N);
      constexpr int M = N / R;
      constexpr TwConst<N> t{};
      float2 s[R][M];
#pragma unroll
      for (int r = 0; r < R; ++r)
#pragma unroll
        for (int m = 0; m < M; ++m) s[r][m] = v[r + R * m];
#pragma unroll
      for (int r = 0; r < R; ++r) RegFft<M>::run(s[r]);
#pragma unroll
      for (int k1 = 0; k1 < M; ++k1) {
        float2 c[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
          const int a = (r * k1) % N;
          c[r] = a == 0 ? s[r][k1] : cmul(s[r][k1], make_float2(t.re[a], t.im[a]));
        }
        Dft<R>::run(c, nullptr, 0);
#pragma unroll
        for (int k2 = 0; k2 < R; ++k2) v[k1 + M * k2] = c[k2];
      }
    }
  }
};

// Marks a loaded value as used here.  Called on every value of a batch of
// loads right after the batch is issued, it keeps the compiler from sinking a
// load into the conditional code that consumes it (which would put a full
// wait after each load); the waits become a count-down over the batch.
__device__ __forceinline__ void keep_loaded(float2& t) { asm volatile("" : "+v"(t.x), "+v"(t.y)); }
__device__ __forceinline__ void keep_loaded(float& t) { asm volatile("" : "+v"(t)); }

// Split L = N1 * N2 (N1 = pass-1 DFT length, N2 = pass-1 workers), both in
// [2, 16] and register-supported; N1 as large as possible.  N1 == 0: none.
constexpr int split_n1(int L) {
  for (int n1 = 16; n1 >= 2; --n1)
    if (L % n1 == 0 && L / n1 >= 2 && L / n1 <= 16 && reg_supported(n1) && reg_supported(L / n1)) return n1;
  return 0;
}

// The two-level path runs for 16-channel chunks (16 workers per channel, as
// many as the longest pass needs) and a compile-time length with a split.
template <int KC, int LC>
constexpr bool two_level_ok() {
  return KC == 16 && LC > 0 && split_n1(LC) > 0 && kBlock / KC == 16;
}

// Threads of a two-level line workgroup: KC x max(N1, N2) workers rounded up to whole waves, so a
// length whose split needs fewer than 16 workers per channel (100 = 10 x 10: 160 of 256 threads)
// leaves no wave idle and more workgroups fit a CU's wave slots.  Other paths use kBlock.
template <int KC, int LC>
constexpr int two_level_threads() {
  if constexpr (!two_level_ok<KC, LC>()) {
    return kBlock;
  } else {
    constexpr int n1 = split_n1(LC), n2 = LC / n1;
    constexpr int t = KC * (n1 > n2 ? n1 : n2);
    return (t + 63) / 64 * 64;
  }
}

// Pass 1 of the two-level FFT for worker j < N2: v = x[j + N2 r] on entry;
// DFT, twiddle w_L^{j k1} (tw = the line's table, tw[i] = w_L^i) and write
// the tile row k1 at buf[(k1 N2 + j) ld + ch].
template <int L>
__host__ __device__ __forceinline__ void two_level_pass1(float2 (&v)[split_n1(L)], float2* buf, int ld, int ch, int j,
                                                const float2* tw) {
  constexpr int N1 = split_n1(L), N2 = L / N1;
  RegFft<N1>::run(v);
#pragma unroll
  for (int k1 = 0; k1 < N1; ++k1) {
    const float2 t = (k1 == 0 || j == 0) ? v[k1] : cmul(v[k1], tw[j * k1]);
    buf[(k1 * N2 + j) * ld + ch] = t;
  }
}

// Pass 2 for worker k1 < N1: read the tile row, DFT; on return
// z[k2] = X[k1 + N1 k2].
template <int L>
__host__ __device__ __forceinline__ void two_level_pass2(float2 (&z)[L / split_n1(L)], const float2* buf, int ld, int ch,
                                                int k1) {
  constexpr int N1 = split_n1(L), N2 = L / N1;
#pragma unroll
  for (int j = 0; j < N2; ++j) z[j] = buf[(k1 * N2 + j) * ld + ch];
  RegFft<N2>::run(z);
}

}  // namespace tr
