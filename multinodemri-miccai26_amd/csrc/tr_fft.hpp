// LDS-resident mixed-radix Stockham FFT along one axis of a channel-innermost grid.
//
// A workgroup holds one line of L points for KC channels in LDS as
// buf[pos * ld + ch] (ld = KC + 1 for KC >= 4 so the row stride is odd in
// 8-byte words).  Lane ch of every worker owns one channel, so every LDS access
// of a stage is conflict-free and every global access is a run of KC
// consecutive complex values.  Stages run the autosort (Stockham) recursion:
// butterfly j of a radix-R stage reads x[j + r L/R], twiddles by
// w_L^{(j mod Ns) r L/(Ns R)}, applies DFT_R and writes
// y[(j / Ns) Ns R + (j mod Ns) + r Ns].  Each worker keeps its butterflies of a
// stage in registers across one barrier, so a single LDS buffer suffices.
// The inverse transform is conj(FFT(conj(x))): callers conjugate on load and
// store.
#pragma once

#include "tr_common.hpp"

namespace tr {

constexpr int kBlock = 256;       // threads per FFT workgroup
constexpr int kPosPerWorker = 16;  // line positions per worker: a chunk of KC channels has
                                   // kBlock / KC workers, so lines of up to 16 kBlock / KC points
constexpr int kMaxLine = kPosPerWorker * kBlock;  // longest line a one-channel workgroup holds
constexpr int kMaxTransform = 2048;  // longest transform accepted (its tile and a row of the x pass
                                     // fit the 64 KiB of LDS a workgroup may use)
constexpr int kMaxStages = 14;

// Longest transform a workgroup of KC channels holds.
__host__ __device__ constexpr int max_line(int kc) { return kMaxLine / kc; }

// One axis of the oversampled grid.  Lengths whose prime factors are at most 13
// are transformed directly (M = L).  Others use Bluestein's chirp-z identity
// nk = (n^2 + k^2 - (k - n)^2) / 2: X_k = w_k sum_n (x_n w_n) conj(w_{k-n}) with
// w_m = exp(-i pi m^2 / L), a circular convolution evaluated with two forward
// FFTs of a length M >= 2L - 1 whose factors are 2, 3 and 5; chirp[n] = w_n
// (n < L) and hhat = FFT_M(b) / M with b_m = conj(w_m), b_{M-m} = conj(w_m)
// (0 <= m < L), both computed in float64 on the host.
struct FftAxis {
  int L;        // logical length
  int M;        // transform length (L, or the Bluestein convolution length)
  int bluestein;
  int nstages;  // radices of the length-M transform
  int radix[kMaxStages];
  const float2* chirp;  // device, L values (Bluestein only)
  const float2* hhat;   // device, M values (Bluestein only)
};

__host__ __device__ constexpr int lds_ld(int kc) { return kc >= 4 ? kc + 1 : kc; }

// ---- small DFT kernels (forward, w = exp(-2 pi i / R)) --------------------
__host__ __device__ __forceinline__ float2 mul_mi(float2 a) { return make_float2(a.y, -a.x); }  // -i * a

template <int R>
struct Dft;

template <>
struct Dft<2> {
  __host__ __device__ static __forceinline__ void run(float2 (&v)[2], const float2*, int) {
    const float2 a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
  }
};

template <>
struct Dft<3> {
  __host__ __device__ static __forceinline__ void run(float2 (&v)[3], const float2*, int) {
    const float s = 0.86602540378443864676f;  // sin(2 pi / 3)
    const float2 t1 = cadd(v[1], v[2]);
    const float2 d = csub(v[1], v[2]);
    const float2 t2 = make_float2(v[0].x - 0.5f * t1.x, v[0].y - 0.5f * t1.y);
    v[0] = cadd(v[0], t1);
    v[1] = make_float2(t2.x + s * d.y, t2.y - s * d.x);
    v[2] = make_float2(t2.x - s * d.y, t2.y + s * d.x);
  }
};

template <>
struct Dft<4> {
  __host__ __device__ static __forceinline__ void run(float2 (&v)[4], const float2*, int) {
    const float2 a = cadd(v[0], v[2]), b = csub(v[0], v[2]);
    const float2 c = cadd(v[1], v[3]), d = mul_mi(csub(v[1], v[3]));
    v[0] = cadd(a, c);
    v[2] = csub(a, c);
    v[1] = cadd(b, d);
    v[3] = csub(b, d);
  }
};

template <>
struct Dft<5> {
  __host__ __device__ static __forceinline__ void run(float2 (&v)[5], const float2*, int) {
    const float c1 = 0.30901699437494742410f;   // cos(2 pi / 5)
    const float c2 = -0.80901699437494742410f;  // cos(4 pi / 5)
    const float s1 = 0.95105651629515357212f;   // sin(2 pi / 5)
    const float s2 = 0.58778525229247312917f;   // sin(4 pi / 5)
    const float2 a1 = cadd(v[1], v[4]), b1 = csub(v[1], v[4]);
    const float2 a2 = cadd(v[2], v[3]), b2 = csub(v[2], v[3]);
    const float2 x0 = v[0];
    const float2 r1 = make_float2(x0.x + c1 * a1.x + c2 * a2.x, x0.y + c1 * a1.y + c2 * a2.y);
    const float2 r2 = make_float2(x0.x + c2 * a1.x + c1 * a2.x, x0.y + c2 * a1.y + c1 * a2.y);
    const float2 i1 = make_float2(s1 * b1.x + s2 * b2.x, s1 * b1.y + s2 * b2.y);
    const float2 i2 = make_float2(s2 * b1.x - s1 * b2.x, s2 * b1.y - s1 * b2.y);
    v[0] = make_float2(x0.x + a1.x + a2.x, x0.y + a1.y + a2.y);
    v[1] = cadd(r1, mul_mi(i1));
    v[4] = csub(r1, mul_mi(i1));
    v[2] = cadd(r2, mul_mi(i2));
    v[3] = csub(r2, mul_mi(i2));
  }
};

template <>
struct Dft<8> {
  __host__ __device__ static __forceinline__ void run(float2 (&v)[8], const float2* tw, int L) {
    const float h = 0.70710678118654752440f;
    float2 e[4] = {v[0], v[2], v[4], v[6]};
    float2 o[4] = {v[1], v[3], v[5], v[7]};
    Dft<4>::run(e, tw, L);
    Dft<4>::run(o, tw, L);
    // o[k] *= w8^k, w8 = (1 - i)/sqrt(2)
    o[1] = make_float2(h * (o[1].x + o[1].y), h * (o[1].y - o[1].x));
    o[2] = mul_mi(o[2]);
    o[3] = make_float2(h * (o[3].y - o[3].x), -h * (o[3].x + o[3].y));
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      v[k] = cadd(e[k], o[k]);
      v[k + 4] = csub(e[k], o[k]);
    }
  }
};

// Generic odd prime radix via the line's twiddle table: w_R^m = tw[m L / R].
template <int R>
struct Dft {
  __host__ __device__ static __forceinline__ void run(float2 (&v)[R], const float2* tw, int L) {
    const int step = L / R;
    float2 out[R];
#pragma unroll
    for (int k = 0; k < R; ++k) {
      float2 acc = v[0];
#pragma unroll
      for (int r = 1; r < R; ++r) {
        const float2 w = tw[((r * k) % R) * step];
        acc = cadd(acc, cmul(v[r], w));
      }
      out[k] = acc;
    }
#pragma unroll
    for (int k = 0; k < R; ++k) v[k] = out[k];
  }
};

// One radix-R stage over the line held in buf.  W workers per channel.
template <int R, int W>
__device__ __forceinline__ void fft_stage(float2* buf, int ld, int ch, int w, int L, int Ns,
                                          const float2* tw) {
  constexpr int NB = (kPosPerWorker * W / R + W - 1) / W;
  const int nb = L / R;
  const int tstride = L / (Ns * R);
  float2 v[NB][R];
#pragma unroll
  for (int b = 0; b < NB; ++b) {
    const int j = w + b * W;
    if (j < nb) {
      const int jm = j % Ns;
#pragma unroll
      for (int r = 0; r < R; ++r) {
        float2 a = buf[(j + r * nb) * ld + ch];
        if (r > 0 && jm != 0) a = cmul(a, tw[jm * r * tstride]);
        v[b][r] = a;
      }
      Dft<R>::run(v[b], tw, L);
    }
  }
  __syncthreads();
#pragma unroll
  for (int b = 0; b < NB; ++b) {
    const int j = w + b * W;
    if (j < nb) {
      const int jm = j % Ns;
      const int base = (j / Ns) * Ns * R + jm;
#pragma unroll
      for (int r = 0; r < R; ++r) buf[(base + r * Ns) * ld + ch] = v[b][r];
    }
  }
  __syncthreads();
}

// Forward FFT of length ax.M of the line in buf (natural order in and out).
template <int W>
__device__ __forceinline__ void fft_stages(float2* buf, int ld, int ch, int w, const FftAxis& ax,
                                           const float2* tw) {
  int Ns = 1;
  for (int s = 0; s < ax.nstages; ++s) {
    const int R = ax.radix[s];
    switch (R) {
      case 2: fft_stage<2, W>(buf, ld, ch, w, ax.M, Ns, tw); break;
      case 3: fft_stage<3, W>(buf, ld, ch, w, ax.M, Ns, tw); break;
      case 4: fft_stage<4, W>(buf, ld, ch, w, ax.M, Ns, tw); break;
      case 5: fft_stage<5, W>(buf, ld, ch, w, ax.M, Ns, tw); break;
      case 7: fft_stage<7, W>(buf, ld, ch, w, ax.M, Ns, tw); break;
      case 8: fft_stage<8, W>(buf, ld, ch, w, ax.M, Ns, tw); break;
      case 11: fft_stage<11, W>(buf, ld, ch, w, ax.M, Ns, tw); break;
      case 13: fft_stage<13, W>(buf, ld, ch, w, ax.M, Ns, tw); break;
      default: break;  // rejected on the host
    }
    Ns *= R;
  }
}

// Full forward DFT of length ax.L of the line in buf (positions 0 .. L-1 in and
// out; a Bluestein axis uses positions L .. M-1 as scratch).  Must be called by
// every thread of the block (contains barriers).
template <int W>
__device__ __forceinline__ void fft_line(float2* buf, int ld, int ch, int w, const FftAxis& ax,
                                         const float2* tw) {
  if (!ax.bluestein) {
    fft_stages<W>(buf, ld, ch, w, ax, tw);
    return;
  }
  const int L = ax.L, M = ax.M;
  for (int n = w; n < M; n += W) {  // a = x w, zero-padded to M
    float2& x = buf[n * ld + ch];
    x = n < L ? cmul(x, ax.chirp[n]) : make_float2(0.f, 0.f);
  }
  __syncthreads();
  fft_stages<W>(buf, ld, ch, w, ax, tw);
  for (int k = w; k < M; k += W) {  // conj(FFT(a) FFT(b) / M)
    float2& x = buf[k * ld + ch];
    x = cconj(cmul(x, ax.hhat[k]));
  }
  __syncthreads();
  fft_stages<W>(buf, ld, ch, w, ax, tw);  // conj of the circular convolution a (*) b
  for (int k = w; k < L; k += W) {
    float2& x = buf[k * ld + ch];
    x = cmul(cconj(x), ax.chirp[k]);
  }
  __syncthreads();
}

// ---- compile-time plans ------------------------------------------------------
// For a length known at compile time every stage's radix, stride and
// butterflies-per-worker are constants, which keeps the stages at a fraction of
// the registers of the runtime-radix path above.
constexpr int kRadixOrder[8] = {8, 4, 2, 3, 5, 7, 11, 13};

constexpr int ct_nstages(int L) {
  int n = 0;
  for (int i = 0; i < 8; ++i)
    while (L % kRadixOrder[i] == 0) {
      L /= kRadixOrder[i];
      ++n;
    }
  return n;
}

constexpr int ct_radix(int L, int s) {
  int k = 0;
  for (int i = 0; i < 8; ++i)
    while (L % kRadixOrder[i] == 0) {
      if (k == s) return kRadixOrder[i];
      L /= kRadixOrder[i];
      ++k;
    }
  return 1;
}

constexpr bool ct_supported(int L) {
  int n = L;
  for (int i = 0; i < 8; ++i)
    while (n % kRadixOrder[i] == 0) n /= kRadixOrder[i];
  return n == 1 && L <= max_line(16);
}

template <int R, int W, int L, int Ns>
__device__ __forceinline__ void ct_stage(float2* buf, int ld, int ch, int w, const float2* tw) {
  constexpr int nb = L / R;
  constexpr int NB = (nb + W - 1) / W;
  constexpr int tstride = L / (Ns * R);
  float2 v[NB][R];
#pragma unroll
  for (int b = 0; b < NB; ++b) {
    const int j = w + b * W;
    if (NB * W == nb || j < nb) {
      const int jm = j % Ns;
#pragma unroll
      for (int r = 0; r < R; ++r) {
        float2 a = buf[(j + r * nb) * ld + ch];
        if (r > 0 && Ns > 1) a = cmul(a, tw[jm * r * tstride]);
        v[b][r] = a;
      }
      Dft<R>::run(v[b], tw, L);
    }
  }
  __syncthreads();
#pragma unroll
  for (int b = 0; b < NB; ++b) {
    const int j = w + b * W;
    if (NB * W == nb || j < nb) {
      const int jm = j % Ns;
      const int base = (j / Ns) * Ns * R + jm;
#pragma unroll
      for (int r = 0; r < R; ++r) buf[(base + r * Ns) * ld + ch] = v[b][r];
    }
  }
  __syncthreads();
}

template <int L, int S, int Ns, int W>
__device__ __forceinline__ void ct_fft(float2* buf, int ld, int ch, int w, const float2* tw) {
  if constexpr (S < ct_nstages(L)) {
    constexpr int R = ct_radix(L, S);
    ct_stage<R, W, L, Ns>(buf, ld, ch, w, tw);
    ct_fft<L, S + 1, Ns * R, W>(buf, ld, ch, w, tw);
  }
}

// L > 0: compile-time plan for length L; L == 0: runtime plan in ax.
template <int W, int L>
__device__ __forceinline__ void fft_any(float2* buf, int ld, int ch, int w, const FftAxis& ax,
                                        const float2* tw) {
  if constexpr (L > 0) {
    ct_fft<L, 0, 1, W>(buf, ld, ch, w, tw);
  } else {
    fft_line<W>(buf, ld, ch, w, ax, tw);
  }
}

// Host: factor n into supported radices (8s first, then 4, 2, odd primes);
// false if a prime factor above 13 remains.
inline bool factor_radices(int n, int* nstages, int* radix) {
  *nstages = 0;
  auto take = [&](int r) {
    while (n % r == 0 && *nstages < kMaxStages) {
      radix[(*nstages)++] = r;
      n /= r;
    }
  };
  take(8);
  take(4);
  take(2);
  take(3);
  take(5);
  take(7);
  take(11);
  take(13);
  return n == 1;
}

// Host: the transform of an axis of length L (tables are filled in by the plan).
// Direct when L's prime factors are at most 13 and L <= kMaxTransform; otherwise
// Bluestein with the smallest M = 2^a 3^b 5^c >= 2L - 1 (at most kMaxTransform).
inline bool make_fft_axis(int L, FftAxis* ax) {
  if (L < 1) return false;
  ax->L = L;
  ax->chirp = nullptr;
  ax->hhat = nullptr;
  if (L <= kMaxTransform && factor_radices(L, &ax->nstages, ax->radix)) {
    ax->M = L;
    ax->bluestein = 0;
    return true;
  }
  int M = 0;
  for (int m = 2 * L - 1; m <= kMaxTransform && !M; ++m) {
    int r = m;
    for (int f : {2, 3, 5})
      while (r % f == 0) r /= f;
    if (r == 1) M = m;
  }
  if (!M) return false;
  ax->M = M;
  ax->bluestein = 1;
  return factor_radices(M, &ax->nstages, ax->radix);
}

}  // namespace tr
