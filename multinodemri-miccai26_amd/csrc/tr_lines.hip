// FFT passes along y and z of the channel-innermost grid.

#include "tr_fft2.hpp"
#include "tr_internal.hpp"

namespace tr {

// ---------------------------------------------------------------------------
// Generic pass along y or z: load a (possibly windowed) line of KC channels,
// FFT (or inverse), store a (possibly windowed) set of positions.
// Block = (channel chunk, line) with the chunk fastest, so concurrent blocks
// cover contiguous channel runs of the same positions; with xcd_chunk, runs of
// whole lines stay on one XCD.
// ---------------------------------------------------------------------------
template <int KC, int LC>
__global__ void __launch_bounds__(kBlock) k_fft_lines(LinesArgs a) {
  extern __shared__ float2 smem[];
  using T = Tile<KC, LC>;
  const int L = LC > 0 ? LC : a.ax.L;
  float2* tw = smem;
  float2* buf = smem + L;
  const int tid = threadIdx.x;
  const int ch = tid % KC, w = tid / KC;
  const int nchunk = (a.K + KC - 1) / KC;
  const unsigned bid = xcd_block(blockIdx.x, gridDim.x, a.xcd_chunk);
  const int line = bid / nchunk;
  const int chunk = bid - line * nchunk;
  const int o = line / a.n_inner;
  const int in = line - o * a.n_inner;
  // support pruning acts on the k-space side: the stores of a forward pass,
  // the loads of an inverse one
  int2 lw = make_int2(0, 0);
  if (a.line_win) {
    lw = a.line_win[o * a.win_so + in * a.win_si];
    if (lw.y == 0) {  // empty line: nothing live to load; its outputs are zero
      if (a.zero_empty) {
        const int nchunk0 = (a.K + KC - 1) / KC;
        const int k00 = (bid % nchunk0) * KC;
        if (k00 + ch < a.K) {
          const int Ls = LC > 0 ? LC : a.ax.L;
          const int ob0 = a.outer_win_n > 0 ? pmod(o - a.outer_win_n / 2, a.outer_win_g) : o;
          float2* b0 = a.data + ob0 * a.outer_stride + in * a.inner_stride + k00 + ch;
          for (int sidx = w; sidx < a.store_n; sidx += T::W) {
            const int p = a.store_n < Ls ? pmod(sidx - a.store_n / 2, Ls) : sidx;
            b0[p * a.pos_stride] = make_float2(0.f, 0.f);
          }
        }
      }
      return;
    }
  }
  auto live_pos = [&](int p) { return !a.line_win || in_window(p, lw, LC > 0 ? LC : a.ax.L); };
  const int ob = a.outer_win_n > 0 ? pmod(o - a.outer_win_n / 2, a.outer_win_g) : o;
  const int k0 = chunk * KC;
  const bool live = k0 + ch < a.K;
  float2* base = a.data + ob * a.outer_stride + in * a.inner_stride + k0 + ch;
  for (int i = tid; i < L; i += kBlock) tw[i] = a.tw[i];
  if constexpr (two_level_ok<KC, LC>()) {
    // two-level register FFT: positions p = w + N2 r in, p = w + N1 k2 out
    constexpr int N1 = split_n1(LC), N2 = LC / N1;
    const int lh = a.load_n / 2, sh = a.store_n / 2;
    const bool wl = a.load_n < LC, ws = a.store_n < LC;
    // Masked loads (padding and pruned positions are never fetched), all issued before any is
    // used so that they are in flight together.
    float2 v[N1];
#pragma unroll
    for (int r = 0; r < N1; ++r) {
      const int p = w + N2 * r;
      // loaded positions are (s - n//2) mod L for s < load_n; the rest are zero padding
      bool ok = w < N2 && live && (!wl || pmod(p + lh, LC) < a.load_n);
      if (a.inverse) ok = ok && live_pos(p);
      v[r] = make_float2(0.f, 0.f);
      if (ok) v[r] = base[p * a.pos_stride];
    }
#pragma unroll
    for (int r = 0; r < N1; ++r) keep_loaded(v[r]);
    if (a.inverse) {
#pragma unroll
      for (int r = 0; r < N1; ++r) v[r].y = -v[r].y;
    }
    __syncthreads();  // twiddle table
    if (w < N2) two_level_pass1<LC>(v, buf, T::ld, ch, w, tw);
    __syncthreads();
    if (w < N1 && live) {
      float2 z[N2];
      two_level_pass2<LC>(z, buf, T::ld, ch, w);
#pragma unroll
      for (int k2 = 0; k2 < N2; ++k2) {
        const int p = w + N1 * k2;
        if ((!ws || pmod(p + sh, LC) < a.store_n) && (a.inverse || live_pos(p))) {
          float2 x = z[k2];
          if (a.inverse) x.y = -x.y;
          base[p * a.pos_stride] = x;
        }
      }
    }
    return;
  }
  const bool win_load = a.load_n < L;
  float2 v[T::PER];
#pragma unroll
  for (int it = 0; it < T::PER; ++it) {
    const int sidx = w + it * T::W;
    v[it] = make_float2(0.f, 0.f);
    if (sidx < a.load_n && live) {
      const int p = win_load ? pmod(sidx - a.load_n / 2, L) : sidx;
      if (!a.inverse || live_pos(p)) v[it] = base[p * a.pos_stride];
    }
  }
#pragma unroll
  for (int it = 0; it < T::PER; ++it) {
    const int sidx = w + it * T::W;
    if (sidx < L) {
      const int p = sidx < a.load_n ? (win_load ? pmod(sidx - a.load_n / 2, L) : sidx)
                                    : (a.load_n - a.load_n / 2) + (sidx - a.load_n);
      float2 x = v[it];
      if (a.inverse) x.y = -x.y;
      buf[p * T::ld + ch] = x;
    }
  }
  __syncthreads();
  fft_any<T::W, LC>(buf, T::ld, ch, w, a.ax, tw);
  if (!live) return;
  const bool win_store = a.store_n < L;
#pragma unroll
  for (int it = 0; it < T::PER; ++it) {
    const int sidx = w + it * T::W;
    if (sidx < a.store_n) {
      const int p = win_store ? pmod(sidx - a.store_n / 2, L) : sidx;
      float2 x = buf[p * T::ld + ch];
      if (a.inverse) x.y = -x.y;
      if (a.inverse || live_pos(p)) base[p * a.pos_stride] = x;
    }
  }
}

struct LinesLaunch {
  template <int KC, int LC>
  static void go(dim3 g, size_t sm, hipStream_t s, const LinesArgs& a) {
    hipLaunchKernelGGL((k_fft_lines<KC, LC>), g, dim3(kBlock), sm, s, a);
  }
};

int launch_lines(int kc, int L, dim3 grid, size_t smem, hipStream_t s, const LinesArgs& a) {
  return launch_fft_kernel<LinesLaunch>(kc, L, grid, smem, s, a);
}

}  // namespace tr
