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
// Position p lies in the circular window [start, start + len) mod L
// (0 <= p, start < L): the unsigned difference p - start is the distance when
// p >= start and wraps past 2^32 - L otherwise, where p - start + L is.
template <int L>
__device__ __forceinline__ bool arc_has(int p, int start, int len) {
  const unsigned d = (unsigned)(p - start);
  return min(d, d + (unsigned)L) < (unsigned)len;
}

// Two-level register FFT of one line tile: thread (ch, w) loads positions
// w + N2 r, runs pass 1 into the LDS tile, and after the barrier pass 2 gives
// it positions w + N1 k2. Every position row r is a uniform base pointer (blk
// plus a multiple of the position stride, kept in scalar registers) and the
// thread adds one 32-bit byte offset, the same for all its rows.
template <int KC, int LC>
__device__ __forceinline__ void lines_two_level(const LinesArgs& a, const float2* tw, float2* buf, int ch,
                                                int w, bool live, int2 lw, char* blk) {
  using T = Tile<KC, LC>;
  constexpr int N1 = split_n1(LC), N2 = LC / N1;
  // loaded positions are (s - load_n//2) mod L for s < load_n, the rest are zero padding;
  // stored ones likewise with store_n
  const int lh = a.load_n / 2, sh = a.store_n / 2;
  const int ls = lh ? LC - lh : 0, ss = sh ? LC - sh : 0;
  const int2 full = make_int2(0, LC);
  const int2 lp = a.inverse ? lw : full;  // support window on the loads (inverse)
  const int2 sp = a.inverse ? full : lw;  // ... or on the stores (forward)
  const int64_t row = a.pos_stride * (int64_t)sizeof(float2);
  const unsigned lane = ((unsigned)w * (unsigned)a.pos_stride + (unsigned)ch) * (unsigned)sizeof(float2);
  float2 v[N1];
  if (w < N2) {  // whole waves past N2 skip the loads
#pragma unroll
    for (int r = 0; r < N1; ++r) {
      const int p = w + N2 * r;
      // bitwise ands: the tests are cheap, short-circuit branches would cost more
      const bool ok = live & arc_has<LC>(p, ls, a.load_n) & arc_has<LC>(p, lp.x, lp.y);
      v[r] = make_float2(0.f, 0.f);
      if (__builtin_amdgcn_ballot_w64(ok))  // a wave with no live lane issues nothing
        v[r] = __builtin_bit_cast(float2, __builtin_amdgcn_raw_buffer_load_b64(buf_rsrc(blk + r * N2 * row),
                                                                                ok ? lane : kBufOob, 0, 0));
    }
#pragma unroll
    for (int r = 0; r < N1; ++r) keep_loaded(v[r]);
    if (a.inverse) {
#pragma unroll
      for (int r = 0; r < N1; ++r) v[r].y = -v[r].y;
    }
  }
  __syncthreads();  // twiddle table
  if (w < N2) two_level_pass1<LC>(v, buf, T::ld, ch, w, tw);
  __syncthreads();
  if (w < N1) {
    float2 z[N2];
    two_level_pass2<LC>(z, buf, T::ld, ch, w);
#pragma unroll
    for (int k2 = 0; k2 < N2; ++k2) {
      const int p = w + N1 * k2;
      const bool ok = live & arc_has<LC>(p, ss, a.store_n) & arc_has<LC>(p, sp.x, sp.y);
      float2 x = z[k2];
      if (a.inverse) x.y = -x.y;
      if (__builtin_amdgcn_ballot_w64(ok))
        __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(__attribute__((ext_vector_type(2))) unsigned, x),
                                              buf_rsrc(blk + k2 * N1 * row), ok ? lane : kBufOob, 0, 0);
    }
  }
}

template <int KC, int LC>
__global__ void __launch_bounds__(kBlock) k_fft_lines(LinesArgs a) {
  if (gate_closed(a.gate)) return;
  extern __shared__ float2 smem[];
  using T = Tile<KC, LC>;
  const int L = LC > 0 ? LC : a.ax.L;   // line length
  const int LF = LC > 0 ? LC : a.ax.M;  // transform length (tile rows, twiddles)
  float2* tw = smem;
  float2* buf = smem + LF;
  const int tid = threadIdx.x;
  const int ch = tid % KC, w = tid / KC;
  // block -> (line, channel chunk); runs of xcd_chunk consecutive logical
  // blocks go to one XCD (hardware dispatch is round-robin over the 8 XCDs)
  const unsigned nchunk = a.div_nchunk.d;
  unsigned bid = blockIdx.x;
  if (bid < a.xcd_full) {
    const unsigned r = bid / kNumXcd, x = bid % kNumXcd, q = a.div_xcd.div(r);
    bid = q * (kNumXcd * a.xcd_chunk) + x * a.xcd_chunk + (r - q * a.xcd_chunk);
  }
  const int line = (int)a.div_nchunk.div(bid);
  const int chunk = (int)(bid - line * nchunk);
  const int o = (int)a.div_inner.div(line);
  const int in = line - o * a.n_inner;
  // windowed outer index: o < outer_win_n <= outer_win_g, so one wrap suffices
  int ob = o;
  if (a.outer_win_n > 0) {
    ob = o - a.outer_win_n / 2;
    if (ob < 0) ob += a.outer_win_g;
  }
  const int k0 = chunk * KC;
  // support pruning acts on the k-space side: the stores of a forward pass,
  // the loads of an inverse one
  int2 lw = make_int2(0, L);
  if (a.line_win) {
    lw = a.line_win[o * a.win_so + in * a.win_si];
    if (lw.y == 0) {  // empty line: nothing live to load; its outputs are zero
      if (a.zero_empty && k0 + ch < a.K) {
        float2* b0 = a.data + ob * a.outer_stride + in * a.inner_stride + k0 + ch;
        const int nw = (int)blockDim.x / KC;  // workers per channel of this launch
        for (int sidx = w; sidx < a.store_n; sidx += nw) {
          const int p = a.store_n < L ? pmod(sidx - a.store_n / 2, L) : sidx;
          b0[p * a.pos_stride] = make_float2(0.f, 0.f);
        }
      }
      return;
    }
  }
  auto live_pos = [&](int p) { return !a.line_win || in_window(p, lw, L); };
  const bool live = k0 + ch < a.K;
  float2* base = a.data + ob * a.outer_stride + in * a.inner_stride + k0 + ch;
  for (int i = tid; i < LF; i += (int)blockDim.x) tw[i] = a.tw[i];
  // the two-level path may run with fewer than kBlock threads (two_level_threads); the launch
  // gives the LDS Stockham path below all kBlock of them (it is taken only without off32)
  if constexpr (two_level_ok<KC, LC>()) {
    if (a.off32) {
      lines_two_level<KC, LC>(a, tw, buf, ch, w, live, lw,
                              (char*)(a.data + ob * a.outer_stride + in * a.inner_stride + k0));
      return;
    }
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
    const int threads = a.off32 ? two_level_threads<KC, LC>() : kBlock;
    hipLaunchKernelGGL((k_fft_lines<KC, LC>), g, dim3(threads), sm, s, a);
  }
};

int launch_lines(int kc, int L, dim3 grid, size_t smem, hipStream_t s, const LinesArgs& a0) {
  if (a0.K <= 0 || grid.x == 0) return kOk;  // nothing to transform (and no divisor of 0 below)
  LinesArgs a = a0;
  a.gate = gate_current();
  const unsigned S = a.xcd_chunk;
  a.div_nchunk = make_fastdiv((unsigned)((a.K + kc - 1) / kc));
  a.div_inner = make_fastdiv((unsigned)a.n_inner);
  a.div_xcd = make_fastdiv(S ? S : 1u);
  a.xcd_full = S ? grid.x / (kNumXcd * S) * (kNumXcd * S) : 0u;
  a.off32 = (16 * a.pos_stride + a.K) * (int64_t)sizeof(float2) < ((int64_t)1 << 31);
  return launch_fft_kernel<LinesLaunch>(kc, L, grid, smem, s, a);
}

}  // namespace tr
