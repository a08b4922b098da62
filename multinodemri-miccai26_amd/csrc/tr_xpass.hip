// Apodise / pad / FFT along x (forward) and inverse FFT / crop / coil sum
// along x (adjoint), one image row per workgroup.

#include "tr_fft2.hpp"
#include "tr_internal.hpp"

namespace tr {

// ---------------------------------------------------------------------------
// Forward pass along x: apodise, coil-weight, zero-pad and FFT one image row.
// Block = (coil chunk, image row); the block walks all echoes, so the coil
// maps of the row are read once.
// ---------------------------------------------------------------------------
// Two-level lengths ask for 4 blocks per CU (<= 128 VGPRs, 4 waves per SIMD, with the image rows of
// 4 echoes staged at a time so the LDS holds 4 blocks as well); the runtime-length path keeps its
// registers.
template <int KC, int LC>
constexpr int fwd_x_min_blocks() {
  return (LC > 0 && two_level_ok<KC, LC>()) ? 4 : 1;
}

template <int KC, int LC>
__global__ void __launch_bounds__(kBlock, (fwd_x_min_blocks<KC, LC>())) k_fwd_x_pad(FwdXArgs a) {
  if (gate_closed(a.gate)) return;
  extern __shared__ float2 smem[];
  using T = Tile<KC, LC>;
  const int LF = LC > 0 ? LC : a.ax.M;  // transform length (tile rows, twiddles)
  float2* tw = smem;
  float2* buf = smem + LF;
  const int tid = threadIdx.x;
  const int ch = tid % KC, w = tid / KC;
  const int nz = a.gg.n[0], ny = a.gg.n[1], nx = a.gg.n[2];
  const int gz = a.gg.g[0], gy = a.gg.g[1], gx = a.gg.g[2];
  const int nchunk = (a.C + KC - 1) / KC;
  const int row = blockIdx.x / nchunk;
  const int c0 = (blockIdx.x - row * nchunk) * KC;
  const int z = row / ny, y = row - (row / ny) * ny;
  const int64_t N = (int64_t)nz * ny * nx;
  const int c = c0 + ch;
  const bool live = c < a.C;
  for (int i = tid; i < LF; i += kBlock) tw[i] = a.tw[i];
  const int64_t rowvox = ((int64_t)z * ny + y) * nx;
  const float azy = a.apod_z[z] * a.apod_y[y] * a.scale;
  if constexpr (two_level_ok<KC, LC>()) {
    // two-level register FFT: worker w < N2 holds padded positions p = w + N2 r, whose image
    // column is x = (p + nx//2) mod gx (zero padding where x >= nx); the next echo's row is
    // loaded while the current one is transformed
    constexpr int N1 = split_n1(LC), N2 = LC / N1;
    // Every load below reads a clamped, in-bounds address and the padding is applied with a
    // select afterwards, so no load sits in a per-lane branch and all of them are in flight at once.
    // The E image rows (shared by the block's 16 coils) are staged in LDS behind the FFT tile.
    const int E = a.E;
    const int EG = a.eg;  // echoes whose image rows are staged together
    float2* img_row = buf + LC * T::ld;
    auto stage_rows = [&](int e0) {
      const int ne = min(EG, E - e0);
      // 4 loads per thread in flight before their LDS writes (1.27 -> 1.25 ms per launch)
      const int n = ne * nx;
      for (int i0 = tid; i0 < n; i0 += 4 * kBlock) {
        float2 t4[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          const int i = i0 + k * kBlock;
          const int e = i / nx;
          t4[k] = make_float2(0.f, 0.f);
          if (i < n) t4[k] = a.img[(int64_t)(e0 + e) * N + rowvox + (i - e * nx)];  // only i < n is read
        }
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          const int i = i0 + k * kBlock;
          if (i < n) img_row[i] = t4[k];
        }
      }
    };
    stage_rows(0);
    const int cc = min(c, a.C - 1);
    float2 sm[N1];
    float ap[N1];
#pragma unroll
    for (int r = 0; r < N1; ++r) ap[r] = a.apod_x[min(pmod(w + N2 * r + nx / 2, gx), nx - 1)];
    if (a.maps) {
#pragma unroll
      for (int r = 0; r < N1; ++r) sm[r] = a.maps[(int64_t)cc * N + rowvox + min(pmod(w + N2 * r + nx / 2, gx), nx - 1)];
    } else {
#pragma unroll
      for (int r = 0; r < N1; ++r) sm[r] = make_float2(1.f, 0.f);
    }
#pragma unroll
    for (int r = 0; r < N1; ++r) {
      keep_loaded(sm[r]);
      keep_loaded(ap[r]);
    }
#pragma unroll
    for (int r = 0; r < N1; ++r) {
      const int x = pmod(w + N2 * r + nx / 2, gx);
      sm[r] = (w < N2 && x < nx && live) ? cscale(sm[r], azy * ap[r]) : make_float2(0.f, 0.f);
    }
    const int zp = pmod(z - nz / 2, gz), yp = pmod(y - ny / 2, gy);
    __syncthreads();  // image rows staged
    for (int e = 0; e < E; ++e) {
      if (e > 0 && e % EG == 0) {
        // the next group's rows: every read of the previous group's rows lies before the barrier
        // that preceded the previous echo's pass 1, which every thread has passed
        stage_rows(e);
        __syncthreads();
      }
      const int el = e % EG;
      float2 v[N1];
      // sm is zero at padding positions, so the product is the zero-padded, coil-weighted row
#pragma unroll
      for (int r = 0; r < N1; ++r)
        v[r] = cmul(img_row[el * nx + min(pmod(w + N2 * r + nx / 2, gx), nx - 1)], sm[r]);
      __syncthreads();  // twiddles / the previous echo's pass-2 reads
      if (w < N2) two_level_pass1<LC>(v, buf, T::ld, ch, w, tw);
      __syncthreads();
      if (w < N1) {
        float2 zz[N2];
        two_level_pass2<LC>(zz, buf, T::ld, ch, w);
        // the grid row as a raw buffer: one 32-bit lane offset and the position row k2 a scalar
        // offset, instead of 16 64-bit addresses; lanes past the chunk's coils store nothing
        const __amdgpu_buffer_rsrc_t rs =
            buf_rsrc(reinterpret_cast<char*>(a.grid + (((int64_t)zp * gy + yp) * gx) * a.K + (int64_t)e * a.C));
        const unsigned lo = live ? ((unsigned)w * (unsigned)a.K + (unsigned)c) * 8u : kBufOob;
        const unsigned rowb = (unsigned)N1 * (unsigned)a.K * 8u;
#pragma unroll
        for (int k2 = 0; k2 < N2; ++k2)
          __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(__attribute__((ext_vector_type(2))) unsigned, zz[k2]),
                                                rs, lo, (int)(k2 * rowb), 0);
      }
    }
    return;
  }
  // map row of this thread's coil, pre-multiplied by the apodisation
  float2 sm[T::PER];
#pragma unroll
  for (int it = 0; it < T::PER; ++it) {
    const int x = w + it * T::W;
    float2 v = make_float2(0.f, 0.f);
    if (x < nx && live) {
      v = a.maps ? a.maps[(int64_t)c * N + rowvox + x] : make_float2(1.f, 0.f);
      v = cscale(v, azy * a.apod_x[x]);
    }
    sm[it] = v;
  }
  const int zp = pmod(z - nz / 2, gz), yp = pmod(y - ny / 2, gy);
  const int E = a.E;
  for (int e = 0; e < E; ++e) {
    float2 u[T::PER];
#pragma unroll
    for (int it = 0; it < T::PER; ++it) {
      const int x = w + it * T::W;
      u[it] = (x < nx) ? a.img[(int64_t)e * N + rowvox + x] : make_float2(0.f, 0.f);
    }
#pragma unroll
    for (int it = 0; it < T::PER; ++it) {
      const int x = w + it * T::W;
      if (x < nx) buf[pmod(x - nx / 2, gx) * T::ld + ch] = cmul(u[it], sm[it]);
    }
    for (int q = w; q < gx - nx; q += T::W) buf[((nx - nx / 2) + q) * T::ld + ch] = make_float2(0.f, 0.f);
    __syncthreads();
    fft_any<T::W, LC>(buf, T::ld, ch, w, a.ax, tw);
    float2* dst = a.grid + (((int64_t)zp * gy + yp) * gx) * a.K + (int64_t)e * a.C;
    if (live) {
#pragma unroll
      for (int it = 0; it < T::PER; ++it) {
        const int p = w + it * T::W;
        if (p < gx) dst[(int64_t)p * a.K + c] = buf[p * T::ld + ch];
      }
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// Adjoint pass along x: inverse FFT of one grid row, crop, apodise, multiply
// by conj(map) and sum the coils in increasing order.  Block = image row; it
// walks coil chunks (outer) and echoes (inner) so each map row is read once,
// keeping one coil-sum accumulator per echo.
// ---------------------------------------------------------------------------
template <int KC, int LC>
__global__ void __launch_bounds__(kBlock) k_adj_x_crop(AdjXArgs a) {
  if (gate_closed(a.gate)) return;
  extern __shared__ float2 smem[];
  using T = Tile<KC, LC>;
  constexpr int kSlots = (max_line(KC) + kBlock - 1) / kBlock;
  constexpr int kMaxE = kAdjXEchoes;
  const int LF = LC > 0 ? LC : a.ax.M;  // transform length (tile rows, twiddles)
  float2* tw = smem;
  float2* buf = smem + LF;
  const int tid = threadIdx.x;
  const int ch = tid % KC, w = tid / KC;
  const int nz = a.gg.n[0], ny = a.gg.n[1], nx = a.gg.n[2];
  const int gz = a.gg.g[0], gy = a.gg.g[1], gx = a.gg.g[2];
  const int row = blockIdx.x;
  const int z = row / ny, y = row - (row / ny) * ny;
  const int64_t N = (int64_t)nz * ny * nx;
  const int E = a.E;
  for (int i = tid; i < LF; i += kBlock) tw[i] = a.tw[i];
  const int zp = pmod(z - nz / 2, gz), yp = pmod(y - ny / 2, gy);
  const float2* src = a.grid + (((int64_t)zp * gy + yp) * gx) * a.K;
  const int64_t rowvox = ((int64_t)z * ny + y) * nx;
  if constexpr (two_level_ok<KC, LC>()) {
    // two-level register FFT; pass 2 multiplies its outputs by conj(map) and writes them to
    // LDS by image column, where thread x adds the chunk's coils in increasing order (the same
    // products in the same order as the stage-by-stage path)
    constexpr int N1 = split_n1(LC), N2 = LC / N1;
    // The block's rounds (coil chunk, echo) run in order, and each round's 16 grid loads are issued
    // one round ahead, so they are in flight while the previous round is transformed.  The grid row
    // is a raw buffer: each lane adds one 32-bit byte offset and the position row r of a load is a
    // scalar offset (r N2 K 8 bytes); workers past N2 get an offset past the record count and read 0
    // without touching memory.
    const __amdgpu_buffer_rsrc_t rs = buf_rsrc(const_cast<char*>(reinterpret_cast<const char*>(src)));
    const unsigned rowb = (unsigned)N2 * (unsigned)a.K * 8u;
    auto load_round = [&](int c0r, int er, float2 (&dst)[N1]) {
      const int chr = c0r + min(ch, min(KC, a.C - c0r) - 1);
      const unsigned lo = w < N2 ? ((unsigned)w * (unsigned)a.K + (unsigned)(er * a.C + chr)) * 8u : kBufOob;
#pragma unroll
      for (int r = 0; r < N1; ++r)
        dst[r] = __builtin_bit_cast(float2, __builtin_amdgcn_raw_buffer_load_b64(rs, lo, (int)(r * rowb), 0));
    };
    float2 vn[N1];
    load_round(0, 0, vn);
    float2 acc[kMaxE];
#pragma unroll
    for (int e = 0; e < kMaxE; ++e) acc[e] = make_float2(0.f, 0.f);
    for (int c0 = 0; c0 < a.C; c0 += KC) {
      const int cn = min(KC, a.C - c0);
      // clamped, unconditional loads (see the forward pass); results are selected afterwards
      const int chc = c0 + min(ch, cn - 1);
      float2 mp[N2];
      if (a.maps) {
#pragma unroll
        for (int k2 = 0; k2 < N2; ++k2)
          mp[k2] = a.maps[(int64_t)chc * N + rowvox + min(pmod(w + N1 * k2 + nx / 2, gx), nx - 1)];
      } else {
#pragma unroll
        for (int k2 = 0; k2 < N2; ++k2) mp[k2] = make_float2(1.f, 0.f);
      }
#pragma unroll
      for (int k2 = 0; k2 < N2; ++k2) keep_loaded(mp[k2]);
#pragma unroll
      for (int e = 0; e < kMaxE; ++e) {
        if (e < E) {
          float2 v[N1];
          const bool lv = w < N2 && ch < cn;
#pragma unroll
          for (int r = 0; r < N1; ++r) v[r] = vn[r];
#pragma unroll
          for (int r = 0; r < N1; ++r) keep_loaded(v[r]);  // this round's loads have landed
          if (e + 1 < E)
            load_round(c0, e + 1, vn);
          else if (c0 + KC < a.C)
            load_round(c0 + KC, 0, vn);
          __builtin_amdgcn_sched_barrier(0);  // the next round's loads stay ahead of this round's work
#pragma unroll
          for (int r = 0; r < N1; ++r) v[r] = lv ? cconj(v[r]) : make_float2(0.f, 0.f);
          __syncthreads();  // twiddles / the previous accumulation's reads
          if (w < N2) two_level_pass1<LC>(v, buf, T::ld, ch, w, tw);
          __syncthreads();
          float2 zz[N2];
          if (w < N1) two_level_pass2<LC>(zz, buf, T::ld, ch, w);
          __syncthreads();  // every tile read before the products overwrite it
          if (w < N1 && ch < cn) {
#pragma unroll
            for (int k2 = 0; k2 < N2; ++k2) {
              const int x = pmod(w + N1 * k2 + nx / 2, gx);
              if (x < nx) {
                const float2 r = cconj(zz[k2]);
                buf[x * T::ld + ch] = a.maps ? cmulc(mp[k2], r) : r;
              }
            }
          }
          __syncthreads();
          if (tid < nx)
            for (int cc = 0; cc < cn; ++cc) acc[e] = cadd(acc[e], buf[tid * T::ld + cc]);
        }
      }
    }
    if (tid < nx) {
      const int x = tid;
      const float sc = (a.apod_z[z] * a.apod_y[y] * a.apod_x[x]) * a.scale;
#pragma unroll
      for (int e = 0; e < kMaxE; ++e) {
        if (e < E) {
          const float2 r = cscale(acc[e], sc);
          float2* o = a.out + (int64_t)e * N + rowvox + x;
          *o = a.accumulate ? cadd(*o, r) : r;
        }
      }
    }
    return;
  }
  // Echo-outer: one accumulator per image column slot of this thread; per (echo, column) the
  // coils are added in increasing order (chunks in order, coils of a chunk in order), the same
  // products in the same order as the two-level path.
  for (int e = 0; e < E; ++e) {
    float2 acc[kSlots];
#pragma unroll
    for (int s = 0; s < kSlots; ++s) acc[s] = make_float2(0.f, 0.f);
    for (int c0 = 0; c0 < a.C; c0 += KC) {
      const int cn = min(KC, a.C - c0);
      float2 v[T::PER];
#pragma unroll
      for (int it = 0; it < T::PER; ++it) {
        const int p = w + it * T::W;
        v[it] = (p < gx && ch < cn) ? src[(int64_t)p * a.K + (int64_t)e * a.C + c0 + ch] : make_float2(0.f, 0.f);
      }
#pragma unroll
      for (int it = 0; it < T::PER; ++it) {
        const int p = w + it * T::W;
        if (p < gx) buf[p * T::ld + ch] = cconj(v[it]);
      }
      __syncthreads();
      fft_any<T::W, LC>(buf, T::ld, ch, w, a.ax, tw);
#pragma unroll
      for (int s = 0; s < kSlots; ++s) {
        const int x = tid + s * kBlock;
        if (x < nx) {
          const int p = pmod(x - nx / 2, gx);
          for (int cc = 0; cc < cn; ++cc) {
            const float2 r = cconj(buf[p * T::ld + cc]);
            acc[s] = cadd(acc[s], a.maps ? cmulc(a.maps[(int64_t)(c0 + cc) * N + rowvox + x], r) : r);
          }
        }
      }
      __syncthreads();
    }
#pragma unroll
    for (int s = 0; s < kSlots; ++s) {
      const int x = tid + s * kBlock;
      if (x < nx) {
        const float sc = (a.apod_z[z] * a.apod_y[y] * a.apod_x[x]) * a.scale;
        const float2 r = cscale(acc[s], sc);
        float2* o = a.out + (int64_t)e * N + rowvox + x;
        *o = a.accumulate ? cadd(*o, r) : r;
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Inverse pass along x without a coil sum (C == 1, no maps): every grid channel k is its own
// image.  Block = (image row, chunk of KC channels), so the line FFT runs KC channels wide rather
// than one image at a time; the crop, apodisation and scale are those of k_adj_x_crop (the same
// operations in the same order: 0 + conj(X), times (az ay ax) scale).
// ---------------------------------------------------------------------------
template <int KC, int LC>
__global__ void __launch_bounds__(kBlock) k_adj_x_channels(AdjXArgs a) {
  if (gate_closed(a.gate)) return;
  extern __shared__ float2 smem[];
  using T = Tile<KC, LC>;
  const int LF = LC > 0 ? LC : a.ax.M;
  float2* tw = smem;
  float2* buf = smem + LF;
  const int tid = threadIdx.x;
  const int ch = tid % KC, w = tid / KC;
  const int nz = a.gg.n[0], ny = a.gg.n[1], nx = a.gg.n[2];
  const int gz = a.gg.g[0], gy = a.gg.g[1], gx = a.gg.g[2];
  const int row = blockIdx.x;
  const int k0 = blockIdx.y * KC;
  const int cn = min(KC, a.K - k0);
  const int z = row / ny, y = row - (row / ny) * ny;
  const int64_t N = (int64_t)nz * ny * nx;
  for (int i = tid; i < LF; i += kBlock) tw[i] = a.tw[i];
  const int zp = pmod(z - nz / 2, gz), yp = pmod(y - ny / 2, gy);
  const float2* src = a.grid + (((int64_t)zp * gy + yp) * gx) * a.K + k0;
  const int64_t rowvox = ((int64_t)z * ny + y) * nx;
  float2 v[T::PER];
#pragma unroll
  for (int it = 0; it < T::PER; ++it) {
    const int p = w + it * T::W;
    v[it] = (p < gx && ch < cn) ? src[(int64_t)p * a.K + ch] : make_float2(0.f, 0.f);
  }
#pragma unroll
  for (int it = 0; it < T::PER; ++it) {
    const int p = w + it * T::W;
    if (p < gx) buf[p * T::ld + ch] = cconj(v[it]);
  }
  __syncthreads();
  fft_any<T::W, LC>(buf, T::ld, ch, w, a.ax, tw);
  __syncthreads();
  for (int i = tid; i < cn * nx; i += kBlock) {
    const int cc = i / nx, x = i - cc * nx;
    const int p = pmod(x - nx / 2, gx);
    const float sc = (a.apod_z[z] * a.apod_y[y] * a.apod_x[x]) * a.scale;
    const float2 r = cscale(cadd(make_float2(0.f, 0.f), cconj(buf[p * T::ld + cc])), sc);
    float2* o = a.out + (int64_t)(k0 + cc) * N + rowvox + x;
    *o = a.accumulate ? cadd(*o, r) : r;
  }
}

struct FwdXLaunch {
  template <int KC, int LC>
  static void go(dim3 g, size_t sm, hipStream_t s, const FwdXArgs& a) {
    hipLaunchKernelGGL((k_fwd_x_pad<KC, LC>), g, dim3(kBlock), sm, s, a);
  }
};

struct AdjXLaunch {
  template <int KC, int LC>
  static void go(dim3 g, size_t sm, hipStream_t s, const AdjXArgs& a) {
    hipLaunchKernelGGL((k_adj_x_crop<KC, LC>), g, dim3(kBlock), sm, s, a);
  }
};

int launch_fwd_x(int kc, int L, dim3 grid, size_t smem, hipStream_t s, const FwdXArgs& a0) {
  FwdXArgs a = a0;
  a.gate = gate_current();
  return launch_fft_kernel<FwdXLaunch>(kc, L, grid, smem, s, a);
}

struct AdjXChannelsLaunch {
  template <int KC, int LC>
  static void go(dim3 g, size_t sm, hipStream_t s, const AdjXArgs& a) {
    hipLaunchKernelGGL((k_adj_x_channels<KC, LC>), g, dim3(kBlock), sm, s, a);
  }
};

int launch_adj_x_channels(int kc, int L, dim3 grid, size_t smem, hipStream_t s, const AdjXArgs& a0) {
  AdjXArgs a = a0;
  a.gate = gate_current();
  return launch_fft_kernel<AdjXChannelsLaunch>(kc, L, grid, smem, s, a);
}

int launch_adj_x(int kc, int L, dim3 grid, size_t smem, hipStream_t s, const AdjXArgs& a0) {
  AdjXArgs a = a0;
  a.gate = gate_current();
  return launch_fft_kernel<AdjXLaunch>(kc, L, grid, smem, s, a);
}

}  // namespace tr
