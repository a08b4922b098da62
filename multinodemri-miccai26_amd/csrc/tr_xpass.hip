// Apodise / pad / FFT along x (forward) and inverse FFT / crop / coil sum
// along x (adjoint), one image row per workgroup.

#include "tr_internal.hpp"

namespace tr {

// ---------------------------------------------------------------------------
// Forward pass along x: apodise, coil-weight, zero-pad and FFT one image row.
// Block = (coil chunk, image row); the block walks all echoes, so the coil
// maps of the row are read once.
// ---------------------------------------------------------------------------
template <int KC, int LC>
__global__ void __launch_bounds__(kBlock) k_fwd_x_pad(FwdXArgs a) {
  extern __shared__ float2 smem[];
  using T = Tile<KC, LC>;
  const int L = LC > 0 ? LC : a.ax.L;
  float2* tw = smem;
  float2* buf = smem + L;
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
  for (int i = tid; i < L; i += kBlock) tw[i] = a.tw[i];
  const int64_t rowvox = ((int64_t)z * ny + y) * nx;
  const float azy = a.apod_z[z] * a.apod_y[y] * a.scale;
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
  const int E = a.K / a.C;
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
  extern __shared__ float2 smem[];
  using T = Tile<KC, LC>;
  constexpr int kSlots = (kMaxFftLen + kBlock - 1) / kBlock;
  constexpr int kMaxE = 8;
  const int L = LC > 0 ? LC : a.ax.L;
  float2* tw = smem;
  float2* buf = smem + L;
  const int tid = threadIdx.x;
  const int ch = tid % KC, w = tid / KC;
  const int nz = a.gg.n[0], ny = a.gg.n[1], nx = a.gg.n[2];
  const int gz = a.gg.g[0], gy = a.gg.g[1], gx = a.gg.g[2];
  const int row = blockIdx.x;
  const int z = row / ny, y = row - (row / ny) * ny;
  const int64_t N = (int64_t)nz * ny * nx;
  const int E = a.K / a.C;
  for (int i = tid; i < L; i += kBlock) tw[i] = a.tw[i];
  const int zp = pmod(z - nz / 2, gz), yp = pmod(y - ny / 2, gy);
  const float2* src = a.grid + (((int64_t)zp * gy + yp) * gx) * a.K;
  const int64_t rowvox = ((int64_t)z * ny + y) * nx;
  // Host guarantees E <= kMaxE.
  float2 acc[kMaxE][kSlots];
#pragma unroll
  for (int e = 0; e < kMaxE; ++e)
#pragma unroll
    for (int s = 0; s < kSlots; ++s) acc[e][s] = make_float2(0.f, 0.f);
  for (int c0 = 0; c0 < a.C; c0 += KC) {
    const int cn = min(KC, a.C - c0);
#pragma unroll
    for (int e = 0; e < kMaxE; ++e) {
      if (e < E) {
        float2 v[T::PER];
#pragma unroll
        for (int it = 0; it < T::PER; ++it) {
          const int p = w + it * T::W;
          v[it] = (p < gx && ch < cn) ? src[(int64_t)p * a.K + (int64_t)e * a.C + c0 + ch]
                                       : make_float2(0.f, 0.f);
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
              acc[e][s] = cadd(acc[e][s],
                               a.maps ? cmulc(a.maps[(int64_t)(c0 + cc) * N + rowvox + x], r) : r);
            }
          }
        }
        __syncthreads();
      }
    }
  }
#pragma unroll
  for (int s = 0; s < kSlots; ++s) {
    const int x = tid + s * kBlock;
    if (x < nx) {
      const float sc = (a.apod_z[z] * a.apod_y[y] * a.apod_x[x]) * a.scale;
#pragma unroll
      for (int e = 0; e < kMaxE; ++e) {
        if (e < E) {
          const float2 r = cscale(acc[e][s], sc);
          float2* o = a.out + (int64_t)e * N + rowvox + x;
          *o = a.accumulate ? cadd(*o, r) : r;
        }
      }
    }
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

int launch_fwd_x(int kc, int L, dim3 grid, size_t smem, hipStream_t s, const FwdXArgs& a) {
  return launch_fft_kernel<FwdXLaunch>(kc, L, grid, smem, s, a);
}

int launch_adj_x(int kc, int L, dim3 grid, size_t smem, hipStream_t s, const AdjXArgs& a) {
  return launch_fft_kernel<AdjXLaunch>(kc, L, grid, smem, s, a);
}

}  // namespace tr
