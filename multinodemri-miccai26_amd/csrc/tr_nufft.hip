// 3-D Kaiser-Bessel NUFFT (SigPy 0.1.27 semantics, width 4, oversampling 1.25)
// for MI355X: channel-innermost grids, LDS Stockham FFT passes with pruned
// zero-padding, a gather interpolation fused with the PDHG dual update, and a
// deterministic CSR gather for the adjoint spreading.
//
// Grid layout in HBM for one motion bin: [gz][gy][gx][K] complex64 with
// K = E * C channels (k = e * C + c), in FFT-natural order: image voxel m maps
// to position (m - n//2) mod g and SigPy grid index t to (t - g//2) mod g, which
// folds the centred resize / ifftshift / fftshift of sp.nufft into indexing.
// Samples of a bin are kept sorted by 8^3 grid tile ("sorted order");
// per-sample data of the fused solver path lives as [M_b][K].

#include <hipcub/hipcub.hpp>

#include <cmath>
#include <vector>

#include "toporecon_hip.h"
#include "tr_common.hpp"
#include "tr_fft.hpp"

namespace tr {

static thread_local std::string g_last_error;
void set_error(const std::string& msg) { g_last_error = msg; }
const char* last_error() { return g_last_error.c_str(); }

constexpr int kWidth = 4;
constexpr int kTile = 8;  // sample sort tile edge (grid points)
constexpr int kMaxKc = 16;  // channels per FFT workgroup

struct BinData {
  int64_t M = 0;
  int64_t nnz = 0;
  int* perm = nullptr;      // sorted position -> original sample index
  float4* kp = nullptr;     // sorted scaled coordinates (kz', ky', kx', 0)
  int* row_ptr = nullptr;   // G + 1, over FFT-natural grid index
  int* col = nullptr;       // nnz sorted-sample indices
  float* val = nullptr;     // nnz weights (already divided by width^3)
};

}  // namespace tr

struct tr_plan {
  int n[3];   // image z, y, x
  int g[3];   // oversampled grid
  int64_t N, G;
  float beta;
  float coord_scale[3];
  float coord_shift[3];
  float pad_scale;  // 1/sqrt(N)
  tr::FftAxis ax[3];
  float2* tw[3] = {nullptr, nullptr, nullptr};
  float* apod[3] = {nullptr, nullptr, nullptr};
  std::vector<tr::BinData> bins;
  float2* work = nullptr;  // grid workspace, G * work_k complex
  int64_t work_k = 0;
};

namespace tr {

// ---------------------------------------------------------------------------
// Kernel arguments
// ---------------------------------------------------------------------------
struct GridGeom {
  int n[3];
  int g[3];
};

struct FwdXArgs {
  GridGeom gg;
  FftAxis ax;
  const float2* tw;
  const float2* img;   // (E, Z, Y, X) for this bin
  const float2* maps;  // (C, Z, Y, X) or null (C == 1)
  const float* apod_z;
  const float* apod_y;
  const float* apod_x;
  float scale;
  int C;
  int K;  // grid channel count (E * C)
  float2* grid;
};

struct LinesArgs {
  FftAxis ax;
  const float2* tw;
  float2* data;
  int K;
  int64_t pos_stride;
  int n_inner;
  int64_t inner_stride;
  int64_t outer_stride;
  int outer_win_n;  // >0: outer index o maps to (o - n//2) mod outer_win_g
  int outer_win_g;
  int load_n;   // < L: windowed load (positions (s - n//2) mod L), zeros elsewhere
  int store_n;  // < L: windowed store
  int inverse;
};

struct AdjXArgs {
  GridGeom gg;
  FftAxis ax;
  const float2* tw;
  const float2* grid;
  const float2* maps;  // (C, Z, Y, X) or null (C == 1)
  const float* apod_z;
  const float* apod_y;
  const float* apod_x;
  float scale;
  int C;
  int K;
  float2* out;  // (E, Z, Y, X)
  int accumulate;
};

enum InterpMode : int { kInterpPlain = 0, kInterpDual = 1, kInterpSim = 2 };

struct InterpArgs {
  GridGeom gg;
  float beta;
  const float4* kp;
  int64_t M;
  int K;
  const float2* grid;
  int mode;
  // plain: out[k * M + perm[m]]
  const int* perm;
  float2* out;
  // dual / sim: y[m * K + k], q[m * K + k]
  const float2* y;
  float2* q;
  float sigma;
  float one_plus_sigma;
};

struct SpreadArgs {
  int64_t G;
  int K;
  int64_t M;
  const int* row_ptr;
  const int* col;
  const float* val;
  int sorted_src;  // 1: q (sorted, [M][K]); 0: ksp (original order, [K][M])
  // sorted source: q[m * K + k] (optionally times dcf_sorted[m])
  const float2* q;
  const float* dcf_sorted;
  // plain source: ksp[k * M + perm[m]] (optionally times dcf_orig[perm[m]])
  const float2* ksp;
  const float* dcf_orig;
  const int* perm;
  float2* grid;
};

// ---------------------------------------------------------------------------
// Forward pass along x: apodise, coil-weight, zero-pad and FFT one image row.
// Block = (image row z*ny + y, echo e, coil chunk).
// ---------------------------------------------------------------------------
template <int KC>
__global__ void __launch_bounds__(kBlock) k_fwd_x_pad(FwdXArgs a) {
  extern __shared__ float2 smem[];
  constexpr int ld = lds_ld(KC);
  constexpr int W = kBlock / KC;
  const int L = a.ax.L;
  float2* tw = smem;
  float2* buf = smem + L;
  const int tid = threadIdx.x;
  const int nz = a.gg.n[0], ny = a.gg.n[1], nx = a.gg.n[2];
  const int gz = a.gg.g[0], gy = a.gg.g[1], gx = a.gg.g[2];
  const int row = blockIdx.x;
  const int z = row / ny, y = row - (row / ny) * ny;
  const int e = blockIdx.y;
  const int c0 = blockIdx.z * KC;
  const int64_t N = (int64_t)nz * ny * nx;
  for (int i = tid; i < L; i += kBlock) tw[i] = a.tw[i];
  const float azy = a.apod_z[z] * a.apod_y[y];
  const int64_t rowvox = ((int64_t)z * ny + y) * nx;
  for (int i = tid; i < KC * nx; i += kBlock) {
    const int ch = i / nx;
    const int x = i - ch * nx;
    const int c = c0 + ch;
    float2 v = make_float2(0.f, 0.f);
    if (c < a.C) {
      float2 u = a.img[(int64_t)e * N + rowvox + x];
      if (a.maps) u = cmul(u, a.maps[(int64_t)c * N + rowvox + x]);
      v = cscale(u, (azy * a.apod_x[x]) * a.scale);
    }
    buf[pmod(x - nx / 2, gx) * ld + ch] = v;
  }
  for (int i = tid; i < KC * (gx - nx); i += kBlock) {
    const int ch = i % KC;
    const int p = (nx - nx / 2) + i / KC;
    buf[p * ld + ch] = make_float2(0.f, 0.f);
  }
  __syncthreads();
  fft_line<W>(buf, ld, tid % KC, tid / KC, a.ax, tw);
  const int zp = pmod(z - nz / 2, gz), yp = pmod(y - ny / 2, gy);
  float2* dst = a.grid + (((int64_t)zp * gy + yp) * gx) * a.K + (int64_t)e * a.C;
  for (int i = tid; i < KC * gx; i += kBlock) {
    const int ch = i % KC;
    const int p = i / KC;
    const int c = c0 + ch;
    if (c < a.C) dst[(int64_t)p * a.K + c] = buf[p * ld + ch];
  }
}

// ---------------------------------------------------------------------------
// Generic pass along y or z: load a (possibly windowed) line of KC channels,
// FFT (or inverse), store a (possibly windowed) set of positions.
// Block = (line, channel chunk).
// ---------------------------------------------------------------------------
template <int KC>
__global__ void __launch_bounds__(kBlock) k_fft_lines(LinesArgs a) {
  extern __shared__ float2 smem[];
  constexpr int ld = lds_ld(KC);
  constexpr int W = kBlock / KC;
  const int L = a.ax.L;
  float2* tw = smem;
  float2* buf = smem + L;
  const int tid = threadIdx.x;
  const int nchunk = (a.K + KC - 1) / KC;
  const int line = blockIdx.x / nchunk;
  const int chunk = blockIdx.x - line * nchunk;
  const int o = line / a.n_inner;
  const int in = line - o * a.n_inner;
  const int ob = a.outer_win_n > 0 ? pmod(o - a.outer_win_n / 2, a.outer_win_g) : o;
  const int k0 = chunk * KC;
  float2* base = a.data + ob * a.outer_stride + in * a.inner_stride + k0;
  for (int i = tid; i < L; i += kBlock) tw[i] = a.tw[i];
  const bool win_load = a.load_n < L;
  for (int i = tid; i < KC * L; i += kBlock) {
    const int ch = i % KC;
    const int s = i / KC;
    float2 v = make_float2(0.f, 0.f);
    int p;
    if (s < a.load_n) {
      p = win_load ? pmod(s - a.load_n / 2, L) : s;
      if (k0 + ch < a.K) {
        v = base[p * a.pos_stride + ch];
        if (a.inverse) v.y = -v.y;
      }
    } else {
      p = (a.load_n - a.load_n / 2) + (s - a.load_n);
    }
    buf[p * ld + ch] = v;
  }
  __syncthreads();
  fft_line<W>(buf, ld, tid % KC, tid / KC, a.ax, tw);
  const bool win_store = a.store_n < L;
  for (int i = tid; i < KC * a.store_n; i += kBlock) {
    const int ch = i % KC;
    const int s = i / KC;
    const int p = win_store ? pmod(s - a.store_n / 2, L) : s;
    if (k0 + ch < a.K) {
      float2 v = buf[p * ld + ch];
      if (a.inverse) v.y = -v.y;
      base[p * a.pos_stride + ch] = v;
    }
  }
}

// ---------------------------------------------------------------------------
// Adjoint pass along x: inverse FFT of one grid row for all coils of an echo,
// crop, apodise, multiply by conj(map) and sum the coils in order.
// Block = (image row z*ny + y, echo e); all coils in one chunk.
// ---------------------------------------------------------------------------
template <int KC>
__global__ void __launch_bounds__(kBlock) k_adj_x_crop(AdjXArgs a) {
  extern __shared__ float2 smem[];
  constexpr int ld = lds_ld(KC);
  constexpr int W = kBlock / KC;
  constexpr int kSlots = (kMaxFftLen + kBlock - 1) / kBlock;
  const int L = a.ax.L;
  float2* tw = smem;
  float2* buf = smem + L;
  const int tid = threadIdx.x;
  const int nz = a.gg.n[0], ny = a.gg.n[1], nx = a.gg.n[2];
  const int gz = a.gg.g[0], gy = a.gg.g[1], gx = a.gg.g[2];
  const int row = blockIdx.x;
  const int z = row / ny, y = row - (row / ny) * ny;
  const int e = blockIdx.y;
  const int64_t N = (int64_t)nz * ny * nx;
  for (int i = tid; i < L; i += kBlock) tw[i] = a.tw[i];
  const int zp = pmod(z - nz / 2, gz), yp = pmod(y - ny / 2, gy);
  const float2* src = a.grid + (((int64_t)zp * gy + yp) * gx) * a.K + (int64_t)e * a.C;
  const int64_t rowvox = ((int64_t)z * ny + y) * nx;
  float2 acc[kSlots];
#pragma unroll
  for (int s = 0; s < kSlots; ++s) acc[s] = make_float2(0.f, 0.f);
  // Coils are summed in increasing order, one chunk of KC coils at a time.
  for (int c0 = 0; c0 < a.C; c0 += KC) {
    for (int i = tid; i < KC * gx; i += kBlock) {
      const int ch = i % KC;
      const int p = i / KC;
      float2 v = make_float2(0.f, 0.f);
      if (c0 + ch < a.C) {
        v = src[(int64_t)p * a.K + c0 + ch];
        v.y = -v.y;
      }
      buf[p * ld + ch] = v;
    }
    __syncthreads();
    fft_line<W>(buf, ld, tid % KC, tid / KC, a.ax, tw);
#pragma unroll
    for (int s = 0; s < kSlots; ++s) {
      const int x = tid + s * kBlock;
      if (x < nx) {
        const int p = pmod(x - nx / 2, gx);
        const int cn = min(KC, a.C - c0);
        for (int ch = 0; ch < cn; ++ch) {
          const float2 v = cconj(buf[p * ld + ch]);
          if (a.maps) {
            acc[s] = cadd(acc[s], cmulc(a.maps[(int64_t)(c0 + ch) * N + rowvox + x], v));
          } else {
            acc[s] = cadd(acc[s], v);
          }
        }
      }
    }
    __syncthreads();
  }
  const float azy = a.apod_z[z] * a.apod_y[y];
#pragma unroll
  for (int s = 0; s < kSlots; ++s) {
    const int x = tid + s * kBlock;
    if (x < nx) {
      const float2 r = cscale(acc[s], (azy * a.apod_x[x]) * a.scale);
      float2* o = a.out + (int64_t)e * N + rowvox + x;
      *o = a.accumulate ? cadd(*o, r) : r;
    }
  }
}

// ---------------------------------------------------------------------------
// Kaiser-Bessel gather at sorted samples.  Lanes own channels (KCI per
// sample, 64/KCI samples per wave, NCI channel groups per lane).
// ---------------------------------------------------------------------------
template <int KCI, int NCI>
__global__ void __launch_bounds__(256) k_interp(InterpArgs a) {
  constexpr int SPW = 64 / KCI;
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int64_t m = ((int64_t)blockIdx.x * 4 + wave) * SPW + lane / KCI;
  const int ch = lane % KCI;
  if (m >= a.M) return;
  const int gz = a.gg.g[0], gy = a.gg.g[1], gx = a.gg.g[2];
  const float4 kp = a.kp[m];
  float wz[5], wy[5], wx[5];
  int pz[5], py[5], px[5];
  const int nz = axis_taps(kp.x, gz, a.beta, 1.0f, wz, pz);
  const int ny = axis_taps(kp.y, gy, a.beta, 1.0f, wy, py);
  const int nx = axis_taps(kp.z, gx, a.beta, 1.0f / (kWidth * kWidth * kWidth), wx, px);
  float2 acc[NCI];
#pragma unroll
  for (int i = 0; i < NCI; ++i) acc[i] = make_float2(0.f, 0.f);
  const int K = a.K;
#pragma unroll
  for (int dz = 0; dz < 5; ++dz) {
    if (dz < nz) {
#pragma unroll
      for (int dy = 0; dy < 5; ++dy) {
        if (dy < ny) {
          const float wzy = wz[dz] * wy[dy];
          const int64_t rb = ((int64_t)pz[dz] * gy + py[dy]) * gx;
#pragma unroll
          for (int dx = 0; dx < 5; ++dx) {
            if (dx < nx) {
              const float w = wzy * wx[dx];
              const float2* gp = a.grid + (rb + px[dx]) * K + ch;
#pragma unroll
              for (int i = 0; i < NCI; ++i) {
                if (i * KCI + ch < K) {
                  const float2 v = gp[i * KCI];
                  acc[i].x += w * v.x;
                  acc[i].y += w * v.y;
                }
              }
            }
          }
        }
      }
    }
  }
#pragma unroll
  for (int i = 0; i < NCI; ++i) {
    const int k = i * KCI + ch;
    if (k >= K) continue;
    if (a.mode == kInterpPlain) {
      a.out[(int64_t)k * a.M + a.perm[m]] = acc[i];
    } else if (a.mode == kInterpDual) {
      const int64_t idx = m * K + k;
      const float2 r = csub(acc[i], a.y[idx]);
      const float2 qo = a.q[idx];
      const float2 t = make_float2(qo.x + a.sigma * r.x, qo.y + a.sigma * r.y);
      a.q[idx] = make_float2(t.x / a.one_plus_sigma, t.y / a.one_plus_sigma);
    } else {
      a.q[m * K + k] = acc[i];
    }
  }
}

// ---------------------------------------------------------------------------
// Deterministic adjoint spreading: every grid point gathers its CSR row of
// (sample, weight) in sorted-sample order.  Lanes own channels.
// ---------------------------------------------------------------------------
template <int KCI, int NCI>
__global__ void __launch_bounds__(256) k_spread(SpreadArgs a) {
  constexpr int SPW = 64 / KCI;
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int64_t t = ((int64_t)blockIdx.x * 4 + wave) * SPW + lane / KCI;
  const int ch = lane % KCI;
  if (t >= a.G) return;
  const int K = a.K;
  const int j0 = a.row_ptr[t], j1 = a.row_ptr[t + 1];
  float2 acc[NCI];
#pragma unroll
  for (int i = 0; i < NCI; ++i) acc[i] = make_float2(0.f, 0.f);
  for (int j = j0; j < j1; ++j) {
    const int m = a.col[j];
    const float w = a.val[j];
    if (a.sorted_src) {
      const float d = a.dcf_sorted ? a.dcf_sorted[m] : 1.0f;
#pragma unroll
      for (int i = 0; i < NCI; ++i) {
        const int k = i * KCI + ch;
        if (k < K) {
          float2 v = a.q[(int64_t)m * K + k];
          if (a.dcf_sorted) v = cscale(v, d);
          acc[i].x += w * v.x;
          acc[i].y += w * v.y;
        }
      }
    } else {
      const int mo = a.perm[m];
      const float d = a.dcf_orig ? a.dcf_orig[mo] : 1.0f;
#pragma unroll
      for (int i = 0; i < NCI; ++i) {
        const int k = i * KCI + ch;
        if (k < K) {
          float2 v = a.ksp[(int64_t)k * a.M + mo];
          if (a.dcf_orig) v = cscale(v, d);
          acc[i].x += w * v.x;
          acc[i].y += w * v.y;
        }
      }
    }
  }
#pragma unroll
  for (int i = 0; i < NCI; ++i) {
    const int k = i * KCI + ch;
    if (k < K) a.grid[t * K + k] = acc[i];
  }
}

// ---------------------------------------------------------------------------
// Plan construction kernels
// ---------------------------------------------------------------------------
__global__ void k_scale_keys(const float* coord, int64_t M, float3 s, float3 sh, int3 g, int3 nt,
                             float4* kp, unsigned* key, int* idx) {
  const int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= M) return;
  // k' = fl32(fl32(k * s) + shift), as SigPy's _scale_coord on a float32 trajectory.
  const float kz = __fadd_rn(__fmul_rn(coord[3 * m + 0], s.x), sh.x);
  const float ky = __fadd_rn(__fmul_rn(coord[3 * m + 1], s.y), sh.y);
  const float kx = __fadd_rn(__fmul_rn(coord[3 * m + 2], s.z), sh.z);
  kp[m] = make_float4(kz, ky, kx, 0.f);
  auto tile = [](float k, int g, int n) {
    int t = (int)floorf((k + 2.0f) / (float)kTile);
    return t < 0 ? 0 : (t >= n ? n - 1 : t);
  };
  const int tz = tile(kz, g.x, nt.x), ty = tile(ky, g.y, nt.y), tx = tile(kx, g.z, nt.z);
  key[m] = (unsigned)((tz * nt.y + ty) * nt.z + tx);
  idx[m] = (int)m;
}

__global__ void k_gather_kp(const float4* kp_in, const int* perm, int64_t M, float4* kp_out) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j < M) kp_out[j] = kp_in[perm[j]];
}

__device__ __forceinline__ int tap_count(float k) {
  if (!isfinite(k) || fabsf(k) > 1.0e6f) return 0;  // same rule as axis_taps
  const double kd = (double)k;
  return (int)floor(kd + 2.0) - (int)ceil(kd - 2.0) + 1;
}

__global__ void k_count_taps(const float4* kp, int64_t M, int* cnt) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= M) return;
  const float4 k = kp[j];
  cnt[j] = tap_count(k.x) * tap_count(k.y) * tap_count(k.z);
}

__global__ void k_fill_pairs(const float4* kp, int64_t M, const int* off, int3 g, float beta,
                             unsigned* key, unsigned long long* val) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= M) return;
  const float4 k = kp[j];
  float wz[5], wy[5], wx[5];
  int pz[5], py[5], px[5];
  const int nz = axis_taps(k.x, g.x, beta, 1.0f, wz, pz);
  const int ny = axis_taps(k.y, g.y, beta, 1.0f, wy, py);
  const int nx = axis_taps(k.z, g.z, beta, 1.0f / (kWidth * kWidth * kWidth), wx, px);
  int p = off[j];
  for (int dz = 0; dz < nz; ++dz)
    for (int dy = 0; dy < ny; ++dy) {
      const float wzy = wz[dz] * wy[dy];
      for (int dx = 0; dx < nx; ++dx) {
        const float w = wzy * wx[dx];
        key[p] = (unsigned)(((int64_t)pz[dz] * g.y + py[dy]) * g.z + px[dx]);
        val[p] = ((unsigned long long)(unsigned)j << 32) | (unsigned long long)__float_as_uint(w);
        ++p;
      }
    }
}

__global__ void k_unpack_pairs(const unsigned long long* val, int64_t nnz, int* col, float* w) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nnz) return;
  const unsigned long long v = val[i];
  col[i] = (int)(v >> 32);
  w[i] = __uint_as_float((unsigned)(v & 0xffffffffull));
}

__global__ void k_row_ptr(const unsigned* key, int64_t nnz, int64_t G, int* row_ptr) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t > G) return;
  // lower_bound(key, t)
  int64_t lo = 0, hi = nnz;
  while (lo < hi) {
    const int64_t mid = (lo + hi) >> 1;
    if ((int64_t)key[mid] < t) lo = mid + 1; else hi = mid;
  }
  row_ptr[t] = (int)lo;
}

template <typename T>
__global__ void k_gather_samples(const T* src, const int* perm, int64_t M, int K, T* dst) {
  // dst[j * K + k] = src[k * M + perm[j]]
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= M * K) return;
  const int64_t j = i / K;
  const int k = (int)(i - j * K);
  dst[i] = src[(int64_t)k * M + perm[j]];
}

template <typename T>
__global__ void k_scatter_samples(const T* src, const int* perm, int64_t M, int K, T* dst) {
  // dst[k * M + perm[j]] = src[j * K + k]
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= M * K) return;
  const int64_t j = i / K;
  const int k = (int)(i - j * K);
  dst[(int64_t)k * M + perm[j]] = src[i];
}

// ---------------------------------------------------------------------------
// Host helpers
// ---------------------------------------------------------------------------
static inline unsigned grid1d(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

static int pow2ceil(int v) {
  int p = 1;
  while (p < v) p <<= 1;
  return p;
}

static int ensure_work(tr_plan* p, int64_t K) {
  if (p->work_k >= K) return kOk;
  if (p->work) {
    TR_HIP(hipDeviceSynchronize());
    TR_HIP(hipFree(p->work));
    p->work = nullptr;
    p->work_k = 0;
  }
  TR_HIP(hipMalloc(&p->work, sizeof(float2) * p->G * K));
  p->work_k = K;
  return kOk;
}

#define TR_DISPATCH_KC(KERNEL, kc, grid, smem, stream, args)                                    \
  do {                                                                                        \
    switch (kc) {                                                                             \
      case 1: hipLaunchKernelGGL(KERNEL<1>, grid, dim3(kBlock), smem, stream, args); break;   \
      case 2: hipLaunchKernelGGL(KERNEL<2>, grid, dim3(kBlock), smem, stream, args); break;   \
      case 4: hipLaunchKernelGGL(KERNEL<4>, grid, dim3(kBlock), smem, stream, args); break;   \
      case 8: hipLaunchKernelGGL(KERNEL<8>, grid, dim3(kBlock), smem, stream, args); break;   \
      case 16: hipLaunchKernelGGL(KERNEL<16>, grid, dim3(kBlock), smem, stream, args); break; \
      default: TR_FAIL(kUnsupported, "channel chunk must be a power of two <= 16");          \
    }                                                                                         \
  } while (0)

// Channel chunk of the FFT passes: at most 16 channels per workgroup keeps the
// Stockham stages under ~130 VGPRs and the LDS tile at L * 17 * 8 bytes.
static int fft_kc(int K) { return K >= kMaxKc ? kMaxKc : pow2ceil(K); }

static size_t fft_smem(int L, int kc) { return sizeof(float2) * (size_t)L * (1 + lds_ld(kc)); }


static GridGeom geom(const tr_plan* p) {
  GridGeom g;
  for (int i = 0; i < 3; ++i) {
    g.n[i] = p->n[i];
    g.g[i] = p->g[i];
  }
  return g;
}

// Forward passes x (with pad), y, z into p->work for K = E*C channels.
static int forward_to_grid(tr_plan* p, const float2* img, const float2* maps, int E, int C,
                           hipStream_t s) {
  const int K = E * C;
  int r = ensure_work(p, K);
  if (r) return r;
  const int nz = p->n[0], ny = p->n[1];
  const int gz = p->g[0], gy = p->g[1], gx = p->g[2];
  {
    FwdXArgs a;
    a.gg = geom(p);
    a.ax = p->ax[2];
    a.tw = p->tw[2];
    a.img = img;
    a.maps = maps;
    a.apod_z = p->apod[0];
    a.apod_y = p->apod[1];
    a.apod_x = p->apod[2];
    a.scale = p->pad_scale;
    a.C = C;
    a.K = K;
    a.grid = p->work;
    const int kc = fft_kc(C);
    dim3 grid(nz * ny, E, (C + kc - 1) / kc);
    TR_DISPATCH_KC(k_fwd_x_pad, kc, grid, fft_smem(gx, kc), s, a);
    TR_LAUNCH_CHECK();
  }
  const int kc = fft_kc(K);
  const int nchunk = (K + kc - 1) / kc;
  {  // y pass over the nz occupied z-planes, windowed load along y
    LinesArgs a;
    a.ax = p->ax[1];
    a.tw = p->tw[1];
    a.data = p->work;
    a.K = K;
    a.pos_stride = (int64_t)gx * K;
    a.n_inner = gx;
    a.inner_stride = K;
    a.outer_stride = (int64_t)gy * gx * K;
    a.outer_win_n = nz;
    a.outer_win_g = gz;
    a.load_n = ny;
    a.store_n = gy;
    a.inverse = 0;
    dim3 grid((unsigned)(nz * gx * nchunk));
    TR_DISPATCH_KC(k_fft_lines, kc, grid, fft_smem(gy, kc), s, a);
    TR_LAUNCH_CHECK();
  }
  {  // z pass over every (y, x) line, windowed load along z
    LinesArgs a;
    a.ax = p->ax[0];
    a.tw = p->tw[0];
    a.data = p->work;
    a.K = K;
    a.pos_stride = (int64_t)gy * gx * K;
    a.n_inner = gx;
    a.inner_stride = K;
    a.outer_stride = (int64_t)gx * K;
    a.outer_win_n = 0;
    a.outer_win_g = gy;
    a.load_n = nz;
    a.store_n = gz;
    a.inverse = 0;
    dim3 grid((unsigned)(gy * gx * nchunk));
    TR_DISPATCH_KC(k_fft_lines, kc, grid, fft_smem(gz, kc), s, a);
    TR_LAUNCH_CHECK();
  }
  return kOk;
}

// Inverse passes z, y (cropping) on p->work, then x with crop/coil sum into out.
static int grid_to_image(tr_plan* p, const float2* maps, int E, int C, float2* out, int accumulate,
                         hipStream_t s) {
  const int K = E * C;
  const int nz = p->n[0], ny = p->n[1];
  const int gz = p->g[0], gy = p->g[1], gx = p->g[2];
  const int kc = fft_kc(K);
  const int nchunk = (K + kc - 1) / kc;
  {
    LinesArgs a;
    a.ax = p->ax[0];
    a.tw = p->tw[0];
    a.data = p->work;
    a.K = K;
    a.pos_stride = (int64_t)gy * gx * K;
    a.n_inner = gx;
    a.inner_stride = K;
    a.outer_stride = (int64_t)gx * K;
    a.outer_win_n = 0;
    a.outer_win_g = gy;
    a.load_n = gz;
    a.store_n = nz;
    a.inverse = 1;
    dim3 grid((unsigned)(gy * gx * nchunk));
    TR_DISPATCH_KC(k_fft_lines, kc, grid, fft_smem(gz, kc), s, a);
    TR_LAUNCH_CHECK();
  }
  {
    LinesArgs a;
    a.ax = p->ax[1];
    a.tw = p->tw[1];
    a.data = p->work;
    a.K = K;
    a.pos_stride = (int64_t)gx * K;
    a.n_inner = gx;
    a.inner_stride = K;
    a.outer_stride = (int64_t)gy * gx * K;
    a.outer_win_n = nz;
    a.outer_win_g = gz;
    a.load_n = gy;
    a.store_n = ny;
    a.inverse = 1;
    dim3 grid((unsigned)(nz * gx * nchunk));
    TR_DISPATCH_KC(k_fft_lines, kc, grid, fft_smem(gy, kc), s, a);
    TR_LAUNCH_CHECK();
  }
  {
    AdjXArgs a;
    a.gg = geom(p);
    a.ax = p->ax[2];
    a.tw = p->tw[2];
    a.grid = p->work;
    a.maps = maps;
    a.apod_z = p->apod[0];
    a.apod_y = p->apod[1];
    a.apod_x = p->apod[2];
    a.scale = p->pad_scale;
    a.C = C;
    a.K = K;
    a.out = out;
    a.accumulate = accumulate;
    const int kcc = fft_kc(C);
    dim3 grid(nz * ny, E);
    TR_DISPATCH_KC(k_adj_x_crop, kcc, grid, fft_smem(gx, kcc), s, a);
    TR_LAUNCH_CHECK();
  }
  return kOk;
}

static int launch_interp(const InterpArgs& a, hipStream_t s) {
  const int K = a.K;
  const int kci = K >= 64 ? 64 : pow2ceil(K);
  const int nci = (K + kci - 1) / kci;
  const int spw = 64 / kci;
  const unsigned blocks = grid1d(a.M, 4 * spw);
  if (a.M == 0) return kOk;
#define TR_INTERP(KCI, NCI) hipLaunchKernelGGL((k_interp<KCI, NCI>), dim3(blocks), dim3(256), 0, s, a)
  switch (kci) {
    case 1: TR_INTERP(1, 1); break;
    case 2: TR_INTERP(2, 1); break;
    case 4: TR_INTERP(4, 1); break;
    case 8: TR_INTERP(8, 1); break;
    case 16: TR_INTERP(16, 1); break;
    case 32: TR_INTERP(32, 1); break;
    case 64:
      switch (nci) {
        case 1: TR_INTERP(64, 1); break;
        case 2: TR_INTERP(64, 2); break;
        case 3: TR_INTERP(64, 3); break;
        case 4: TR_INTERP(64, 4); break;
        default: TR_FAIL(kUnsupported, "more than 256 channels per launch");
      }
      break;
  }
#undef TR_INTERP
  TR_LAUNCH_CHECK();
  return kOk;
}

static int launch_spread(const SpreadArgs& a, hipStream_t s) {
  const int K = a.K;
  const int kci = K >= 64 ? 64 : pow2ceil(K);
  const int nci = (K + kci - 1) / kci;
  const int spw = 64 / kci;
  const unsigned blocks = grid1d(a.G, 4 * spw);
#define TR_SPREAD(KCI, NCI) hipLaunchKernelGGL((k_spread<KCI, NCI>), dim3(blocks), dim3(256), 0, s, a)
  switch (kci) {
    case 1: TR_SPREAD(1, 1); break;
    case 2: TR_SPREAD(2, 1); break;
    case 4: TR_SPREAD(4, 1); break;
    case 8: TR_SPREAD(8, 1); break;
    case 16: TR_SPREAD(16, 1); break;
    case 32: TR_SPREAD(32, 1); break;
    case 64:
      switch (nci) {
        case 1: TR_SPREAD(64, 1); break;
        case 2: TR_SPREAD(64, 2); break;
        case 3: TR_SPREAD(64, 3); break;
        case 4: TR_SPREAD(64, 4); break;
        default: TR_FAIL(kUnsupported, "more than 256 channels per launch");
      }
      break;
  }
#undef TR_SPREAD
  TR_LAUNCH_CHECK();
  return kOk;
}

static int build_bin(tr_plan* p, BinData& b, const float* coord, int64_t M, hipStream_t s) {
  b.M = M;
  TR_HIP(hipMalloc(&b.row_ptr, sizeof(int) * (p->G + 1)));
  if (M == 0) {
    TR_HIP(hipMemsetAsync(b.row_ptr, 0, sizeof(int) * (p->G + 1), s));
    return kOk;
  }
  // 125 taps per sample must index an int32 CSR.
  if (M > (int64_t)17000000) TR_FAIL(kUnsupported, "more than 17M samples in one bin");
  float4* kp_raw = nullptr;
  unsigned *key_in = nullptr, *key_out = nullptr;
  int* idx = nullptr;
  TR_HIP(hipMalloc(&kp_raw, sizeof(float4) * M));
  TR_HIP(hipMalloc(&key_in, sizeof(unsigned) * M));
  TR_HIP(hipMalloc(&key_out, sizeof(unsigned) * M));
  TR_HIP(hipMalloc(&idx, sizeof(int) * M));
  TR_HIP(hipMalloc(&b.perm, sizeof(int) * M));
  TR_HIP(hipMalloc(&b.kp, sizeof(float4) * M));
  int3 nt;
  nt.x = p->g[0] / kTile + 2;
  nt.y = p->g[1] / kTile + 2;
  nt.z = p->g[2] / kTile + 2;
  const int3 g3 = make_int3(p->g[0], p->g[1], p->g[2]);
  hipLaunchKernelGGL(k_scale_keys, dim3(grid1d(M, 256)), dim3(256), 0, s, coord, M,
                     make_float3(p->coord_scale[0], p->coord_scale[1], p->coord_scale[2]),
                     make_float3(p->coord_shift[0], p->coord_shift[1], p->coord_shift[2]), g3, nt,
                     kp_raw, key_in, idx);
  TR_LAUNCH_CHECK();
  int key_bits = 1;
  while (((int64_t)1 << key_bits) < (int64_t)nt.x * nt.y * nt.z) ++key_bits;
  size_t tmp_bytes = 0;
  TR_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp_bytes, key_in, key_out, idx, b.perm, (int)M, 0,
                                            key_bits, s));
  void* tmp = nullptr;
  TR_HIP(hipMalloc(&tmp, tmp_bytes));
  TR_HIP(hipcub::DeviceRadixSort::SortPairs(tmp, tmp_bytes, key_in, key_out, idx, b.perm, (int)M, 0,
                                            key_bits, s));
  hipLaunchKernelGGL(k_gather_kp, dim3(grid1d(M, 256)), dim3(256), 0, s, kp_raw, b.perm, M, b.kp);
  TR_LAUNCH_CHECK();
  // tap counts -> offsets
  int* cnt = nullptr;
  int* off = nullptr;
  TR_HIP(hipMalloc(&cnt, sizeof(int) * (M + 1)));
  TR_HIP(hipMalloc(&off, sizeof(int) * (M + 1)));
  TR_HIP(hipMemsetAsync(cnt + M, 0, sizeof(int), s));
  hipLaunchKernelGGL(k_count_taps, dim3(grid1d(M, 256)), dim3(256), 0, s, b.kp, M, cnt);
  TR_LAUNCH_CHECK();
  size_t scan_bytes = 0;
  TR_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, scan_bytes, cnt, off, (int)(M + 1), s));
  void* scan_tmp = nullptr;
  TR_HIP(hipMalloc(&scan_tmp, scan_bytes));
  TR_HIP(hipcub::DeviceScan::ExclusiveSum(scan_tmp, scan_bytes, cnt, off, (int)(M + 1), s));
  int nnz_i = 0;
  TR_HIP(hipMemcpyAsync(&nnz_i, off + M, sizeof(int), hipMemcpyDeviceToHost, s));
  TR_HIP(hipStreamSynchronize(s));
  const int64_t nnz = nnz_i;
  b.nnz = nnz;
  if (nnz == 0) {
    TR_HIP(hipMemsetAsync(b.row_ptr, 0, sizeof(int) * (p->G + 1), s));
    TR_HIP(hipStreamSynchronize(s));
    void* frees0[] = {kp_raw, key_in, key_out, idx, tmp, cnt, off, scan_tmp};
    for (void* f : frees0) TR_HIP(hipFree(f));
    return kOk;
  }
  unsigned *pk_in = nullptr, *pk_out = nullptr;
  unsigned long long *pv_in = nullptr, *pv_out = nullptr;
  TR_HIP(hipMalloc(&pk_in, sizeof(unsigned) * nnz));
  TR_HIP(hipMalloc(&pk_out, sizeof(unsigned) * nnz));
  TR_HIP(hipMalloc(&pv_in, sizeof(unsigned long long) * nnz));
  TR_HIP(hipMalloc(&pv_out, sizeof(unsigned long long) * nnz));
  hipLaunchKernelGGL(k_fill_pairs, dim3(grid1d(M, 256)), dim3(256), 0, s, b.kp, M, off, g3, p->beta,
                     pk_in, pv_in);
  TR_LAUNCH_CHECK();
  int gbits = 1;
  while (((int64_t)1 << gbits) < p->G) ++gbits;
  size_t tmp2_bytes = 0;
  TR_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp2_bytes, pk_in, pk_out, pv_in, pv_out,
                                            (int)nnz, 0, gbits, s));
  void* tmp2 = nullptr;
  TR_HIP(hipMalloc(&tmp2, tmp2_bytes));
  TR_HIP(hipcub::DeviceRadixSort::SortPairs(tmp2, tmp2_bytes, pk_in, pk_out, pv_in, pv_out, (int)nnz,
                                            0, gbits, s));
  TR_HIP(hipMalloc(&b.col, sizeof(int) * nnz));
  TR_HIP(hipMalloc(&b.val, sizeof(float) * nnz));
  hipLaunchKernelGGL(k_unpack_pairs, dim3(grid1d(nnz, 256)), dim3(256), 0, s, pv_out, nnz, b.col,
                     b.val);
  TR_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_row_ptr, dim3(grid1d(p->G + 1, 256)), dim3(256), 0, s, pk_out, nnz, p->G,
                     b.row_ptr);
  TR_LAUNCH_CHECK();
  TR_HIP(hipStreamSynchronize(s));
  void* frees[] = {kp_raw, key_in, key_out, idx, tmp, cnt, off, scan_tmp, pk_in, pk_out, pv_in, pv_out, tmp2};
  for (void* f : frees) TR_HIP(hipFree(f));
  return kOk;
}

static void free_plan(tr_plan* p) {
  if (!p) return;
  for (auto& b : p->bins) {
    (void)hipFree(b.perm);
    (void)hipFree(b.kp);
    (void)hipFree(b.row_ptr);
    (void)hipFree(b.col);
    (void)hipFree(b.val);
  }
  for (int i = 0; i < 3; ++i) {
    (void)hipFree(p->tw[i]);
    (void)hipFree(p->apod[i]);
  }
  (void)hipFree(p->work);
  delete p;
}

static int check_plan_bin(const tr_plan* p, int bin) {
  if (!p) TR_FAIL(kInvalidArgument, "null plan");
  if (bin < 0 || bin >= (int)p->bins.size())
    TR_FAIL(kInvalidArgument, "bin_index " + std::to_string(bin) + " is outside [0, " +
                                  std::to_string(p->bins.size()) + ").");
  return kOk;
}

}  // namespace tr

using namespace tr;

extern "C" {

const char* tr_last_error(void) { return tr::last_error(); }

int tr_version(void) { return TR_ABI_VERSION; }

int tr_plan_create(tr_plan** out, int nz, int ny, int nx, int nbins, const float* const* coords,
                   const int64_t* samples_per_bin, void* stream) {
  if (!out) TR_FAIL(kInvalidArgument, "null output pointer");
  *out = nullptr;
  if (nz < 1 || ny < 1 || nx < 1) TR_FAIL(kInvalidArgument, "img_shape must be three positive values");
  if (nbins < 1) TR_FAIL(kInvalidArgument, "At least one trajectory bin is required.");
  hipStream_t s = (hipStream_t)stream;
  tr_plan* p = new tr_plan();
  const int n[3] = {nz, ny, nx};
  const double beta = M_PI * std::sqrt(std::pow((kWidth / 1.25) * (1.25 - 0.5), 2) - 0.8);
  p->beta = (float)beta;
  p->N = (int64_t)nz * ny * nx;
  p->pad_scale = (float)(1.0 / std::sqrt((double)p->N));
  for (int i = 0; i < 3; ++i) {
    p->n[i] = n[i];
    p->g[i] = (int)std::ceil(1.25 * n[i]);
    p->coord_scale[i] = (float)((double)p->g[i] / n[i]);
    p->coord_shift[i] = (float)(p->g[i] / 2);
    if (!make_fft_axis(p->g[i], &p->ax[i])) {
      free_plan(p);
      TR_FAIL(kUnsupported, "oversampled grid length " + std::to_string(p->g[i]) +
                                " is not supported (max 256, radices 2,3,5,7,11,13)");
    }
  }
  p->G = (int64_t)p->g[0] * p->g[1] * p->g[2];
  for (int i = 0; i < 3; ++i) {
    const int L = p->g[i];
    std::vector<float2> tw(L);
    for (int t = 0; t < L; ++t) {
      const double a = -2.0 * M_PI * (double)t / L;
      tw[t] = make_float2((float)std::cos(a), (float)std::sin(a));
    }
    std::vector<float> ap(n[i]);
    for (int j = 0; j < n[i]; ++j) {
      const double x = M_PI * kWidth * (double)(j - n[i] / 2) / p->g[i];
      const double sarg = std::sqrt(beta * beta - x * x);
      ap[j] = (float)(sarg / std::sinh(sarg));
    }
    if (hipMalloc(&p->tw[i], sizeof(float2) * L) != hipSuccess ||
        hipMalloc(&p->apod[i], sizeof(float) * n[i]) != hipSuccess) {
      free_plan(p);
      TR_FAIL(kOutOfMemory, "allocating plan tables");
    }
    (void)hipMemcpy(p->tw[i], tw.data(), sizeof(float2) * L, hipMemcpyHostToDevice);
    (void)hipMemcpy(p->apod[i], ap.data(), sizeof(float) * n[i], hipMemcpyHostToDevice);
  }
  p->bins.resize(nbins);
  for (int b = 0; b < nbins; ++b) {
    if (samples_per_bin[b] < 0) {
      free_plan(p);
      TR_FAIL(kInvalidArgument, "negative sample count");
    }
    const int r = build_bin(p, p->bins[b], coords[b], samples_per_bin[b], s);
    if (r) {
      std::string msg = last_error();
      free_plan(p);
      set_error(msg);
      return r;
    }
  }
  *out = p;
  return kOk;
}

int tr_plan_destroy(tr_plan* plan) {
  if (plan) {
    (void)hipDeviceSynchronize();
    free_plan(plan);
  }
  return kOk;
}

int tr_plan_info(const tr_plan* p, int bin, int64_t* info) {
  if (!p || !info) TR_FAIL(kInvalidArgument, "null argument");
  for (int i = 0; i < 3; ++i) {
    info[i] = p->n[i];
    info[3 + i] = p->g[i];
  }
  info[6] = (int64_t)p->bins.size();
  if (bin >= 0 && bin < (int)p->bins.size()) {
    info[7] = p->bins[bin].M;
    info[8] = p->bins[bin].nnz;
  } else {
    info[7] = info[8] = -1;
  }
  info[9] = p->work_k;
  return kOk;
}

int tr_plan_release_workspace(tr_plan* p) {
  if (!p) TR_FAIL(kInvalidArgument, "null plan");
  if (p->work) {
    TR_HIP(hipDeviceSynchronize());
    TR_HIP(hipFree(p->work));
  }
  p->work = nullptr;
  p->work_k = 0;
  return kOk;
}

int tr_plan_gather_samples(const tr_plan* p, int bin, const tr_c64* src, int K, tr_c64* dst, void* stream) {
  if (int r = check_plan_bin(p, bin)) return r;
  const BinData& b = p->bins[bin];
  if (b.M == 0 || K == 0) return kOk;
  hipLaunchKernelGGL(k_gather_samples<float2>, dim3(grid1d(b.M * K, 256)), dim3(256), 0,
                     (hipStream_t)stream, F2(src), b.perm, b.M, K, F2(dst));
  TR_LAUNCH_CHECK();
  return kOk;
}

int tr_plan_scatter_samples(const tr_plan* p, int bin, const tr_c64* src, int K, tr_c64* dst, void* stream) {
  if (int r = check_plan_bin(p, bin)) return r;
  const BinData& b = p->bins[bin];
  if (b.M == 0 || K == 0) return kOk;
  hipLaunchKernelGGL(k_scatter_samples<float2>, dim3(grid1d(b.M * K, 256)), dim3(256), 0,
                     (hipStream_t)stream, F2(src), b.perm, b.M, K, F2(dst));
  TR_LAUNCH_CHECK();
  return kOk;
}

int tr_plan_gather_weights(const tr_plan* p, int bin, const float* src, float* dst, void* stream) {
  if (int r = check_plan_bin(p, bin)) return r;
  const BinData& b = p->bins[bin];
  if (b.M == 0) return kOk;
  hipLaunchKernelGGL(k_gather_samples<float>, dim3(grid1d(b.M, 256)), dim3(256), 0,
                     (hipStream_t)stream, src, b.perm, b.M, 1, dst);
  TR_LAUNCH_CHECK();
  return kOk;
}

int tr_nufft_forward(tr_plan* p, int bin, const tr_c64* img, const tr_c64* maps, int E, int C,
                     tr_c64* out, void* stream) {
  if (int r = check_plan_bin(p, bin)) return r;
  if (E < 1 || C < 1 || (!maps && C != 1)) TR_FAIL(kInvalidArgument, "need E >= 1 and C == 1 without maps");
  hipStream_t s = (hipStream_t)stream;
  const BinData& b = p->bins[bin];
  if (b.M == 0) return kOk;
  if (int r = forward_to_grid(p, F2(img), F2(maps), E, C, s)) return r;
  InterpArgs a = {};
  a.gg = geom(p);
  a.beta = p->beta;
  a.kp = b.kp;
  a.M = b.M;
  a.K = E * C;
  a.grid = p->work;
  a.mode = kInterpPlain;
  a.perm = b.perm;
  a.out = F2(out);
  return launch_interp(a, s);
}

int tr_nufft_adjoint(tr_plan* p, int bin, const tr_c64* ksp, const float* dcf, const tr_c64* maps, int E,
                     int C, tr_c64* out, int accumulate, void* stream) {
  if (int r = check_plan_bin(p, bin)) return r;
  if (E < 1 || C < 1 || (!maps && C != 1)) TR_FAIL(kInvalidArgument, "need E >= 1 and C == 1 without maps");
  hipStream_t s = (hipStream_t)stream;
  const BinData& b = p->bins[bin];
  const int K = E * C;
  if (int r = ensure_work(p, K)) return r;
  SpreadArgs a = {};
  a.G = p->G;
  a.K = K;
  a.M = b.M;
  a.row_ptr = b.row_ptr;
  a.col = b.col;
  a.val = b.val;
  a.sorted_src = 0;
  a.ksp = F2(ksp);
  a.dcf_orig = dcf;
  a.perm = b.perm;
  a.grid = p->work;
  if (int r = launch_spread(a, s)) return r;
  return grid_to_image(p, F2(maps), E, C, F2(out), accumulate, s);
}

int tr_data_dual_update(tr_plan* p, int bin, const tr_c64* ubar, const tr_c64* maps, int E, int C,
                        const tr_c64* y_sorted, tr_c64* q_sorted, float sigma, float one_plus_sigma,
                        void* stream) {
  if (int r = check_plan_bin(p, bin)) return r;
  hipStream_t s = (hipStream_t)stream;
  const BinData& b = p->bins[bin];
  if (b.M == 0) return kOk;
  if (int r = forward_to_grid(p, F2(ubar), F2(maps), E, C, s)) return r;
  InterpArgs a = {};
  a.gg = geom(p);
  a.beta = p->beta;
  a.kp = b.kp;
  a.M = b.M;
  a.K = E * C;
  a.grid = p->work;
  a.mode = kInterpDual;
  a.y = F2(y_sorted);
  a.q = F2(q_sorted);
  a.sigma = sigma;
  a.one_plus_sigma = one_plus_sigma;
  return launch_interp(a, s);
}

int tr_data_simulate(tr_plan* p, int bin, const tr_c64* x, const tr_c64* maps, int E, int C,
                     tr_c64* y_sorted, void* stream) {
  if (int r = check_plan_bin(p, bin)) return r;
  hipStream_t s = (hipStream_t)stream;
  const BinData& b = p->bins[bin];
  if (b.M == 0) return kOk;
  if (int r = forward_to_grid(p, F2(x), F2(maps), E, C, s)) return r;
  InterpArgs a = {};
  a.gg = geom(p);
  a.beta = p->beta;
  a.kp = b.kp;
  a.M = b.M;
  a.K = E * C;
  a.grid = p->work;
  a.mode = kInterpSim;
  a.q = F2(y_sorted);
  return launch_interp(a, s);
}

int tr_data_adjoint(tr_plan* p, int bin, const tr_c64* q_sorted, const float* dcf_sorted,
                    const tr_c64* maps, int E, int C, tr_c64* g, int accumulate, void* stream) {
  if (int r = check_plan_bin(p, bin)) return r;
  hipStream_t s = (hipStream_t)stream;
  const BinData& b = p->bins[bin];
  const int K = E * C;
  if (int r = ensure_work(p, K)) return r;
  SpreadArgs a = {};
  a.G = p->G;
  a.K = K;
  a.M = b.M;
  a.row_ptr = b.row_ptr;
  a.col = b.col;
  a.val = b.val;
  a.sorted_src = 1;
  a.q = F2(q_sorted);
  a.dcf_sorted = dcf_sorted;
  a.grid = p->work;
  if (int r = launch_spread(a, s)) return r;
  return grid_to_image(p, F2(maps), E, C, F2(g), accumulate, s);
}

}  // extern "C"
