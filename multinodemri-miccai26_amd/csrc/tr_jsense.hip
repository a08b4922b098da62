// JSENSE coil-sensitivity estimation: the operators of SigPy 0.1.27's JsenseRecon on gfx950.
//
// The reference estimates coil maps with sigpy.mri.app.JsenseRecon on the first echo
// (src/toporecon/preprocessing/jsense.py:91).  Its two linear operators (ConvSense / ConvImage of
// sigpy.mri.linop) are   y_c = sqrt(w) NUFFT(IFFT(mps_ker_c (*) img_ker))   with (*) SigPy's 'valid'
// N-D convolution (signal-processing definition, kernel flipped):
//
//     out[c, n] = sum_k ker[c, k] img[n + W - 1 - k],   n in [0, I - W + 1)^3,  k in [0, W)^3.
//
// This file holds the three convolution kernels (forward, adjoint in the image kernel, adjoint in
// the map kernel), a "matrix along one axis" kernel that carries the centred DFTs on the small
// calibration grid and the k-space resizes of the output maps, and the root-sum-of-squares
// normalisation.  The NUFFT is the library's own plan (tr_data_simulate / tr_data_adjoint).
//
// The convolutions are FP32 FMA on the vector ALUs (about 8 C G^3 W^3 flops per launch, 6.4 GFLOP at
// 32 coils, a 23^3 calibration grid and 16^3 map kernels).  Work that would leave the chip idle is
// split into slabs (forward: tap planes; map-kernel adjoint: sample planes) or coil groups
// (image-kernel adjoint) whose partial sums land in a workspace and are added in a fixed order, so
// every result is bit-reproducible.  Kernel and data values that are the same for a whole
// workgroup are indexed only by block and loop indices, so they load through the scalar path.

#include "tr_common.hpp"
#include "toporecon_hip.h"

namespace tr {
namespace {

constexpr int kJsBlock = 256;
constexpr int kJsMaxWidth = 64;
constexpr int kJsMaxDim = 1024;
constexpr int kJsTargetBlocks = 2048;

struct JsConv {
  int Iz, Iy, Ix;   // image-kernel dims
  int Gz, Gy, Gx;   // output (calibration grid) dims: I - W + 1
  int W, C;
  int64_t In, Gn, W3;
};

__device__ __forceinline__ float2 cfma(float2 a, float2 b, float2 acc) {
  acc.x = fmaf(a.x, b.x, acc.x);
  acc.x = fmaf(-a.y, b.y, acc.x);
  acc.y = fmaf(a.x, b.y, acc.y);
  acc.y = fmaf(a.y, b.x, acc.y);
  return acc;
}

// acc + conj(a) b
__device__ __forceinline__ float2 cfma_conj(float2 a, float2 b, float2 acc) {
  acc.x = fmaf(a.x, b.x, acc.x);
  acc.x = fmaf(a.y, b.y, acc.x);
  acc.y = fmaf(a.x, b.y, acc.y);
  acc.y = fmaf(-a.y, b.x, acc.y);
  return acc;
}

// Forward valid convolution, taps kz in [kz0, kz1) of slab blockIdx.z, CPT coils per thread.
// part[(slab * C + c) * Gn + n].
template <int CPT>
__global__ __launch_bounds__(kJsBlock) void k_js_conv(JsConv g, const float2* __restrict__ img,
                                                     const float2* __restrict__ ker, float2* __restrict__ part,
                                                     int kz_per_slab) {
  const int64_t n = (int64_t)blockIdx.x * kJsBlock + threadIdx.x;
  if (n >= g.Gn) return;
  const int c0 = blockIdx.y * CPT;
  const int slab = blockIdx.z;
  const int kz0 = slab * kz_per_slab;
  const int kz1 = min(g.W, kz0 + kz_per_slab);
  const int nx = (int)(n % g.Gx);
  const int64_t t = n / g.Gx;
  const int ny = (int)(t % g.Gy);
  const int nz = (int)(t / g.Gy);
  const int W = g.W;
  float2 acc[CPT];
#pragma unroll
  for (int c = 0; c < CPT; ++c) acc[c] = make_float2(0.f, 0.f);
  for (int kz = kz0; kz < kz1; ++kz) {
    for (int ky = 0; ky < W; ++ky) {
      const float2* row = img + ((int64_t)(nz + W - 1 - kz) * g.Iy + (ny + W - 1 - ky)) * g.Ix + (nx + W - 1);
      const float2* kr = ker + (int64_t)c0 * g.W3 + (int64_t)(kz * W + ky) * W;
      for (int kx = 0; kx < W; ++kx) {
        const float2 v = row[-kx];
#pragma unroll
        for (int c = 0; c < CPT; ++c) acc[c] = cfma(kr[(int64_t)c * g.W3 + kx], v, acc[c]);
      }
    }
  }
#pragma unroll
  for (int c = 0; c < CPT; ++c) part[((int64_t)slab * g.C + c0 + c) * g.Gn + n] = acc[c];
}

// Rows of z padded with W - 1 zeros on both sides in x (and kJsAdjRows - 1 more at the end), so the
// image-kernel adjoint reads whole tap windows without bounds checks: zx[c, nz, ny, p] =
// z[c, nz, ny, p - (W - 1)] inside, 0 outside; rows of px values.
__global__ __launch_bounds__(kJsBlock) void k_js_pad_x(JsConv g, const float2* __restrict__ z,
                                                      float2* __restrict__ zx, int px) {
  const int64_t n = (int64_t)g.C * g.Gz * g.Gy * px;
  const int64_t stride = (int64_t)gridDim.x * kJsBlock;
  for (int64_t j = (int64_t)blockIdx.x * kJsBlock + threadIdx.x; j < n; j += stride) {
    const int p = (int)(j % px);
    const int64_t row = j / px;
    const int x = p - (g.W - 1);
    zx[j] = (x >= 0 && x < g.Gx) ? z[row * g.Gx + x] : make_float2(0.f, 0.f);
  }
}

// Adjoint in the image kernel, coils [c0, c0 + cpg) of group blockIdx.y, R consecutive outputs in x
// per thread: part[grp * In + i] = sum_c sum_k conj(ker[c, k]) z[c, i - (W - 1) + k] over the valid
// (kz, ky); every kx runs against the x-padded rows.  With the width known at compile time (W_T > 0)
// a thread loads its R + W - 1 window values of a row once and reuses them for R * W products.
// The (kz, ky) loops run over the full range with a per-lane predicate, so the tap indices are
// uniform (kernel values load through the scalar path) and a plane no lane needs is skipped.
template <int W_T, int R>
__global__ __launch_bounds__(kJsBlock) void k_js_conv_adj_img(JsConv g, const float2* __restrict__ zx, int px,
                                                             const float2* __restrict__ ker,
                                                             float2* __restrict__ part, int cpg, int nxg) {
  const int64_t t = (int64_t)blockIdx.x * kJsBlock + threadIdx.x;
  if (t >= (int64_t)g.Iz * g.Iy * nxg) return;
  const int W = W_T > 0 ? W_T : g.W;
  const int xg = (int)(t % nxg);
  const int64_t row = t / nxg;
  const int iy = (int)(row % g.Iy);
  const int iz = (int)(row / g.Iy);
  const int ix0 = xg * R;
  const int grp = blockIdx.y;
  const int c0 = grp * cpg;
  const int c1 = min(g.C, c0 + cpg);
  // valid taps: 0 <= i - (W - 1) + k < G per axis
  const int zlo = max(0, W - 1 - iz), zhi = min(W - 1, W - 1 + g.Gz - 1 - iz);
  const int ylo = max(0, W - 1 - iy), yhi = min(W - 1, W - 1 + g.Gy - 1 - iy);
  float2 acc[R];
#pragma unroll
  for (int r = 0; r < R; ++r) acc[r] = make_float2(0.f, 0.f);
  for (int c = c0; c < c1; ++c) {
    for (int kz = 0; kz < W; ++kz) {
      if (kz < zlo || kz > zhi) continue;
      const int nz = iz - (W - 1) + kz;
      for (int ky = 0; ky < W; ++ky) {
        if (ky < ylo || ky > yhi) continue;
        const int ny = iy - (W - 1) + ky;
        const float2* zr = zx + (((int64_t)c * g.Gz + nz) * g.Gy + ny) * px + ix0;
        const float2* kr = ker + (int64_t)c * g.W3 + (int64_t)(kz * W + ky) * W;
        if constexpr (W_T > 0) {
          float2 w[R + W_T - 1];
#pragma unroll
          for (int q = 0; q < R + W_T - 1; ++q) w[q] = zr[q];
#pragma unroll
          for (int kx = 0; kx < W_T; ++kx) {
            const float2 kv = kr[kx];
#pragma unroll
            for (int r = 0; r < R; ++r) acc[r] = cfma_conj(kv, w[r + kx], acc[r]);
          }
        } else {
          for (int kx = 0; kx < W; ++kx) {
            const float2 kv = kr[kx];
#pragma unroll
            for (int r = 0; r < R; ++r) acc[r] = cfma_conj(kv, zr[r + kx], acc[r]);
          }
        }
      }
    }
  }
  float2* out = part + (int64_t)grp * g.In + ((int64_t)iz * g.Iy + iy) * g.Ix;
#pragma unroll
  for (int r = 0; r < R; ++r) {
    if (ix0 + r < g.Ix) out[ix0 + r] = acc[r];
  }
}

// Adjoint in the map kernel, sample planes [nz0, nz1) of slab blockIdx.z, CPT coils per thread:
// part[(slab * C + c) * W3 + k] = sum_n z[c, n] conj(img[n + W - 1 - k]).
template <int CPT>
__global__ __launch_bounds__(kJsBlock) void k_js_conv_adj_ker(JsConv g, const float2* __restrict__ z,
                                                             const float2* __restrict__ img,
                                                             float2* __restrict__ part, int nz_per_slab) {
  const int64_t k = (int64_t)blockIdx.x * kJsBlock + threadIdx.x;
  if (k >= g.W3) return;
  const int c0 = blockIdx.y * CPT;
  const int slab = blockIdx.z;
  const int nz0 = slab * nz_per_slab;
  const int nz1 = min(g.Gz, nz0 + nz_per_slab);
  const int W = g.W;
  const int kx = (int)(k % W);
  const int ky = (int)((k / W) % W);
  const int kz = (int)(k / ((int64_t)W * W));
  float2 acc[CPT];
#pragma unroll
  for (int c = 0; c < CPT; ++c) acc[c] = make_float2(0.f, 0.f);
  for (int nz = nz0; nz < nz1; ++nz) {
    for (int ny = 0; ny < g.Gy; ++ny) {
      const float2* ir = img + ((int64_t)(nz + W - 1 - kz) * g.Iy + (ny + W - 1 - ky)) * g.Ix + (W - 1 - kx);
      const float2* zr = z + (((int64_t)c0 * g.Gz + nz) * g.Gy + ny) * g.Gx;
      for (int nx = 0; nx < g.Gx; ++nx) {
        const float2 v = ir[nx];
#pragma unroll
        for (int c = 0; c < CPT; ++c) acc[c] = cfma_conj(v, zr[(int64_t)c * g.Gn + nx], acc[c]);
      }
    }
  }
#pragma unroll
  for (int c = 0; c < CPT; ++c) part[((int64_t)slab * g.C + c0 + c) * g.W3 + k] = acc[c];
}

// out[j] = sum_{s < S} part[s * n + j], s in increasing order.
__global__ __launch_bounds__(kJsBlock) void k_js_sum_slabs(const float2* __restrict__ part, int S, int64_t n,
                                                          float2* __restrict__ out) {
  const int64_t stride = (int64_t)gridDim.x * kJsBlock;
  for (int64_t j = (int64_t)blockIdx.x * kJsBlock + threadIdx.x; j < n; j += stride) {
    float2 a = part[j];
    for (int s = 1; s < S; ++s) {
      const float2 b = part[(int64_t)s * n + j];
      a.x += b.x;
      a.y += b.y;
    }
    out[j] = a;
  }
}

// (outer, n_in, inner) -> (outer, n_out, inner): out[o, j, t] = sum_i mat[j, i] x[o, i, t].
// Each thread owns kJsRows consecutive rows j of one (o, t) column and reads x once per i.
constexpr int kJsRows = 4;

__global__ __launch_bounds__(kJsBlock) void k_js_axis(const float2* __restrict__ x, const float2* __restrict__ mat,
                                                     float2* __restrict__ out, int64_t outer, int n_in,
                                                     int n_out, int64_t inner, int64_t items) {
  const int nj = (n_out + kJsRows - 1) / kJsRows;
  const int64_t stride = (int64_t)gridDim.x * kJsBlock;
  for (int64_t q = (int64_t)blockIdx.x * kJsBlock + threadIdx.x; q < items; q += stride) {
    const int64_t tt = q % inner;
    const int64_t r = q / inner;
    const int jb = (int)(r % nj);
    const int64_t o = r / nj;
    const int j0 = jb * kJsRows;
    const float2* xp = x + o * n_in * inner + tt;
    float2 acc[kJsRows];
#pragma unroll
    for (int u = 0; u < kJsRows; ++u) acc[u] = make_float2(0.f, 0.f);
    for (int i = 0; i < n_in; ++i) {
      const float2 v = xp[(int64_t)i * inner];
#pragma unroll
      for (int u = 0; u < kJsRows; ++u) {
        const int j = min(j0 + u, n_out - 1);
        acc[u] = cfma(mat[(int64_t)j * n_in + i], v, acc[u]);
      }
    }
#pragma unroll
    for (int u = 0; u < kJsRows; ++u) {
      if (j0 + u < n_out) out[(o * n_out + j0 + u) * inner + tt] = acc[u];
    }
  }
}

// maps[c, v] /= sqrt(sum_c |maps[c, v]|^2)  (no floor, as JsenseRecon's output step).
__global__ __launch_bounds__(kJsBlock) void k_js_rss(float2* __restrict__ m, int C, int64_t N) {
  const int64_t stride = (int64_t)gridDim.x * kJsBlock;
  for (int64_t v = (int64_t)blockIdx.x * kJsBlock + threadIdx.x; v < N; v += stride) {
    float s = 0.f;
    for (int c = 0; c < C; ++c) {
      const float2 a = m[(int64_t)c * N + v];
      s = fmaf(a.x, a.x, fmaf(a.y, a.y, s));
    }
    const float r = sqrtf(s);
    for (int c = 0; c < C; ++c) {
      float2 a = m[(int64_t)c * N + v];
      a.x = a.x / r;
      a.y = a.y / r;
      m[(int64_t)c * N + v] = a;
    }
  }
}

// y += a x (the lambda I term of a regularised normal operator).
__global__ __launch_bounds__(kJsBlock) void k_js_axpy(float2* __restrict__ y, const float2* __restrict__ x,
                                                     int64_t n, float a) {
  const int64_t stride = (int64_t)gridDim.x * kJsBlock;
  for (int64_t j = (int64_t)blockIdx.x * kJsBlock + threadIdx.x; j < n; j += stride) {
    float2 v = y[j];
    const float2 u = x[j];
    v.x = fmaf(a, u.x, v.x);
    v.y = fmaf(a, u.y, v.y);
    y[j] = v;
  }
}

int make_conv(int iz, int iy, int ix, int C, int W, JsConv* g) {
  if (W < 1 || W > kJsMaxWidth) TR_FAIL(kInvalidArgument, "kernel width outside [1, 64]");
  if (C < 1) TR_FAIL(kInvalidArgument, "need at least one coil");
  if (iz < W || iy < W || ix < W || iz > kJsMaxDim || iy > kJsMaxDim || ix > kJsMaxDim)
    TR_FAIL(kInvalidArgument, "image-kernel dims must lie in [W, 1024]");
  g->Iz = iz;
  g->Iy = iy;
  g->Ix = ix;
  g->Gz = iz - W + 1;
  g->Gy = iy - W + 1;
  g->Gx = ix - W + 1;
  g->W = W;
  g->C = C;
  g->In = (int64_t)iz * iy * ix;
  g->Gn = (int64_t)g->Gz * g->Gy * g->Gx;
  g->W3 = (int64_t)W * W * W;
  return kOk;
}

int coils_per_thread(int C) { return C % 4 == 0 ? 4 : (C % 2 == 0 ? 2 : 1); }

int div_up(int64_t a, int64_t b) { return (int)((a + b - 1) / b); }

// Slabs of `extent` planes so that about kJsTargetBlocks workgroups run.
int slab_count(int64_t base_blocks, int extent) {
  int s = div_up(kJsTargetBlocks, base_blocks);
  return s < 1 ? 1 : (s > extent ? extent : s);
}

struct ConvLaunch {
  int cpt, per_slab, slabs;
};

ConvLaunch fwd_launch(const JsConv& g) {
  ConvLaunch l;
  l.cpt = coils_per_thread(g.C);
  const int64_t base = (int64_t)div_up(g.Gn, kJsBlock) * (g.C / l.cpt);
  const int s = slab_count(base, g.W);
  l.per_slab = div_up(g.W, s);
  l.slabs = div_up(g.W, l.per_slab);
  return l;
}

ConvLaunch adj_ker_launch(const JsConv& g) {
  ConvLaunch l;
  l.cpt = coils_per_thread(g.C);
  const int64_t base = (int64_t)div_up(g.W3, kJsBlock) * (g.C / l.cpt);
  const int s = slab_count(base, g.Gz);
  l.per_slab = div_up(g.Gz, s);
  l.slabs = div_up(g.Gz, l.per_slab);
  return l;
}

// Image-kernel adjoint launch: outputs per thread in x, the x-padded row length, threads per coil
// group, coils per group.
constexpr int kJsAdjRows = 4;
constexpr int kJsAdjWidth = 16;  // the map-kernel width with a compile-time window (SigPy's default)

struct AdjImgLaunch {
  int R, px, nxg, cpg, groups;
  int64_t threads;
};

AdjImgLaunch adj_img_launch(const JsConv& g) {
  AdjImgLaunch l;
  l.R = g.W == kJsAdjWidth ? kJsAdjRows : 1;
  l.px = g.Gx + 2 * (g.W - 1) + l.R - 1;
  l.nxg = div_up(g.Ix, l.R);
  l.threads = (int64_t)g.Iz * g.Iy * l.nxg;
  const int64_t base = div_up(l.threads, kJsBlock);
  int groups = div_up(kJsTargetBlocks, base);
  if (groups < 1) groups = 1;
  if (groups > g.C) groups = g.C;
  l.cpg = div_up(g.C, groups);
  l.groups = div_up(g.C, l.cpg);
  return l;
}

int64_t work_count(const JsConv& g) {
  const ConvLaunch f = fwd_launch(g), a = adj_ker_launch(g);
  const AdjImgLaunch m = adj_img_launch(g);
  int64_t n = (int64_t)f.slabs * g.C * g.Gn;
  n = n > (int64_t)a.slabs * g.C * g.W3 ? n : (int64_t)a.slabs * g.C * g.W3;
  const int64_t ni = m.groups * g.In + (int64_t)g.C * g.Gz * g.Gy * m.px;
  return n > ni ? n : ni;
}

unsigned ew_blocks(int64_t n) {
  const int64_t b = (n + kJsBlock - 1) / kJsBlock;
  return (unsigned)(b < 1 ? 1 : (b > 65536 ? 65536 : b));
}

int sum_slabs(const float2* part, int S, int64_t n, float2* out, hipStream_t s) {
  hipLaunchKernelGGL(k_js_sum_slabs, dim3(ew_blocks(n)), dim3(kJsBlock), 0, s, part, S, n, out);
  TR_LAUNCH_CHECK();
  return kOk;
}

}  // namespace
}  // namespace tr

using namespace tr;

extern "C" {

int64_t tr_js_work_count(int iz, int iy, int ix, int C, int W) {
  JsConv g;
  if (make_conv(iz, iy, ix, C, W, &g)) return -1;
  return work_count(g);
}

int tr_js_conv(const tr_c64* img, int iz, int iy, int ix, const tr_c64* ker, int C, int W, tr_c64* out,
               tr_c64* work, void* stream) {
  JsConv g;
  if (int r = make_conv(iz, iy, ix, C, W, &g)) return r;
  if (!img || !ker || !out || !work) TR_FAIL(kInvalidArgument, "null argument");
  hipStream_t s = (hipStream_t)stream;
  const ConvLaunch l = fwd_launch(g);
  ProfScope ps(kKJsConv, 8.0 * (g.In + g.C * g.W3 + g.C * g.Gn), s);
  const dim3 grid(div_up(g.Gn, kJsBlock), g.C / l.cpt, l.slabs);
  float2* part = (float2*)work;
  if (l.cpt == 4)
    hipLaunchKernelGGL(k_js_conv<4>, grid, dim3(kJsBlock), 0, s, g, (const float2*)img, (const float2*)ker, part,
                       l.per_slab);
  else if (l.cpt == 2)
    hipLaunchKernelGGL(k_js_conv<2>, grid, dim3(kJsBlock), 0, s, g, (const float2*)img, (const float2*)ker, part,
                       l.per_slab);
  else
    hipLaunchKernelGGL(k_js_conv<1>, grid, dim3(kJsBlock), 0, s, g, (const float2*)img, (const float2*)ker, part,
                       l.per_slab);
  TR_LAUNCH_CHECK();
  return sum_slabs(part, l.slabs, (int64_t)g.C * g.Gn, (float2*)out, s);
}

int tr_js_conv_adj_img(const tr_c64* z, int iz, int iy, int ix, const tr_c64* ker, int C, int W, tr_c64* out,
                       tr_c64* work, void* stream) {
  JsConv g;
  if (int r = make_conv(iz, iy, ix, C, W, &g)) return r;
  if (!z || !ker || !out || !work) TR_FAIL(kInvalidArgument, "null argument");
  hipStream_t s = (hipStream_t)stream;
  const AdjImgLaunch l = adj_img_launch(g);
  float2* part = (float2*)work;
  float2* zx = part + (int64_t)l.groups * g.In;
  ProfScope ps(kKJsConvAdjImg, 8.0 * (g.C * g.Gn + g.C * g.W3 + g.In), s);
  hipLaunchKernelGGL(k_js_pad_x, dim3(ew_blocks((int64_t)g.C * g.Gz * g.Gy * l.px)), dim3(kJsBlock), 0, s, g,
                     (const float2*)z, zx, l.px);
  TR_LAUNCH_CHECK();
  const dim3 grid(div_up(l.threads, kJsBlock), l.groups, 1);
  if (l.R == kJsAdjRows)
    hipLaunchKernelGGL((k_js_conv_adj_img<kJsAdjWidth, kJsAdjRows>), grid, dim3(kJsBlock), 0, s, g,
                       (const float2*)zx, l.px, (const float2*)ker, part, l.cpg, l.nxg);
  else
    hipLaunchKernelGGL((k_js_conv_adj_img<0, 1>), grid, dim3(kJsBlock), 0, s, g, (const float2*)zx, l.px,
                       (const float2*)ker, part, l.cpg, l.nxg);
  TR_LAUNCH_CHECK();
  return sum_slabs(part, l.groups, g.In, (float2*)out, s);
}

int tr_js_conv_adj_ker(const tr_c64* z, const tr_c64* img, int iz, int iy, int ix, int C, int W, tr_c64* out,
                       tr_c64* work, void* stream) {
  JsConv g;
  if (int r = make_conv(iz, iy, ix, C, W, &g)) return r;
  if (!z || !img || !out || !work) TR_FAIL(kInvalidArgument, "null argument");
  hipStream_t s = (hipStream_t)stream;
  const ConvLaunch l = adj_ker_launch(g);
  ProfScope ps(kKJsConvAdjKer, 8.0 * (g.C * g.Gn + g.In + g.C * g.W3), s);
  const dim3 grid(div_up(g.W3, kJsBlock), g.C / l.cpt, l.slabs);
  float2* part = (float2*)work;
  if (l.cpt == 4)
    hipLaunchKernelGGL(k_js_conv_adj_ker<4>, grid, dim3(kJsBlock), 0, s, g, (const float2*)z, (const float2*)img,
                       part, l.per_slab);
  else if (l.cpt == 2)
    hipLaunchKernelGGL(k_js_conv_adj_ker<2>, grid, dim3(kJsBlock), 0, s, g, (const float2*)z, (const float2*)img,
                       part, l.per_slab);
  else
    hipLaunchKernelGGL(k_js_conv_adj_ker<1>, grid, dim3(kJsBlock), 0, s, g, (const float2*)z, (const float2*)img,
                       part, l.per_slab);
  TR_LAUNCH_CHECK();
  return sum_slabs(part, l.slabs, (int64_t)g.C * g.W3, (float2*)out, s);
}

int tr_js_axis_apply(const tr_c64* x, int64_t outer, int n_in, int64_t inner, const tr_c64* mat, int n_out,
                     tr_c64* out, void* stream) {
  if (outer < 1 || inner < 1 || n_in < 1 || n_out < 1) TR_FAIL(kInvalidArgument, "empty shape");
  if (!x || !mat || !out) TR_FAIL(kInvalidArgument, "null argument");
  if (x == out) TR_FAIL(kInvalidArgument, "out must not alias x");
  hipStream_t s = (hipStream_t)stream;
  const int64_t items = outer * ((n_out + kJsRows - 1) / kJsRows) * inner;
  ProfScope ps(kKJsAxis, 8.0 * (outer * inner * (n_in + (int64_t)n_out) + (int64_t)n_in * n_out), s);
  hipLaunchKernelGGL(k_js_axis, dim3(ew_blocks(items)), dim3(kJsBlock), 0, s, (const float2*)x,
                     (const float2*)mat, (float2*)out, outer, n_in, n_out, inner, items);
  TR_LAUNCH_CHECK();
  return kOk;
}

int tr_js_rss_normalize(tr_c64* maps, int C, int64_t N, void* stream) {
  if (C < 1 || N < 0) TR_FAIL(kInvalidArgument, "bad shape");
  if (N == 0) return kOk;
  if (!maps) TR_FAIL(kInvalidArgument, "null argument");
  hipStream_t s = (hipStream_t)stream;
  ProfScope ps(kKJsRss, 16.0 * C * N, s);
  hipLaunchKernelGGL(k_js_rss, dim3(ew_blocks(N)), dim3(kJsBlock), 0, s, (float2*)maps, C, N);
  TR_LAUNCH_CHECK();
  return kOk;
}

int tr_js_axpy(tr_c64* y, const tr_c64* x, int64_t n, float a, void* stream) {
  if (n < 0) TR_FAIL(kInvalidArgument, "negative length");
  if (n == 0) return kOk;
  if (!y || !x) TR_FAIL(kInvalidArgument, "null argument");
  hipLaunchKernelGGL(k_js_axpy, dim3(ew_blocks(n)), dim3(kJsBlock), 0, (hipStream_t)stream, (float2*)y,
                     (const float2*)x, n, a);
  TR_LAUNCH_CHECK();
  return kOk;
}

}  // extern "C"
