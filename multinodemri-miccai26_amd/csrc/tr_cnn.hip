// Warm-start network kernels: 3x3x3 'same' convolutions over channels-last
// bf16 volumes on MFMA (v_mfma_f32_32x32x8_bf16, fp32 accumulation), the
// weight gradient as a split-K MFMA contraction with a fixed-order reduction,
// and the element-wise pieces of training (loss, Adam, packing).
//
// Layouts.  Activations: [n][Z][Y][X][C] bf16, C = 8 (network input: re, im,
// 6 zeros) or 32.  Master weights: fp32 [27][Cin][32], tap t = 9(dz+1) +
// 3(dy+1) + (dx+1).  MFMA fragments: wfrag[((t * Cin/8 + cb) * 64 + lane) * 4
// + j] = W[t][cb*8 + 4*(lane/32) + j][lane%32], the B operand of one
// 32x32x8 step (A = 32 x-voxels x 8 input channels, D = 32 voxels x 32
// output channels; lane l of D item v holds voxel 8(v/4) + 4(l/32) + v%4,
// channel l%32).

#include <cmath>

#include "tr_internal.hpp"

namespace tr {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

constexpr int kCnnTX = 32;  // x voxels of one output row (MFMA M)
constexpr int kCnnTY = 4;   // output rows per workgroup, one per wave
constexpr int kCnnCo = 32;  // output channels (MFMA N)

__device__ __forceinline__ uint16_t f2bf(float f) {
  uint32_t u = __float_as_uint(f);
  if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)((u >> 16) | ((u & 0xffffu) ? 0x40u : 0u));
  u += 0x7fffu + ((u >> 16) & 1u);  // round to nearest even
  return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

struct ConvArgs {
  const uint16_t* x;  // [n][Z][Y][X][CIN]
  const bf16x4* wfrag;
  const float* bias;  // [32] (modes 0, 1)
  const uint16_t* mask;  // [n][Z][Y][X][32] (mode 2)
  void* out;
  int n, Z, Y, X;
  int mode;       // 0: +bias, ReLU -> bf16 [..][32]; 1: +bias -> f32 [..][cout_keep]; 2: x (mask > 0) -> bf16
  int cout_keep;  // mode 1
};

struct ConvTile {
  int img, z, y0, x0;
};

__device__ __forceinline__ ConvTile tile_of(int64_t b, int Z, int Y, int X) {
  const int nxt = (X + kCnnTX - 1) / kCnnTX, nyt = (Y + kCnnTY - 1) / kCnnTY;
  ConvTile t;
  const int xt = (int)(b % nxt);
  b /= nxt;
  const int yt = (int)(b % nyt);
  b /= nyt;
  t.z = (int)(b % Z);
  t.img = (int)(b / Z);
  t.x0 = xt * kCnnTX;
  t.y0 = yt * kCnnTY;
  return t;
}

// Stage the (3, kCnnTY + 2, kCnnTX + 2) voxel neighbourhood of a tile into LDS
// with voxel stride VS bf16 (zeros outside the volume).
template <int CIN, int VS>
__device__ __forceinline__ void stage_patch(uint16_t* patch, const uint16_t* x, const ConvTile& t, int Z, int Y,
                                            int X) {
  constexpr int NCB = CIN / 8, PX = kCnnTX + 2, PY = kCnnTY + 2, PZ = 3;
  const uint16_t* src = x + (int64_t)t.img * Z * Y * X * CIN;
  // every load of the thread's share is issued before any is written to LDS (a 256-thread block)
  constexpr int NTOT = PZ * PY * PX * NCB, NPER = (NTOT + 255) / 256;
  uint4 val[NPER];
#pragma unroll
  for (int i = 0; i < NPER; ++i) {
    const int c = (int)threadIdx.x + i * 256;
    const int cb = c % NCB;
    int v = c / NCB;
    const int px = v % PX;
    v /= PX;
    const int py = v % PY;
    const int pz = v / PY;
    const int gz = t.z + pz - 1, gy = t.y0 + py - 1, gx = t.x0 + px - 1;
    val[i] = make_uint4(0u, 0u, 0u, 0u);
    if (c < NTOT && gz >= 0 && gz < Z && gy >= 0 && gy < Y && gx >= 0 && gx < X)
      val[i] = *(const uint4*)(src + (((int64_t)gz * Y + gy) * X + gx) * CIN + cb * 8);
  }
#pragma unroll
  for (int i = 0; i < NPER; ++i) {
    const int c = (int)threadIdx.x + i * 256;
    if (c < NTOT) {
      const int cb = c % NCB;
      int v = c / NCB;
      const int px = v % PX;
      v /= PX;
      const int py = v % PY;
      const int pz = v / PY;
      uint2* dst = (uint2*)(patch + ((pz * PY + py) * PX + px) * VS + cb * 8);
      dst[0] = make_uint2(val[i].x, val[i].y);
      dst[1] = make_uint2(val[i].z, val[i].w);
    }
  }
}

template <int CIN>
__global__ void __launch_bounds__(256) k_conv3(ConvArgs a) {
  constexpr int NCB = CIN / 8;
  constexpr int VS = CIN + 4;  // padded voxel stride: 8-B fragment reads hit distinct banks
  constexpr int PX = kCnnTX + 2, PY = kCnnTY + 2;
  __shared__ __align__(16) uint16_t patch[3 * PY * PX * VS];
  const ConvTile t = tile_of(blockIdx.x, a.Z, a.Y, a.X);
  stage_patch<CIN, VS>(patch, a.x, t, a.Z, a.Y, a.X);
  __syncthreads();
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int xi = lane & 31, kh = lane >> 5;
  f32x16 acc;
#pragma unroll
  for (int v = 0; v < 16; ++v) acc[v] = 0.f;
  for (int tap = 0; tap < 27; ++tap) {
    const int dz = tap / 9, dy = (tap / 3) % 3, dx = tap % 3;
    const uint16_t* pa = patch + ((dz * PY + wave + dy) * PX + xi + dx) * VS + kh * 4;
#pragma unroll
    for (int cb = 0; cb < NCB; ++cb) {
      const bf16x4 av = *(const bf16x4*)(pa + cb * 8);
      const bf16x4 bv = a.wfrag[(tap * NCB + cb) * 64 + lane];
      acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(av, bv, acc, 0, 0, 0);
    }
  }
  const int y = t.y0 + wave;
  if (y >= a.Y) return;
  const int co = lane & 31;
  const int64_t row = (((int64_t)t.img * a.Z + t.z) * a.Y + y) * a.X;
  const float bias = (a.mode == 2) ? 0.f : a.bias[co];
#pragma unroll
  for (int v = 0; v < 16; ++v) {
    const int gx = t.x0 + 8 * (v / 4) + 4 * kh + (v % 4);
    if (gx >= a.X) continue;
    const int64_t vox = row + gx;
    const float r = acc[v] + bias;
    if (a.mode == 0) {
      ((uint16_t*)a.out)[vox * kCnnCo + co] = f2bf(fmaxf(r, 0.f));
    } else if (a.mode == 1) {
      if (co < a.cout_keep) ((float*)a.out)[vox * a.cout_keep + co] = r;
    } else {
      const bool on = bf2f(a.mask[vox * kCnnCo + co]) > 0.f;
      ((uint16_t*)a.out)[vox * kCnnCo + co] = on ? f2bf(r) : (uint16_t)0;
    }
  }
}

struct WgradArgs {
  const uint16_t* x;   // layer input [n][Z][Y][X][CIN]
  const uint16_t* dy;  // output gradient [n][Z][Y][X][32] (ReLU mask applied)
  float* partial;      // [nblocks][27][32 co][32 ci]
  float* partial_db;   // [nblocks][32]
  int n, Z, Y, X;
  int64_t ntiles;
};

// dW[t][ci][co] = sum_v x[v + off(t)][ci] dy[v][co]: per tap a 32 (co) x 32 (ci)
// MFMA contraction over the tile's voxels (A = dy^T, B = shifted x).  Block
// (b, dz) stages the one input plane z + dz - 1 of tiles b, b + gridDim.x, ...
// and its 4 waves share the 9 (dy, dx) taps of that plane (at most 3 each);
// every block writes its own partial sums and k_cnn_wgrad_reduce adds the
// blocks in index order.
template <int CIN>
__global__ void __launch_bounds__(256) k_conv3_wgrad(WgradArgs a) {
  constexpr int VS = CIN + 4, PX = kCnnTX + 2, PY = kCnnTY + 2;
  constexpr int DS = kCnnCo + 4;
  constexpr int NT = 3;  // taps per wave: w, w + 4, w + 8 (< 9)
  __shared__ __align__(16) uint16_t plane[PY * PX * VS];
  __shared__ __align__(16) uint16_t dyt[kCnnTY * kCnnTX * DS];
  const int dzi = blockIdx.y;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int l32 = lane & 31, kh = lane >> 5;
  f32x16 acc[NT];
#pragma unroll
  for (int i = 0; i < NT; ++i)
#pragma unroll
    for (int v = 0; v < 16; ++v) acc[i][v] = 0.f;
  float dbsum = 0.f;
  for (int64_t tb = blockIdx.x; tb < a.ntiles; tb += gridDim.x) {
    const ConvTile t = tile_of(tb, a.Z, a.Y, a.X);
    __syncthreads();
    {  // input plane z + dz - 1, rows y0 - 1 .. y0 + kCnnTY, columns x0 - 1 .. x0 + kCnnTX
      constexpr int NCB = CIN / 8;
      const uint16_t* src = a.x + (int64_t)t.img * a.Z * a.Y * a.X * CIN;
      const int gz = t.z + dzi - 1;
      for (int c = threadIdx.x; c < PY * PX * NCB; c += blockDim.x) {
        const int cb = c % NCB, v = c / NCB;
        const int px = v % PX, py = v / PX;
        const int gy = t.y0 + py - 1, gx = t.x0 + px - 1;
        uint4 val = make_uint4(0u, 0u, 0u, 0u);
        if (gz >= 0 && gz < a.Z && gy >= 0 && gy < a.Y && gx >= 0 && gx < a.X)
          val = *(const uint4*)(src + (((int64_t)gz * a.Y + gy) * a.X + gx) * CIN + cb * 8);
        uint2* dst = (uint2*)(plane + (py * PX + px) * VS + cb * 8);
        dst[0] = make_uint2(val.x, val.y);
        dst[1] = make_uint2(val.z, val.w);
      }
    }
    for (int c = threadIdx.x; c < kCnnTY * kCnnTX * 4; c += blockDim.x) {
      const int cb = c & 3, v = c >> 2;
      const int gy = t.y0 + v / kCnnTX, gx = t.x0 + v % kCnnTX;
      uint4 val = make_uint4(0u, 0u, 0u, 0u);
      if (gy < a.Y && gx < a.X)
        val = *(const uint4*)(a.dy + ((((int64_t)t.img * a.Z + t.z) * a.Y + gy) * a.X + gx) * kCnnCo + cb * 8);
      uint2* dst = (uint2*)(dyt + v * DS + cb * 8);
      dst[0] = make_uint2(val.x, val.y);
      dst[1] = make_uint2(val.z, val.w);
    }
    __syncthreads();
    if (dzi == 0 && threadIdx.x < kCnnCo)
      for (int v = 0; v < kCnnTY * kCnnTX; ++v) dbsum += bf2f(dyt[v * DS + threadIdx.x]);
#pragma unroll
    for (int i = 0; i < NT; ++i) {
      const int t9 = wave + 4 * i;
      if (t9 < 9) {
        const int dy = t9 / 3, dx = t9 % 3;
#pragma unroll 1
        for (int r = 0; r < kCnnTY; ++r) {
#pragma unroll
          for (int ks = 0; ks < kCnnTX / 8; ++ks) {
            bf16x4 av, bv;
#pragma unroll
            for (int j = 0; j < 4; ++j) {
              const int xv = ks * 8 + 4 * kh + j;
              av[j] = (short)dyt[(r * kCnnTX + xv) * DS + l32];
              bv[j] = l32 < CIN ? (short)plane[((r + dy) * PX + xv + dx) * VS + l32] : (short)0;
            }
            acc[i] = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(av, bv, acc[i], 0, 0, 0);
          }
        }
      }
    }
  }
  float* dst = a.partial + (int64_t)blockIdx.x * 27 * kCnnCo * kCnnCo;
#pragma unroll
  for (int i = 0; i < NT; ++i) {
    const int t9 = wave + 4 * i;
    if (t9 < 9) {
      const int tap = dzi * 9 + t9;
#pragma unroll
      for (int v = 0; v < 16; ++v) {
        const int co = 8 * (v / 4) + 4 * kh + (v % 4);
        dst[((int64_t)tap * kCnnCo + co) * kCnnCo + l32] = acc[i][v];
      }
    }
  }
  if (dzi == 0 && threadIdx.x < kCnnCo) a.partial_db[(int64_t)blockIdx.x * kCnnCo + threadIdx.x] = dbsum;
}

__global__ void k_cnn_wgrad_reduce(const float* partial, const float* partial_db, int nblocks, int cin,
                                   float* dw, float* db) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int nw = 27 * cin * kCnnCo;
  if (i < nw) {
    const int co = i % kCnnCo, ci = (i / kCnnCo) % cin, tap = i / (kCnnCo * cin);
    float s = 0.f;
    for (int b = 0; b < nblocks; ++b) s += partial[(((int64_t)b * 27 + tap) * kCnnCo + co) * kCnnCo + ci];
    dw[i] = s;  // dw[tap][ci][co]
  } else if (i < nw + kCnnCo) {
    const int co = i - nw;
    float s = 0.f;
    for (int b = 0; b < nblocks; ++b) s += partial_db[(int64_t)b * kCnnCo + co];
    db[co] = s;
  }
}

// Fragments of W (fp32 [27][cin][32]) for the forward pass, or (flip = 1) of
// W'[t][co][ci] = W[26 - t][ci][co] for the data gradient (a forward
// convolution of dy with the flipped, transposed kernel; cin must be 32).
__global__ void k_cnn_pack(const float* w, int cin, int flip, uint16_t* frag) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int ncb = (flip ? kCnnCo : cin) / 8;
  if (i >= 27 * ncb * 64 * 4) return;
  const int j = i & 3, lane = (i >> 2) & 63, cb = (i >> 8) % ncb, tap = (i >> 8) / ncb;
  const int k = cb * 8 + 4 * (lane >> 5) + j, n = lane & 31;
  const float v = flip ? w[((26 - tap) * cin + n) * kCnnCo + k] : w[(tap * cin + k) * kCnnCo + n];
  frag[i] = f2bf(v);
}

// dy[v][c] = scale * (out[v][c] - target[v][c]) for c < 2, 0 for c >= 2 (bf16),
// and per-block float64 sums of (out - target)^2 as (sum, 0) pairs.
__global__ void k_cnn_mse(const float* out, const float* target, int64_t V, float scale, uint16_t* dy,
                          double2* partial) {
  __shared__ double red[256];
  double s = 0.0;
  for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < V; v += (int64_t)gridDim.x * blockDim.x) {
    const float d0 = out[2 * v] - target[2 * v], d1 = out[2 * v + 1] - target[2 * v + 1];
    s += (double)d0 * d0 + (double)d1 * d1;
    uint4* row = (uint4*)(dy + v * kCnnCo);
    const uint32_t lo = (uint32_t)f2bf(scale * d0) | ((uint32_t)f2bf(scale * d1) << 16);
    row[0] = make_uint4(lo, 0u, 0u, 0u);
    row[1] = make_uint4(0u, 0u, 0u, 0u);
    row[2] = make_uint4(0u, 0u, 0u, 0u);
    row[3] = make_uint4(0u, 0u, 0u, 0u);
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = make_double2(red[0], 0.0);
}

__global__ void k_cnn_adam(float* p, const float* g, float* m, float* v, int64_t n, float lr, float b1,
                           float b2, float eps, float bc1, float bc2) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  p[i] -= lr * (mi / bc1) / (sqrtf(vi / bc2) + eps);
}

// x[i][v] = (re, im, 0 x 6) of img[i][v] / scale[i], bf16.
__global__ void k_cnn_pack_input(const float2* img, const float* scale, int64_t V, int n, uint16_t* x) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= V * n) return;
  const float s = scale[i / V];
  const float2 c = img[i];
  const float inv = s > 0.f ? 1.f / s : 0.f;
  const uint32_t lo = (uint32_t)f2bf(c.x * inv) | ((uint32_t)f2bf(c.y * inv) << 16);
  *(uint4*)(x + i * 8) = make_uint4(lo, 0u, 0u, 0u);
}

// img[i][v] = scale[i] * (x[i][v][0:2] + out[i][v][0:2]) as complex64.
__global__ void k_cnn_unpack(const uint16_t* x, const float* out, const float* scale, int64_t V, int n,
                             float2* img) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= V * n) return;
  const float s = scale[i / V];
  img[i] = make_float2(s * (bf2f(x[i * 8]) + out[2 * i]), s * (bf2f(x[i * 8 + 1]) + out[2 * i + 1]));
}

// Per-image float64 sum of |img|^2 (one block per image, fixed-order tree).
__global__ void k_cnn_sumsq(const float2* img, int64_t V, double* out) {
  __shared__ double red[256];
  const float2* p = img + (int64_t)blockIdx.x * V;
  double s = 0.0;
  for (int64_t v = threadIdx.x; v < V; v += blockDim.x) s += (double)p[v].x * p[v].x + (double)p[v].y * p[v].y;
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[blockIdx.x] = red[0];
}

// Data-consistency sums per channel over a bin's sorted samples:
// out[k] = (sum_m d_m Re(conj(a_mk) y_mk), sum_m d_m |a_mk|^2).  Block c sums
// samples c, c + gridDim.x, ... for every channel; k_fit_reduce adds blocks in order.
__global__ void k_fit_partial(const float2* a, const float2* y, const float* dcf, int64_t M, int K,
                              double2* partial) {
  for (int k = threadIdx.x; k < K; k += blockDim.x) {
    double s1 = 0.0, s2 = 0.0;
    for (int64_t m = blockIdx.x; m < M; m += gridDim.x) {
      const float2 av = a[m * K + k], yv = y[m * K + k];
      const double d = dcf ? (double)dcf[m] : 1.0;
      s1 += d * ((double)av.x * yv.x + (double)av.y * yv.y);
      s2 += d * ((double)av.x * av.x + (double)av.y * av.y);
    }
    partial[(int64_t)blockIdx.x * K + k] = make_double2(s1, s2);
  }
}

__global__ void k_fit_reduce(const double2* partial, int nblocks, int K, double2* out) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= K) return;
  double s1 = 0.0, s2 = 0.0;
  for (int b = 0; b < nblocks; ++b) {
    s1 += partial[(int64_t)b * K + k].x;
    s2 += partial[(int64_t)b * K + k].y;
  }
  out[k] = make_double2(s1, s2);
}

// q[m, k] *= scale[k] over sorted samples (M, K): per-channel scaling of k-space.
__global__ void k_scale_channels(float2* __restrict__ q, int64_t n, int K, const float* __restrict__ scale) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const float c = scale[i % K];
    float2 v = q[i];
    v.x *= c;
    v.y *= c;
    q[i] = v;
  }
}

}  // namespace tr

using namespace tr;

extern "C" {

int tr_cnn_conv(const uint16_t* x, int n, int Z, int Y, int X, int cin, const uint16_t* wfrag,
                const float* bias, int mode, const uint16_t* mask, void* out, int cout_keep, void* stream) {
  if (!x || !wfrag || !out) TR_FAIL(kInvalidArgument, "null argument");
  if (n < 1 || Z < 1 || Y < 1 || X < 1) TR_FAIL(kInvalidArgument, "volume dimensions must be positive");
  if (cin != 8 && cin != 32) TR_FAIL(kUnsupported, "input channels must be 8 or 32");
  if (mode < 0 || mode > 2) TR_FAIL(kInvalidArgument, "mode must be 0, 1 or 2");
  if (mode != 2 && !bias) TR_FAIL(kInvalidArgument, "bias required");
  if (mode == 2 && !mask) TR_FAIL(kInvalidArgument, "mask required");
  if (mode == 1 && (cout_keep < 1 || cout_keep > kCnnCo)) TR_FAIL(kInvalidArgument, "cout_keep outside [1, 32]");
  ConvArgs a = {x, (const bf16x4*)wfrag, bias, mask, out, n, Z, Y, X, mode, cout_keep};
  const int64_t tiles = (int64_t)n * Z * ((Y + kCnnTY - 1) / kCnnTY) * ((X + kCnnTX - 1) / kCnnTX);
  if (tiles > 0x7fffffffLL) TR_FAIL(kUnsupported, "volume too large");
  hipStream_t s = (hipStream_t)stream;
  if (cin == 8) hipLaunchKernelGGL(k_conv3<8>, dim3((unsigned)tiles), dim3(256), 0, s, a);
  else hipLaunchKernelGGL(k_conv3<32>, dim3((unsigned)tiles), dim3(256), 0, s, a);
  TR_LAUNCH_CHECK();
  return kOk;
}

int64_t tr_cnn_wgrad_work_size(int nblocks) { return (int64_t)nblocks * (27 * kCnnCo * kCnnCo + kCnnCo); }

int tr_cnn_wgrad(const uint16_t* x, const uint16_t* dy, int n, int Z, int Y, int X, int cin, int nblocks,
                 float* work, float* dw, float* db, void* stream) {
  if (!x || !dy || !work || !dw || !db) TR_FAIL(kInvalidArgument, "null argument");
  if (n < 1 || Z < 1 || Y < 1 || X < 1) TR_FAIL(kInvalidArgument, "volume dimensions must be positive");
  if (cin != 8 && cin != 32) TR_FAIL(kUnsupported, "input channels must be 8 or 32");
  if (nblocks < 1 || nblocks > 65536) TR_FAIL(kInvalidArgument, "nblocks outside [1, 65536]");
  WgradArgs a;
  a.x = x;
  a.dy = dy;
  a.partial = work;
  a.partial_db = work + (int64_t)nblocks * 27 * kCnnCo * kCnnCo;
  a.n = n;
  a.Z = Z;
  a.Y = Y;
  a.X = X;
  a.ntiles = (int64_t)n * Z * ((Y + kCnnTY - 1) / kCnnTY) * ((X + kCnnTX - 1) / kCnnTX);
  hipStream_t s = (hipStream_t)stream;
  if (cin == 8) hipLaunchKernelGGL(k_conv3_wgrad<8>, dim3(nblocks, 3), dim3(256), 0, s, a);
  else hipLaunchKernelGGL(k_conv3_wgrad<32>, dim3(nblocks, 3), dim3(256), 0, s, a);
  TR_LAUNCH_CHECK();
  const int nout = 27 * cin * kCnnCo + kCnnCo;
  hipLaunchKernelGGL(k_cnn_wgrad_reduce, dim3(grid1d(nout, 256)), dim3(256), 0, s, a.partial, a.partial_db,
                     nblocks, cin, dw, db);
  TR_LAUNCH_CHECK();
  return kOk;
}

int tr_cnn_pack_weights(const float* w, int cin, int flip, uint16_t* frag, void* stream) {
  if (!w || !frag) TR_FAIL(kInvalidArgument, "null argument");
  if (cin != 8 && cin != 32) TR_FAIL(kUnsupported, "input channels must be 8 or 32");
  if (flip && cin != 32) TR_FAIL(kUnsupported, "the data-gradient kernel needs 32 input channels");
  const int n = 27 * ((flip ? kCnnCo : cin) / 8) * 64 * 4;
  hipLaunchKernelGGL(k_cnn_pack, dim3(grid1d(n, 256)), dim3(256), 0, (hipStream_t)stream, w, cin, flip, frag);
  TR_LAUNCH_CHECK();
  return kOk;
}

int tr_cnn_mse(const float* out, const float* target, int64_t V, float scale, uint16_t* dy, double* partial,
               int nblocks, void* stream) {
  if (!out || !target || !dy || !partial) TR_FAIL(kInvalidArgument, "null argument");
  if (nblocks < 1 || nblocks > 65536) TR_FAIL(kInvalidArgument, "nblocks outside [1, 65536]");
  hipLaunchKernelGGL(k_cnn_mse, dim3(nblocks), dim3(256), 0, (hipStream_t)stream, out, target, V, scale, dy,
                     (double2*)partial);
  TR_LAUNCH_CHECK();
  return kOk;
}

int tr_cnn_adam(float* p, const float* g, float* m, float* v, int64_t n, float lr, float b1, float b2,
                float eps, int step, void* stream) {
  if (!p || !g || !m || !v) TR_FAIL(kInvalidArgument, "null argument");
  if (step < 1) TR_FAIL(kInvalidArgument, "step counts from 1");
  const float bc1 = (float)(1.0 - std::pow((double)b1, step));
  const float bc2 = (float)(1.0 - std::pow((double)b2, step));
  if (n == 0) return kOk;
  hipLaunchKernelGGL(k_cnn_adam, dim3(grid1d(n, 256)), dim3(256), 0, (hipStream_t)stream, p, g, m, v, n, lr, b1,
                     b2, eps, bc1, bc2);
  TR_LAUNCH_CHECK();
  return kOk;
}

int tr_cnn_pack_input(const tr_c64* img, const float* scale, int64_t V, int n, uint16_t* x, void* stream) {
  if (!img || !scale || !x) TR_FAIL(kInvalidArgument, "null argument");
  if (V * n == 0) return kOk;
  hipLaunchKernelGGL(k_cnn_pack_input, dim3(grid1d(V * n, 256)), dim3(256), 0, (hipStream_t)stream, F2(img),
                     scale, V, n, x);
  TR_LAUNCH_CHECK();
  return kOk;
}

int tr_cnn_unpack(const uint16_t* x, const float* out, const float* scale, int64_t V, int n, tr_c64* img,
                  void* stream) {
  if (!x || !out || !scale || !img) TR_FAIL(kInvalidArgument, "null argument");
  if (V * n == 0) return kOk;
  hipLaunchKernelGGL(k_cnn_unpack, dim3(grid1d(V * n, 256)), dim3(256), 0, (hipStream_t)stream, x, out, scale, V,
                     n, F2(img));
  TR_LAUNCH_CHECK();
  return kOk;
}

int tr_cnn_sumsq(const tr_c64* img, int64_t V, int n, double* out, void* stream) {
  if (!img || !out) TR_FAIL(kInvalidArgument, "null argument");
  if (n < 1) return kOk;
  hipLaunchKernelGGL(k_cnn_sumsq, dim3(n), dim3(256), 0, (hipStream_t)stream, F2(img), V, out);
  TR_LAUNCH_CHECK();
  return kOk;
}

int tr_data_scale_channels(tr_c64* q_sorted, int64_t M, int K, const float* scale, void* stream) {
  if (M < 0 || K < 1) TR_FAIL(kInvalidArgument, "need M >= 0 and K >= 1");
  if (M == 0) return kOk;
  if (!q_sorted || !scale) TR_FAIL(kInvalidArgument, "null argument");
  const int64_t n = M * K;
  const int64_t b = (n + 255) / 256;
  hipLaunchKernelGGL(k_scale_channels, dim3((unsigned)(b > 65536 ? 65536 : b)), dim3(256), 0, (hipStream_t)stream,
                     F2(q_sorted), n, K, scale);
  TR_LAUNCH_CHECK();
  return kOk;
}

int tr_data_fit_sums(const tr_c64* a_sorted, const tr_c64* y_sorted, const float* dcf_sorted, int64_t M, int K,
                     double* work, int nblocks, double* out, void* stream) {
  if (!a_sorted || !y_sorted || !work || !out) TR_FAIL(kInvalidArgument, "null argument");
  if (K < 1 || nblocks < 1 || nblocks > 65536) TR_FAIL(kInvalidArgument, "need K >= 1 and nblocks in [1, 65536]");
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(k_fit_partial, dim3(nblocks), dim3(256), 0, s, F2(a_sorted), F2(y_sorted), dcf_sorted, M, K,
                     (double2*)work);
  TR_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_fit_reduce, dim3(grid1d(K, 256)), dim3(256), 0, s, (const double2*)work, nblocks, K,
                     (double2*)out);
  TR_LAUNCH_CHECK();
  return kOk;
}

}  // extern "C"
