// Kaiser-Bessel interpolation (gather at samples, fused with the PDHG dual
// update) and deterministic CSR spreading (gather at grid points).

#include "tr_internal.hpp"

namespace tr {

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
// (sample, weight) in sorted-sample order.  Lanes own channels.  Rows longer
// than kSpreadSeg entries (the k-space centre, where every interleave passes)
// are cut at plan time into fixed segments that separate waves reduce; their
// partial sums are then added in segment order, so the result depends only on
// the plan, never on the launch configuration or scheduling.
// ---------------------------------------------------------------------------
template <int NCI>
__device__ __forceinline__ void spread_load(const SpreadArgs& a, int m, int ch, int kci, float2 (&v)[NCI]) {
  const int K = a.K;
  if (a.sorted_src) {
    const float d = a.dcf_sorted ? a.dcf_sorted[m] : 1.0f;
#pragma unroll
    for (int i = 0; i < NCI; ++i) {
      const int k = i * kci + ch;
      v[i] = k < K ? cscale(a.q[(int64_t)m * K + k], d) : make_float2(0.f, 0.f);
    }
  } else {
    const int mo = a.perm[m];
    const float d = a.dcf_orig ? a.dcf_orig[mo] : 1.0f;
#pragma unroll
    for (int i = 0; i < NCI; ++i) {
      const int k = i * kci + ch;
      v[i] = k < K ? cscale(a.ksp[(int64_t)k * a.M + mo], d) : make_float2(0.f, 0.f);
    }
  }
}

// acc += sum_{j0 <= j < j1} val[j] * src(col[j]) in increasing j (loads of four
// entries are issued together; the additions stay in order).
template <int KCI, int NCI>
__device__ __forceinline__ void spread_accumulate(const SpreadArgs& a, int j0, int j1, int ch,
                                                  float2 (&acc)[NCI]) {
  int j = j0;
  for (; j + 4 <= j1; j += 4) {
    int m[4];
    float w[4];
    float2 v[4][NCI];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      m[u] = a.col[j + u];
      w[u] = a.val[j + u];
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) spread_load<NCI>(a, m[u], ch, KCI, v[u]);
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
      for (int i = 0; i < NCI; ++i) {
        acc[i].x += w[u] * v[u][i].x;
        acc[i].y += w[u] * v[u][i].y;
      }
  }
  for (; j < j1; ++j) {
    float2 v[NCI];
    const float w = a.val[j];
    spread_load<NCI>(a, a.col[j], ch, KCI, v);
#pragma unroll
    for (int i = 0; i < NCI; ++i) {
      acc[i].x += w * v[i].x;
      acc[i].y += w * v[i].y;
    }
  }
}

// Rows with at most kSpreadSeg entries (including empty rows, written as 0).
template <int KCI, int NCI>
__global__ void __launch_bounds__(256) k_spread(SpreadArgs a) {
  constexpr int SPW = 64 / KCI;
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int64_t t = ((int64_t)blockIdx.x * 4 + wave) * SPW + lane / KCI;
  const int ch = lane % KCI;
  if (t >= a.G) return;
  const int j0 = a.row_ptr[t], j1 = a.row_ptr[t + 1];
  if (j1 - j0 > kSpreadSeg) return;  // heavy row: written by k_spread_combine
  float2 acc[NCI];
#pragma unroll
  for (int i = 0; i < NCI; ++i) acc[i] = make_float2(0.f, 0.f);
  spread_accumulate<KCI, NCI>(a, j0, j1, ch, acc);
#pragma unroll
  for (int i = 0; i < NCI; ++i) {
    const int k = i * KCI + ch;
    if (k < a.K) a.grid[t * a.K + k] = acc[i];
  }
}

// One segment of a heavy row -> partial[seg].
template <int KCI, int NCI>
__global__ void __launch_bounds__(256) k_spread_segments(SpreadArgs a) {
  constexpr int SPW = 64 / KCI;
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int64_t sg = ((int64_t)blockIdx.x * 4 + wave) * SPW + lane / KCI;
  const int ch = lane % KCI;
  if (sg >= a.nseg) return;
  const int row = a.seg_row[sg];
  const int j0 = a.seg_j0[sg];
  const int j1 = min(j0 + kSpreadSeg, a.row_ptr[row + 1]);
  float2 acc[NCI];
#pragma unroll
  for (int i = 0; i < NCI; ++i) acc[i] = make_float2(0.f, 0.f);
  spread_accumulate<KCI, NCI>(a, j0, j1, ch, acc);
#pragma unroll
  for (int i = 0; i < NCI; ++i) {
    const int k = i * KCI + ch;
    if (k < a.K) a.partial[sg * a.K + k] = acc[i];
  }
}

// Heavy row = sum of its segment partials in segment order.
template <int KCI, int NCI>
__global__ void __launch_bounds__(256) k_spread_combine(SpreadArgs a) {
  constexpr int SPW = 64 / KCI;
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int64_t h = ((int64_t)blockIdx.x * 4 + wave) * SPW + lane / KCI;
  const int ch = lane % KCI;
  if (h >= a.nheavy) return;
  const int row = a.heavy_row[h];
  float2 acc[NCI];
#pragma unroll
  for (int i = 0; i < NCI; ++i) acc[i] = make_float2(0.f, 0.f);
  for (int sg = a.heavy_seg_ptr[h]; sg < a.heavy_seg_ptr[h + 1]; ++sg) {
#pragma unroll
    for (int i = 0; i < NCI; ++i) {
      const int k = i * KCI + ch;
      if (k < a.K) acc[i] = cadd(acc[i], a.partial[(int64_t)sg * a.K + k]);
    }
  }
#pragma unroll
  for (int i = 0; i < NCI; ++i) {
    const int k = i * KCI + ch;
    if (k < a.K) a.grid[(int64_t)row * a.K + k] = acc[i];
  }
}

int launch_interp(const InterpArgs& a, hipStream_t s) {
  const int K = a.K;
  const double G = (double)a.gg.g[0] * a.gg.g[1] * a.gg.g[2];
  const double per_sample = a.mode == kInterpDual ? 24.0 * K : (a.mode == kInterpPlain ? 8.0 * K + 4.0 : 8.0 * K);
  ProfScope ps(kKInterp, 8.0 * K * G + a.M * (12.0 + per_sample), s);
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

int launch_spread(SpreadArgs a, const BinData& b, tr_plan* p, hipStream_t s) {
  const int K = a.K;
  ProfScope ps(kKSpread, 8.0 * K * a.M + 16.0 * a.M + 8.0 * K * (double)a.G, s);
  const int kci = K >= 64 ? 64 : pow2ceil(K);
  const int nci = (K + kci - 1) / kci;
  if (nci > 4) TR_FAIL(kUnsupported, "more than 256 channels per launch");
  const int spw = 64 / kci;
  a.nseg = b.nseg;
  a.seg_row = b.seg_row;
  a.seg_j0 = b.seg_j0;
  a.nheavy = b.nheavy;
  a.heavy_row = b.heavy_row;
  a.heavy_seg_ptr = b.heavy_seg_ptr;
  if (b.nseg > 0) {
    if (int r = ensure_partial(p, b.nseg * (int64_t)K)) return r;
    a.partial = p->partial;
  }
#define TR_SPREAD_ALL(KERN, n)                                                                   \
  do {                                                                                          \
    const unsigned blocks = grid1d((n), 4 * spw);                                               \
    switch (kci) {                                                                              \
      case 1: hipLaunchKernelGGL((KERN<1, 1>), dim3(blocks), dim3(256), 0, s, a); break;        \
      case 2: hipLaunchKernelGGL((KERN<2, 1>), dim3(blocks), dim3(256), 0, s, a); break;        \
      case 4: hipLaunchKernelGGL((KERN<4, 1>), dim3(blocks), dim3(256), 0, s, a); break;        \
      case 8: hipLaunchKernelGGL((KERN<8, 1>), dim3(blocks), dim3(256), 0, s, a); break;        \
      case 16: hipLaunchKernelGGL((KERN<16, 1>), dim3(blocks), dim3(256), 0, s, a); break;      \
      case 32: hipLaunchKernelGGL((KERN<32, 1>), dim3(blocks), dim3(256), 0, s, a); break;      \
      default:                                                                                  \
        switch (nci) {                                                                          \
          case 1: hipLaunchKernelGGL((KERN<64, 1>), dim3(blocks), dim3(256), 0, s, a); break;   \
          case 2: hipLaunchKernelGGL((KERN<64, 2>), dim3(blocks), dim3(256), 0, s, a); break;   \
          case 3: hipLaunchKernelGGL((KERN<64, 3>), dim3(blocks), dim3(256), 0, s, a); break;   \
          default: hipLaunchKernelGGL((KERN<64, 4>), dim3(blocks), dim3(256), 0, s, a); break;  \
        }                                                                                       \
    }                                                                                           \
    TR_LAUNCH_CHECK();                                                                          \
  } while (0)
  TR_SPREAD_ALL(k_spread, a.G);
  if (b.nseg > 0) {
    TR_SPREAD_ALL(k_spread_segments, b.nseg);
    TR_SPREAD_ALL(k_spread_combine, b.nheavy);
  }
#undef TR_SPREAD_ALL
  return kOk;
}

}  // namespace tr
