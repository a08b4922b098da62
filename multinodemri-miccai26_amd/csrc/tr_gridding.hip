// Kaiser-Bessel interpolation (gather at samples, fused with the PDHG dual
// update) and deterministic spreading (gather over grid-point clusters).

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
// Deterministic adjoint spreading over clusters of kClZ x kClY x kClX grid
// points.  A wave owns one cluster and KCI channels (lanes: KCI channels x
// 64/KCI point slots) and keeps the cluster's sums in registers while it walks
// the cluster's (sample, weights) list once, in sorted-sample order.  Lists
// longer than kSpreadSeg pairs (the k-space centre, where every interleave
// passes) are cut at plan time into fixed segments reduced by separate waves;
// their partial sums are then added in segment order, so the result depends
// only on the plan, never on the launch configuration or scheduling.
// ---------------------------------------------------------------------------
template <int KCI>
struct ClLanes {
  static constexpr int NPL = 64 / KCI;                   // point slots per wave
  static constexpr int NP = (kClPts + NPL - 1) / NPL;    // points per lane
};

template <int N>
__device__ __forceinline__ float pick(const float (&w)[N], int i) {
  float r = w[0];
#pragma unroll
  for (int t = 1; t < N; ++t) r = i == t ? w[t] : r;
  return r;
}

// source value of sorted sample m, channel k (zero past K)
template <int SORTED>
__device__ __forceinline__ float2 spread_src(const SpreadArgs& a, int m, int k) {
  if (k >= a.K) return make_float2(0.f, 0.f);
  if (SORTED) {
    const float d = a.dcf_sorted ? a.dcf_sorted[m] : 1.0f;
    return cscale(a.q[(int64_t)m * a.K + k], d);
  }
  const int mo = a.perm[m];
  const float d = a.dcf_orig ? a.dcf_orig[mo] : 1.0f;
  return cscale(a.ksp[(int64_t)k * a.M + mo], d);
}

template <int KCI>
__device__ __forceinline__ void cl_add(const ClusterPair& r, float2 v, int pl, float2 (&acc)[ClLanes<KCI>::NP]) {
  constexpr int NPL = ClLanes<KCI>::NPL;
#pragma unroll
  for (int i = 0; i < ClLanes<KCI>::NP; ++i) {
    const int pt = pl + i * NPL;
    const int iz = pt / (kClY * kClX), iy = (pt / kClX) % kClY, ix = pt % kClX;
    const float w = (pick(r.wz, iz) * pick(r.wy, iy)) * pick(r.wx, ix);
    acc[i].x += w * v.x;
    acc[i].y += w * v.y;
  }
}

// acc += sum over pairs [j0, j1) in increasing order (four pairs' loads are
// issued together; the additions stay in order)
template <int KCI, int SORTED>
__device__ __forceinline__ void cl_run(const SpreadArgs& a, int j0, int j1, int k, int pl,
                                       float2 (&acc)[ClLanes<KCI>::NP]) {
  int j = j0;
  for (; j + 4 <= j1; j += 4) {
    ClusterPair r[4];
    float2 v[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) r[u] = a.pairs[j + u];
#pragma unroll
    for (int u = 0; u < 4; ++u) v[u] = spread_src<SORTED>(a, r[u].m, k);
#pragma unroll
    for (int u = 0; u < 4; ++u) cl_add<KCI>(r[u], v[u], pl, acc);
  }
  for (; j < j1; ++j) {
    const ClusterPair r = a.pairs[j];
    cl_add<KCI>(r, spread_src<SORTED>(a, r.m, k), pl, acc);
  }
}

// Write a cluster's sums to the grid (points past the grid edge are dropped).
template <int KCI>
__device__ __forceinline__ void cl_store(const SpreadArgs& a, int64_t cl, int k, int pl,
                                         const float2 (&acc)[ClLanes<KCI>::NP]) {
  constexpr int NPL = ClLanes<KCI>::NPL;
  if (k >= a.K) return;
  const int ncy = a.ncl[1], ncx = a.ncl[2];
  const int cz = (int)(cl / ((int64_t)ncy * ncx));
  const int rem = (int)(cl - (int64_t)cz * ncy * ncx);
  const int cy = rem / ncx, cx = rem - (rem / ncx) * ncx;
  const int gz = a.gg.g[0], gy = a.gg.g[1], gx = a.gg.g[2];
#pragma unroll
  for (int i = 0; i < ClLanes<KCI>::NP; ++i) {
    const int pt = pl + i * NPL;
    if (pt >= kClPts) continue;
    const int z = cz * kClZ + pt / (kClY * kClX);
    const int y = cy * kClY + (pt / kClX) % kClY;
    const int x = cx * kClX + pt % kClX;
    if (z < gz && y < gy && x < gx) a.grid[(((int64_t)z * gy + y) * gx + x) * a.K + k] = acc[i];
  }
}

// Clusters with at most kSpreadSeg pairs (empty clusters are written as 0).
template <int KCI, int SORTED>
__global__ void __launch_bounds__(256) k_spread(SpreadArgs a) {
  constexpr int NP = ClLanes<KCI>::NP;
  const int lane = threadIdx.x & 63;
  const int64_t item = (int64_t)blockIdx.x * 4 + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  if (item >= a.NCL * a.ngrp) return;
  const int64_t cl = item / a.ngrp;
  const int grp = (int)(item - cl * a.ngrp);
  const int j0 = a.row_ptr[cl], j1 = a.row_ptr[cl + 1];
  if (j1 - j0 > kSpreadSeg) return;  // heavy cluster: written by k_spread_combine
  const int k = grp * KCI + lane % KCI;
  const int pl = lane / KCI;
  float2 acc[NP];
#pragma unroll
  for (int i = 0; i < NP; ++i) acc[i] = make_float2(0.f, 0.f);
  cl_run<KCI, SORTED>(a, j0, j1, k, pl, acc);
  cl_store<KCI>(a, cl, k, pl, acc);
}

// One segment of a heavy cluster -> partial[seg][point][K].
template <int KCI, int SORTED>
__global__ void __launch_bounds__(256) k_spread_segments(SpreadArgs a) {
  constexpr int NP = ClLanes<KCI>::NP;
  constexpr int NPL = ClLanes<KCI>::NPL;
  const int lane = threadIdx.x & 63;
  const int64_t item = (int64_t)blockIdx.x * 4 + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  if (item >= a.nseg * a.ngrp) return;
  const int64_t sg = item / a.ngrp;
  const int grp = (int)(item - sg * a.ngrp);
  const int cl = a.seg_row[sg];
  const int j0 = a.seg_j0[sg];
  const int j1 = min(j0 + kSpreadSeg, a.row_ptr[cl + 1]);
  const int k = grp * KCI + lane % KCI;
  const int pl = lane / KCI;
  float2 acc[NP];
#pragma unroll
  for (int i = 0; i < NP; ++i) acc[i] = make_float2(0.f, 0.f);
  cl_run<KCI, SORTED>(a, j0, j1, k, pl, acc);
  if (k >= a.K) return;
#pragma unroll
  for (int i = 0; i < NP; ++i) {
    const int pt = pl + i * NPL;
    if (pt < kClPts) a.partial[(sg * kClPts + pt) * a.K + k] = acc[i];
  }
}

// Heavy cluster = sum of its segment partials in segment order.
template <int KCI, int SORTED>
__global__ void __launch_bounds__(256) k_spread_combine(SpreadArgs a) {
  constexpr int NP = ClLanes<KCI>::NP;
  constexpr int NPL = ClLanes<KCI>::NPL;
  const int lane = threadIdx.x & 63;
  const int64_t item = (int64_t)blockIdx.x * 4 + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  if (item >= a.nheavy * a.ngrp) return;
  const int64_t h = item / a.ngrp;
  const int grp = (int)(item - h * a.ngrp);
  const int cl = a.heavy_row[h];
  const int k = grp * KCI + lane % KCI;
  const int pl = lane / KCI;
  if (k >= a.K) return;
  float2 acc[NP];
#pragma unroll
  for (int i = 0; i < NP; ++i) acc[i] = make_float2(0.f, 0.f);
  for (int sg = a.heavy_seg_ptr[h]; sg < a.heavy_seg_ptr[h + 1]; ++sg) {
#pragma unroll
    for (int i = 0; i < NP; ++i) {
      const int pt = pl + i * NPL;
      if (pt < kClPts) acc[i] = cadd(acc[i], a.partial[((int64_t)sg * kClPts + pt) * a.K + k]);
    }
  }
  cl_store<KCI>(a, cl, k, pl, acc);
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
  const double G = (double)p->G;
  ProfScope ps(kKSpread, 8.0 * K * a.M + 16.0 * a.M + 8.0 * K * G, s);
  const int kci = K >= 64 ? 64 : pow2ceil(K);
  a.gg = geom(p);
  for (int i = 0; i < 3; ++i) a.ncl[i] = p->ncl[i];
  a.NCL = p->NCL;
  a.row_ptr = b.row_ptr;
  a.pairs = b.pairs;
  a.ngrp = (K + kci - 1) / kci;
  a.nseg = b.nseg;
  a.seg_row = b.seg_row;
  a.seg_j0 = b.seg_j0;
  a.nheavy = b.nheavy;
  a.heavy_row = b.heavy_row;
  a.heavy_seg_ptr = b.heavy_seg_ptr;
  if (b.nseg > 0) {
    if (int r = ensure_partial(p, b.nseg * (int64_t)kClPts * K)) return r;
    a.partial = p->partial;
  }
#define TR_SPREAD_KCI(KERN, SRC, n)                                                                \
  do {                                                                                            \
    const unsigned blocks = grid1d((n) * a.ngrp, 4);                                              \
    switch (kci) {                                                                                \
      case 1: hipLaunchKernelGGL((KERN<1, SRC>), dim3(blocks), dim3(256), 0, s, a); break;        \
      case 2: hipLaunchKernelGGL((KERN<2, SRC>), dim3(blocks), dim3(256), 0, s, a); break;        \
      case 4: hipLaunchKernelGGL((KERN<4, SRC>), dim3(blocks), dim3(256), 0, s, a); break;        \
      case 8: hipLaunchKernelGGL((KERN<8, SRC>), dim3(blocks), dim3(256), 0, s, a); break;        \
      case 16: hipLaunchKernelGGL((KERN<16, SRC>), dim3(blocks), dim3(256), 0, s, a); break;      \
      case 32: hipLaunchKernelGGL((KERN<32, SRC>), dim3(blocks), dim3(256), 0, s, a); break;      \
      default: hipLaunchKernelGGL((KERN<64, SRC>), dim3(blocks), dim3(256), 0, s, a); break;      \
    }                                                                                             \
    TR_LAUNCH_CHECK();                                                                            \
  } while (0)
#define TR_SPREAD_ALL(KERN, n)                          \
  do {                                                  \
    if (a.sorted_src) TR_SPREAD_KCI(KERN, 1, n);        \
    else TR_SPREAD_KCI(KERN, 0, n);                     \
  } while (0)
  TR_SPREAD_ALL(k_spread, a.NCL);
  if (b.nseg > 0) {
    TR_SPREAD_ALL(k_spread_segments, b.nseg);
    TR_SPREAD_ALL(k_spread_combine, b.nheavy);
  }
#undef TR_SPREAD_ALL
#undef TR_SPREAD_KCI
  return kOk;
}

}  // namespace tr
