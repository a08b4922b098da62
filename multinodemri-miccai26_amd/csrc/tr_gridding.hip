// Kaiser-Bessel interpolation (gather at samples, fused with the PDHG dual
// update) and deterministic spreading (gather over grid-point clusters).

#include "tr_internal.hpp"

namespace tr {

// ---------------------------------------------------------------------------
// Kaiser-Bessel gather at sorted samples.  Lanes own channels (KCI per
// sample, 64/KCI samples per wave, NCI channel groups per lane).  The taps
// come from the plan's per-sample records; with KCI = 64 the sample index is
// wave-uniform, so the record and all tap addresses live in scalar registers.
// ---------------------------------------------------------------------------
__device__ __forceinline__ int wrap2(int p, int g) {
  // p < p0 + 5 with 0 <= p0 < g and g >= 2: two conditional subtractions suffice
  p = p >= g ? p - g : p;
  return p >= g ? p - g : p;
}

template <int KCI, int NCI>
__global__ void __launch_bounds__(256) k_interp(InterpArgs a) {
  if (gate_closed(a.gate)) return;
  constexpr int SPW = 64 / KCI;
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const unsigned blk = xcd_block(blockIdx.x, gridDim.x, a.xcd_chunk);
  const int64_t m = ((int64_t)blk * 4 + wave) * SPW + lane / KCI;
  const int ch = lane % KCI;
  if (m >= a.M) return;
  const int gz = a.gg.g[0], gy = a.gg.g[1], gx = a.gg.g[2];
  const int kbase = blockIdx.y * KCI * NCI;  // channel group (gridDim.y > 1 in group-major mode)
  const TapRec& r = a.tap[m];
  const int packed = r.n;
  const int nx = (packed >> 8) & 15, ny = (packed >> 4) & 15;
  const int nz = (nx == 0 || ny == 0) ? 0 : (packed & 15);  // no taps on one axis: no taps at all
  const int p0z = r.p0[0], p0y = r.p0[1], p0x = r.p0[2];
  float2 acc[NCI];
#pragma unroll
  for (int i = 0; i < NCI; ++i) acc[i] = make_float2(0.f, 0.f);
  const int K = a.K;
  // channels past K read channel K-1 (in bounds, never stored), so that no
  // load sits behind a per-lane branch and a whole tap row is in flight at once
  int kk[NCI];
#pragma unroll
  for (int i = 0; i < NCI; ++i) kk[i] = min(kbase + i * KCI + ch, K - 1);
  int px[5];
#pragma unroll
  for (int dx = 0; dx < 5; ++dx) px[dx] = wrap2(p0x + dx, gx);
  // One z plane at a time: the 4 x 4 taps of its first four y rows (every sample has them) are
  // loaded together before any is used, so 16 * NCI loads are in flight per wave; the rare fifth
  // y row / x tap (a coordinate on a grid line) follow.  Products are added in (dz, dy, dx) order.
  for (int dz = 0; dz < nz; ++dz) {
    const int pz = wrap2(p0z + dz, gz);
    const float wz = r.w[dz];
    float2 v[4][4][NCI];
#pragma unroll
    for (int dy = 0; dy < 4; ++dy) {
      const float2* row = a.grid + (((int64_t)pz * gy + wrap2(p0y + dy, gy)) * gx) * K;
#pragma unroll
      for (int dx = 0; dx < 4; ++dx)
#pragma unroll
        for (int i = 0; i < NCI; ++i) v[dy][dx][i] = row[(int64_t)px[dx] * K + kk[i]];
    }
#pragma unroll
    for (int dy = 0; dy < 5; ++dy) {
      if (dy == 4 && ny < 5) break;
      const float wzy = wz * r.w[5 + dy];
      const float2* row = a.grid + (((int64_t)pz * gy + wrap2(p0y + dy, gy)) * gx) * K;
      if (dy < 4) {
#pragma unroll
        for (int dx = 0; dx < 4; ++dx) {
          const float w = wzy * r.w[10 + dx];
#pragma unroll
          for (int i = 0; i < NCI; ++i) {
            acc[i].x += w * v[dy][dx][i].x;
            acc[i].y += w * v[dy][dx][i].y;
          }
        }
      } else {
#pragma unroll
        for (int dx = 0; dx < 4; ++dx) {
          const float w = wzy * r.w[10 + dx];
#pragma unroll
          for (int i = 0; i < NCI; ++i) {
            const float2 u = row[(int64_t)px[dx] * K + kk[i]];
            acc[i].x += w * u.x;
            acc[i].y += w * u.y;
          }
        }
      }
      if (nx == 5) {
        const float w = wzy * r.w[14];
#pragma unroll
        for (int i = 0; i < NCI; ++i) {
          const float2 u = row[(int64_t)px[4] * K + kk[i]];
          acc[i].x += w * u.x;
          acc[i].y += w * u.y;
        }
      }
    }
  }
#pragma unroll
  for (int i = 0; i < NCI; ++i) {
    const int k = kbase + i * KCI + ch;
    if (k >= K) continue;
    if (a.mode == kInterpPlain) {
      a.out[(int64_t)k * a.M + a.perm[m]] = acc[i];
    } else if (a.mode == kInterpDual) {
      const int64_t idx = m * K + k;
      const float2 rr = csub(acc[i], a.y[idx]);
      const float2 qo = a.q[idx];
      const float2 t = make_float2(qo.x + a.sigma * rr.x, qo.y + a.sigma * rr.y);
      a.q[idx] = make_float2(t.x / a.one_plus_sigma, t.y / a.one_plus_sigma);
    } else if (a.mode == kInterpResid) {
      const int64_t idx = m * K + k;
      a.q[idx] = csub(acc[i], a.y[idx]);
    } else {
      a.q[m * K + k] = acc[i];
    }
  }
}

// ---------------------------------------------------------------------------
// The same gather with 16-byte loads: lane l of the wave owns the channel pair
// (kbase + 2l, kbase + 2l + 1) of a group of G = gridDim.y-th of the channels
// (G even, at most 128), so one load instruction moves up to 1 KB of a grid
// point's channel record.  One sample per wave (record in scalar registers);
// channel groups are launch rows (group-major).  Per channel the products are
// added in the same (dz, dy, dx) order as k_interp, so results are identical.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_interp_v4(InterpArgs a, int G) {
  if (gate_closed(a.gate)) return;
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const unsigned blk = xcd_block(blockIdx.x, gridDim.x, a.xcd_chunk);
  const int64_t m = (int64_t)blk * 4 + wave;
  if (m >= a.M) return;
  const int gz = a.gg.g[0], gy = a.gg.g[1], gx = a.gg.g[2];
  const int K = a.K;
  const int k2 = blockIdx.y * G + 2 * lane;         // first channel of this lane's pair
  const bool live = 2 * lane < G && k2 < K;
  // float4 index of the lane's pair; a dead lane reads the group's first pair, a line the live lanes
  // fetch anyway, so the gather touches no cache line outside the group's channels
  const int kk = min(live ? k2 : (int)blockIdx.y * G, K - 2) >> 1;
  const TapRec& r = a.tap[m];
  const int packed = r.n;
  const int nx = (packed >> 8) & 15, ny = (packed >> 4) & 15;
  const int nz = (nx == 0 || ny == 0) ? 0 : (packed & 15);
  const int p0z = r.p0[0], p0y = r.p0[1], p0x = r.p0[2];
  const int K2 = K >> 1;
  int px[5];
#pragma unroll
  for (int dx = 0; dx < 5; ++dx) px[dx] = wrap2(p0x + dx, gx);
  const float4* grid4 = reinterpret_cast<const float4*>(a.grid);
  float4 acc = make_float4(0.f, 0.f, 0.f, 0.f);
  for (int dz = 0; dz < nz; ++dz) {
    const int pz = wrap2(p0z + dz, gz);
    const float wz = r.w[dz];
    float4 v[4][4];
#pragma unroll
    for (int dy = 0; dy < 4; ++dy) {
      const float4* row = grid4 + (((int64_t)pz * gy + wrap2(p0y + dy, gy)) * gx) * K2;
#pragma unroll
      for (int dx = 0; dx < 4; ++dx) v[dy][dx] = row[(int64_t)px[dx] * K2 + kk];
    }
#pragma unroll
    for (int dy = 0; dy < 5; ++dy) {
      if (dy == 4 && ny < 5) break;
      const float wzy = wz * r.w[5 + dy];
      const float4* row = grid4 + (((int64_t)pz * gy + wrap2(p0y + dy, gy)) * gx) * K2;
#pragma unroll
      for (int dx = 0; dx < 4; ++dx) {
        const float w = wzy * r.w[10 + dx];
        const float4 u = dy < 4 ? v[dy][dx] : row[(int64_t)px[dx] * K2 + kk];
        acc.x += w * u.x;
        acc.y += w * u.y;
        acc.z += w * u.z;
        acc.w += w * u.w;
      }
      if (nx == 5) {
        const float w = wzy * r.w[14];
        const float4 u = row[(int64_t)px[4] * K2 + kk];
        acc.x += w * u.x;
        acc.y += w * u.y;
        acc.z += w * u.z;
        acc.w += w * u.w;
      }
    }
  }
  if (!live) return;
  const float2 a0 = make_float2(acc.x, acc.y), a1 = make_float2(acc.z, acc.w);
  if (a.mode == kInterpPlain) {
    const int64_t o = a.perm[m];
    a.out[(int64_t)k2 * a.M + o] = a0;
    a.out[(int64_t)(k2 + 1) * a.M + o] = a1;
  } else {
    float4* q4 = reinterpret_cast<float4*>(a.q) + m * K2 + (k2 >> 1);
    if (a.mode == kInterpDual) {
      const float4 yy = reinterpret_cast<const float4*>(a.y)[m * K2 + (k2 >> 1)];
      const float4 qo = *q4;
      const float2 r0 = csub(a0, make_float2(yy.x, yy.y)), r1 = csub(a1, make_float2(yy.z, yy.w));
      const float t0x = qo.x + a.sigma * r0.x, t0y = qo.y + a.sigma * r0.y;
      const float t1x = qo.z + a.sigma * r1.x, t1y = qo.w + a.sigma * r1.y;
      *q4 = make_float4(t0x / a.one_plus_sigma, t0y / a.one_plus_sigma, t1x / a.one_plus_sigma,
                        t1y / a.one_plus_sigma);
    } else if (a.mode == kInterpResid) {
      const float4 yy = reinterpret_cast<const float4*>(a.y)[m * K2 + (k2 >> 1)];
      *q4 = make_float4(a0.x - yy.x, a0.y - yy.y, a1.x - yy.z, a1.y - yy.w);
    } else {
      *q4 = make_float4(a0.x, a0.y, a1.x, a1.y);
    }
  }
}

// ---------------------------------------------------------------------------
// Grouped gather: one wave per group of up to kTapGroup consecutive sorted
// samples whose taps fit in one box of at most 5 positions per axis (dense
// k-space regions hold many), 16-byte loads as k_interp_v4.  The group's
// records (a.tapg) give every member the box's origin and extent and the
// member's weights shifted into the box, zero where it has no tap.  Each tap
// is loaded once for the group and added into every member's sums with the
// member's own weights, in the same (dz, dy, dx) order and with the same
// products as the per-sample kernels; the extra terms have weight 0, so every
// sample's result is bit-identical to theirs.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void acc4(float4& acc, float w, const float4& u) {
  acc.x += w * u.x;
  acc.y += w * u.y;
  acc.z += w * u.z;
  acc.w += w * u.w;
}

__global__ void __launch_bounds__(256) k_interp_grp(InterpArgs a, int G) {
  if (gate_closed(a.gate)) return;
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const unsigned blk = xcd_block(blockIdx.x, gridDim.x, a.xcd_chunk);
  const int64_t gi = (int64_t)blk * 4 + wave;
  if (gi >= a.ngroups) return;
  const int m0 = a.grp[gi];
  const int gs = a.grp[gi + 1] - m0;  // 1 .. kTapGroup, wave-uniform
  const int gz = a.gg.g[0], gy = a.gg.g[1], gx = a.gg.g[2];
  const int K = a.K;
  const int k2 = blockIdx.y * G + 2 * lane;
  const bool live = 2 * lane < G && k2 < K;
  const int kk = min(live ? k2 : (int)blockIdx.y * G, K - 2) >> 1;  // dead lanes: the group's first pair
  const TapRec* rec = a.tapg + m0;
  const int packed = rec[0].n;
  const int nx = (packed >> 8) & 15, ny = (packed >> 4) & 15;
  const int nz = (nx == 0 || ny == 0) ? 0 : (packed & 15);
  const int p0z = rec[0].p0[0], p0y = rec[0].p0[1], p0x = rec[0].p0[2];
  const int K2 = K >> 1;
  int px[5];
#pragma unroll
  for (int dx = 0; dx < 5; ++dx) px[dx] = wrap2(p0x + dx, gx);
  const float4* grid4 = reinterpret_cast<const float4*>(a.grid);
  float4 acc[kTapGroup];
#pragma unroll
  for (int t = 0; t < kTapGroup; ++t) acc[t] = make_float4(0.f, 0.f, 0.f, 0.f);
  constexpr int RB = kInterpRows;
  // The members' y and x weights are read once per group into scalar registers and the z weights
  // once per plane, so no row waits for a scalar load (reading them from the records inside the
  // row loop put a scalar load and its wait in front of every member's row: 2.71 -> 2.62 ms).
  // Members past the group size read the last member's record and are never used.
  float wyv[kTapGroup][5], wxv[kTapGroup][5];
#pragma unroll
  for (int t = 0; t < kTapGroup; ++t) {
    const TapRec& r = rec[min(t, gs - 1)];
#pragma unroll
    for (int j = 0; j < 5; ++j) {
      wyv[t][j] = r.w[5 + j];
      wxv[t][j] = r.w[10 + j];
    }
  }
  for (int dz = 0; dz < nz; ++dz) {
    const int pz = wrap2(p0z + dz, gz);
    float wzv[kTapGroup];
#pragma unroll
    for (int t = 0; t < kTapGroup; ++t) wzv[t] = rec[min(t, gs - 1)].w[dz];
#pragma unroll
    for (int d0 = 0; d0 < 4; d0 += RB) {
      float4 v[RB][4];
#pragma unroll
      for (int i = 0; i < RB; ++i) {
        const float4* row = grid4 + (((int64_t)pz * gy + wrap2(p0y + d0 + i, gy)) * gx) * K2;
#pragma unroll
        for (int dx = 0; dx < 4; ++dx) v[i][dx] = row[(int64_t)px[dx] * K2 + kk];
      }
#pragma unroll
      for (int i = 0; i < RB + 1; ++i) {
        const int dy = d0 + i;
        if (i == RB && (dy < 4 || ny < 5)) break;  // the fifth row follows the last batch
        const float4* row = grid4 + (((int64_t)pz * gy + wrap2(p0y + dy, gy)) * gx) * K2;
        float4 u[5];
#pragma unroll
        for (int dx = 0; dx < 4; ++dx) u[dx] = i < RB ? v[i < RB ? i : 0][dx] : row[(int64_t)px[dx] * K2 + kk];
        u[4] = nx == 5 ? row[(int64_t)px[4] * K2 + kk] : make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
        for (int t = 0; t < kTapGroup; ++t) {
          if (t < gs) {
            const float wzy = wzv[t] * wyv[t][dy];
#pragma unroll
            for (int dx = 0; dx < 4; ++dx) acc4(acc[t], wzy * wxv[t][dx], u[dx]);
            if (nx == 5) acc4(acc[t], wzy * wxv[t][4], u[4]);
          }
        }
      }
    }
  }
  if (!live) return;
#pragma unroll
  for (int t = 0; t < kTapGroup; ++t) {
    if (t >= gs) break;
    const int64_t m = m0 + t;
    const float2 a0 = make_float2(acc[t].x, acc[t].y), a1 = make_float2(acc[t].z, acc[t].w);
    if (a.mode == kInterpPlain) {
      const int64_t o = a.perm[m];
      a.out[(int64_t)k2 * a.M + o] = a0;
      a.out[(int64_t)(k2 + 1) * a.M + o] = a1;
    } else {
      float4* q4 = reinterpret_cast<float4*>(a.q) + m * K2 + (k2 >> 1);
      if (a.mode == kInterpDual) {
        const float4 yy = reinterpret_cast<const float4*>(a.y)[m * K2 + (k2 >> 1)];
        const float4 qo = *q4;
        const float2 r0 = csub(a0, make_float2(yy.x, yy.y)), r1 = csub(a1, make_float2(yy.z, yy.w));
        const float t0x = qo.x + a.sigma * r0.x, t0y = qo.y + a.sigma * r0.y;
        const float t1x = qo.z + a.sigma * r1.x, t1y = qo.w + a.sigma * r1.y;
        *q4 = make_float4(t0x / a.one_plus_sigma, t0y / a.one_plus_sigma, t1x / a.one_plus_sigma,
                          t1y / a.one_plus_sigma);
      } else if (a.mode == kInterpResid) {
        const float4 yy = reinterpret_cast<const float4*>(a.y)[m * K2 + (k2 >> 1)];
        *q4 = make_float4(a0.x - yy.x, a0.y - yy.y, a1.x - yy.z, a1.y - yy.w);
      } else {
        *q4 = make_float4(a0.x, a0.y, a1.x, a1.y);
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Deterministic adjoint spreading over clusters of kClZ x kClY x kClX grid
// points.  A wave owns one cluster and KCI channels (lanes: KCI channels x
// 64/KCI point slots) and keeps the cluster's sums in registers while it walks
// the cluster's (sample, weights) list once, in sorted-sample order.  Lists
// longer than the plan's segment length (the k-space centre, where every interleave
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

// source value of sorted sample m, channel kk (kk is clamped to K-1 by the
// caller: lanes past K load a valid channel whose sums are never stored)
template <int SORTED, int DCF>
__device__ __forceinline__ float2 spread_src(const SpreadArgs& a, int m, int kk) {
  if (SORTED) {
    const float2 v = a.q[(int64_t)m * a.K + kk];
    return DCF ? cscale(v, a.dcf_sorted[m]) : v;
  }
  const int mo = a.perm[m];
  const float2 v = a.ksp[(int64_t)kk * a.M + mo];
  return DCF ? cscale(v, a.dcf_orig[mo]) : v;
}

// acc[z][y][x] += (wz[z] wy[y]) * (wx[x] v): the point weights factor over
// the axes, so each pair costs 2 multiply-adds per point.
template <int KCI>
__device__ __forceinline__ void cl_add(const ClusterPair& r, float2 v, int pl, float2 (&acc)[ClLanes<KCI>::NP]) {
  constexpr int NPL = ClLanes<KCI>::NPL;
  float2 tx[kClX];
#pragma unroll
  for (int ix = 0; ix < kClX; ++ix) tx[ix] = make_float2(r.wx[ix] * v.x, r.wx[ix] * v.y);
  if constexpr (NPL == 1) {
    // one wave per (cluster, 64 channels): the pair and its row mask are wave-uniform, so the
    // rows its taps miss are skipped with scalar branches
#pragma unroll
    for (int row = 0; row < kClZ * kClY; ++row) {
      if (r.rows & (1 << row)) {
        const float wzy = r.wz[row / kClY] * r.wy[row % kClY];
#pragma unroll
        for (int ix = 0; ix < kClX; ++ix) {
          acc[row * kClX + ix].x += wzy * tx[ix].x;
          acc[row * kClX + ix].y += wzy * tx[ix].y;
        }
      }
    }
  } else {
#pragma unroll
    for (int i = 0; i < ClLanes<KCI>::NP; ++i) {
      const int pt = pl + i * NPL;
      const int iz = pt / (kClY * kClX), iy = (pt / kClX) % kClY, ix = pt % kClX;
      const float wzy = pick(r.wz, iz) * pick(r.wy, iy);
      float2 t = tx[0];
#pragma unroll
      for (int c = 1; c < kClX; ++c) t = ix == c ? tx[c] : t;
      acc[i].x += wzy * t.x;
      acc[i].y += wzy * t.y;
    }
  }
}

// acc += sum over pairs [j0, j1) in increasing order (the sample loads of
// kSpreadUnroll pairs are issued together; the additions stay in order)
template <int KCI, int SORTED, int DCF>
__device__ __forceinline__ void cl_run_impl(const SpreadArgs& a, int j0, int j1, int kk, int pl,
                                            float2 (&acc)[ClLanes<KCI>::NP]) {
  int j = j0;
  constexpr int U = kSpreadUnroll;
  for (; j + U <= j1; j += U) {
    ClusterPair r[U];
    float2 v[U];
#pragma unroll
    for (int u = 0; u < U; ++u) r[u] = a.pairs[j + u];
#pragma unroll
    for (int u = 0; u < U; ++u) v[u] = spread_src<SORTED, DCF>(a, r[u].m, kk);
#pragma unroll
    for (int u = 0; u < U; ++u) cl_add<KCI>(r[u], v[u], pl, acc);
  }
  for (; j < j1; ++j) {
    const ClusterPair r = a.pairs[j];
    cl_add<KCI>(r, spread_src<SORTED, DCF>(a, r.m, kk), pl, acc);
  }
}

template <int KCI, int SORTED>
__device__ __forceinline__ void cl_run(const SpreadArgs& a, int j0, int j1, int k, int pl,
                                       float2 (&acc)[ClLanes<KCI>::NP]) {
  const int kk = min(k, a.K - 1);
  if (SORTED ? a.dcf_sorted != nullptr : a.dcf_orig != nullptr)
    cl_run_impl<KCI, SORTED, 1>(a, j0, j1, kk, pl, acc);
  else
    cl_run_impl<KCI, SORTED, 0>(a, j0, j1, kk, pl, acc);
}

// Write a cluster's sums to the grid (points past the grid edge are dropped).
template <int KCI>
__device__ __forceinline__ void cl_store(const SpreadArgs& a, int64_t cl, int k, int pl,
                                         const float2 (&acc)[ClLanes<KCI>::NP]) {
  constexpr int NPL = ClLanes<KCI>::NPL;
  if (k >= a.K) return;
  const int gz = a.gg.g[0], gy = a.gg.g[1], gx = a.gg.g[2];
  if constexpr (NPL == 1) {
    // one wave per (cluster, 64 channels): everything but the channel is wave-uniform, so each
    // (y, x) column reads its z window once and its kClZ points sit at one uniform stride from a
    // column base (scalar registers); per point that leaves a window test and a store.  Cluster
    // coordinates come from precomputed inverses of the cluster-grid sizes.
    const unsigned ucl = (unsigned)cl;
    const int cz = (int)a.div_plane.div(ucl);
    const unsigned rem = ucl - (unsigned)cz * a.div_plane.d;
    const int cy = (int)a.div_ncx.div(rem), cx = (int)(rem - (unsigned)cy * a.div_ncx.d);
    const int64_t zs = (int64_t)gy * gx * a.K;
#pragma unroll
    for (int iy = 0; iy < kClY; ++iy) {
#pragma unroll
      for (int ix = 0; ix < kClX; ++ix) {
        const int y = cy * kClY + iy, x = cx * kClX + ix;
        if (y >= gy || x >= gx) continue;
        const int2 w = a.col_win ? a.col_win[y * gx + x] : make_int2(0, gz);
        float2* col = a.grid + ((int64_t)y * gx + x) * a.K + (int64_t)(cz * kClZ) * zs;
#pragma unroll
        for (int iz = 0; iz < kClZ; ++iz) {
          const int z = cz * kClZ + iz;
          if (z < gz && in_window(z, w, gz)) col[iz * zs + k] = acc[(iz * kClY + iy) * kClX + ix];
        }
      }
    }
  } else {
    const int ncy = a.ncl[1], ncx = a.ncl[2];
    const int cz = (int)(cl / ((int64_t)ncy * ncx));
    const int rem = (int)(cl - (int64_t)cz * ncy * ncx);
    const int cy = rem / ncx, cx = rem - (rem / ncx) * ncx;
#pragma unroll
    for (int i = 0; i < ClLanes<KCI>::NP; ++i) {
      const int pt = pl + i * NPL;
      if (pt >= kClPts) continue;
      const int z = cz * kClZ + pt / (kClY * kClX);
      const int y = cy * kClY + (pt / kClX) % kClY;
      const int x = cx * kClX + pt % kClX;
      if (z < gz && y < gy && x < gx && (!a.col_win || in_window(z, a.col_win[y * gx + x], gz)))
        a.grid[(((int64_t)z * gy + y) * gx + x) * a.K + k] = acc[i];
    }
  }
}

// Clusters with at most a.seg pairs (empty clusters are written as 0).
// With a.seg_blocks > 0 the first seg_blocks workgroups run the heavy-cluster segments (the
// longest waves first, their sums to a.partial), the rest the light clusters; both share one
// pair walk so the kernel keeps the register footprint of either.
template <int KCI, int SORTED>
__global__ void __launch_bounds__(256) k_spread(SpreadArgs a) {
  if (gate_closed(a.gate)) return;
  constexpr int NP = ClLanes<KCI>::NP;
  constexpr int NPL = ClLanes<KCI>::NPL;
  const int lane = threadIdx.x & 63;
  const unsigned wave = (unsigned)__builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  unsigned blk = blockIdx.x;
  const bool seg = blk < a.seg_blocks;
  int64_t cl, sg = 0;
  int grp, j0, j1;
  if (seg) {
    const int64_t item = (int64_t)blk * 4 + wave;
    if (item >= a.nseg * a.ngrp) return;
    const unsigned usg = a.div_ngrp.div((unsigned)item);
    sg = usg;
    grp = (int)((unsigned)item - usg * (unsigned)a.ngrp);
    cl = a.seg_row[sg];
    j0 = a.seg_j0[sg];
    j1 = min(j0 + a.seg, a.row_ptr[cl + 1]);
  } else {
    blk -= a.seg_blocks;
    if (blk < a.xcd_full) {  // runs of xcd_chunk consecutive logical blocks per XCD
      const unsigned r = blk / kNumXcd, x = blk % kNumXcd, q = a.div_xcd.div(r);
      blk = q * (kNumXcd * (unsigned)a.xcd_chunk) + x * (unsigned)a.xcd_chunk + (r - q * (unsigned)a.xcd_chunk);
    }
    const unsigned item = blk * 4u + wave;
    const unsigned nitems = a.active ? (unsigned)(a.nactive * a.ngrp) : (unsigned)(a.NCL * a.ngrp);
    if (item >= nitems) return;
    const unsigned uit = a.div_ngrp.div(item);
    grp = (int)(item - uit * (unsigned)a.ngrp);
    cl = a.active ? (unsigned)a.active[uit] : uit;
    j0 = a.row_ptr[cl];
    j1 = a.row_ptr[cl + 1];
    if (!a.active && j1 - j0 > a.seg) return;  // heavy cluster: written by k_spread_combine
  }
  const int k = grp * KCI + lane % KCI;
  const int pl = lane / KCI;
  float2 acc[NP];
#pragma unroll
  for (int i = 0; i < NP; ++i) acc[i] = make_float2(0.f, 0.f);
  cl_run<KCI, SORTED>(a, j0, j1, k, pl, acc);
  if (!seg) {
    cl_store<KCI>(a, cl, k, pl, acc);
  } else if (k < a.K) {
#pragma unroll
    for (int i = 0; i < NP; ++i) {
      const int pt = pl + i * NPL;
      if (pt < kClPts) a.partial[(sg * kClPts + pt) * a.K + k] = acc[i];
    }
  }
}

// One segment of a heavy cluster -> partial[seg][point][K].
template <int KCI, int SORTED>
__device__ __forceinline__ void spread_segment(const SpreadArgs& a, int64_t item, int lane) {
  constexpr int NP = ClLanes<KCI>::NP;
  constexpr int NPL = ClLanes<KCI>::NPL;
  const int64_t sg = item / a.ngrp;
  const int grp = (int)(item - sg * a.ngrp);
  const int cl = a.seg_row[sg];
  const int j0 = a.seg_j0[sg];
  const int j1 = min(j0 + a.seg, a.row_ptr[cl + 1]);
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

template <int KCI, int SORTED>
__global__ void __launch_bounds__(256) k_spread_segments(SpreadArgs a) {
  if (gate_closed(a.gate)) return;
  const int64_t item = (int64_t)blockIdx.x * 4 + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  if (item >= a.nseg * a.ngrp) return;
  spread_segment<KCI, SORTED>(a, item, threadIdx.x & 63);
}

// Heavy cluster (or cell) = sum of its segment partials in segment order; one
// thread per (heavy cluster, point, channel), channel fastest.
template <int CZ, int CY, int CX>
__global__ void __launch_bounds__(256) k_spread_combine(SpreadArgs a) {
  if (gate_closed(a.gate)) return;
  constexpr int NPT = CZ * CY * CX;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int K = a.K;
  if (i >= a.nheavy * NPT * K) return;
  const int k = (int)(i % K);
  const int64_t hp = i / K;
  const int pt = (int)(hp % NPT);
  const int64_t h = hp / NPT;
  const int s0 = a.heavy_seg_ptr[h], s1 = a.heavy_seg_ptr[h + 1];
  const float2* src = a.partial + (int64_t)pt * K + k;
  const int64_t stride = (int64_t)NPT * K;
  float2 acc = make_float2(0.f, 0.f);
  int sg = s0;
  for (; sg + 4 <= s1; sg += 4) {
    float2 v[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) v[u] = src[(int64_t)(sg + u) * stride];
#pragma unroll
    for (int u = 0; u < 4; ++u) acc = cadd(acc, v[u]);
  }
  for (; sg < s1; ++sg) acc = cadd(acc, src[(int64_t)sg * stride]);
  const int64_t cl = a.heavy_row[h];
  const int ncy = a.ncl[1], ncx = a.ncl[2];
  const int cz = (int)(cl / ((int64_t)ncy * ncx));
  const int rem = (int)(cl - (int64_t)cz * ncy * ncx);
  const int cy = rem / ncx, cx = rem - (rem / ncx) * ncx;
  const int z = cz * CZ + pt / (CY * CX);
  const int y = cy * CY + (pt / CX) % CY;
  const int x = cx * CX + pt % CX;
  const int gz = a.gg.g[0], gy = a.gg.g[1], gx = a.gg.g[2];
  if (z < gz && y < gy && x < gx && (!a.col_win || in_window(z, a.col_win[y * gx + x], gz)))
    a.grid[(((int64_t)z * gy + y) * gx + x) * K + k] = acc;
}

// ---------------------------------------------------------------------------
// Distinct grid points read by a bin's interpolation (its taps' union), for
// the algorithmic byte count of the gather kernels: mark every tap point with
// a byte store (all stores write 1, so no ordering is needed), then count.
// ---------------------------------------------------------------------------
__global__ void k_mark_taps(const TapRec* tap, int64_t M, GridGeom gg, unsigned char* mark) {
  const int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= M) return;
  const int gz = gg.g[0], gy = gg.g[1], gx = gg.g[2];
  const TapRec r = tap[m];
  const int nx = (r.n >> 8) & 15, ny = (r.n >> 4) & 15;
  const int nz = (nx == 0 || ny == 0) ? 0 : (r.n & 15);
  for (int dz = 0; dz < nz; ++dz) {
    const int pz = wrap2(r.p0[0] + dz, gz);
    for (int dy = 0; dy < ny; ++dy) {
      const int py = wrap2(r.p0[1] + dy, gy);
      for (int dx = 0; dx < nx; ++dx) mark[((int64_t)pz * gy + py) * gx + wrap2(r.p0[2] + dx, gx)] = 1;
    }
  }
}

__global__ void __launch_bounds__(256) k_count_marks(const unsigned char* mark, int64_t n, int64_t* partial) {
  int64_t c = 0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    c += mark[i];
  __shared__ int64_t sh[256];
  sh[threadIdx.x] = c;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = sh[0];
}

int count_touched(const tr_plan* p, BinData& b, hipStream_t s) {
  b.touched = 0;
  if (b.M == 0) return kOk;
  constexpr int kParts = 256;
  unsigned char* mark = nullptr;
  int64_t* part = nullptr;
  TR_HIP(hipMalloc(&mark, (size_t)p->G));
  if (hipMalloc(&part, sizeof(int64_t) * kParts) != hipSuccess) {
    (void)hipFree(mark);
    TR_FAIL(kOutOfMemory, "touched-point counters");
  }
  int64_t host[kParts];
  hipError_t e = hipMemsetAsync(mark, 0, (size_t)p->G, s);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_mark_taps, dim3(grid1d(b.M, 256)), dim3(256), 0, s, b.tap, b.M, geom(p), mark);
    hipLaunchKernelGGL(k_count_marks, dim3(kParts), dim3(256), 0, s, mark, p->G, part);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpyAsync(host, part, sizeof(host), hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  (void)hipFree(mark);
  (void)hipFree(part);
  TR_HIP(e);
  for (int i = 0; i < kParts; ++i) b.touched += host[i];
  return kOk;
}

int launch_interp(const InterpArgs& a0, hipStream_t s) {
  InterpArgs a = a0;
  a.gate = gate_current();
  const int K = a.K;
  const double per_sample = a.mode == kInterpDual ? 24.0 * K
                            : a.mode == kInterpResid ? 16.0 * K
                            : (a.mode == kInterpPlain ? 8.0 * K + 4.0 : 8.0 * K);
  ProfScope ps(kKInterp, 8.0 * K * (double)a.touched + a.M * (12.0 + per_sample), s);
  const int kci = K >= 64 ? 64 : pow2ceil(K);
  const int nci = (K + kci - 1) / kci;
  const int spw = 64 / kci;
  const unsigned blocks = grid1d(a.M, 4 * spw);
  if (a.M == 0) return kOk;
  const bool aligned16 = (((uintptr_t)a.grid | (uintptr_t)a.q | (uintptr_t)a.y) & 15) == 0;
  if (a.kgroups >= 2 && K % 2 == 0 && K >= 96 && aligned16) {  // 16-byte loads, groups of <= 128
    const int ng = (K + 127) / 128;
    int G = (K + ng - 1) / ng;
    G += G & 1;
    if (a.kgroups == 3 && a.grp && a.tapg && a.ngroups > 0)  // one wave per tap group
      hipLaunchKernelGGL(k_interp_grp, dim3(grid1d(a.ngroups, 4), ng), dim3(256), 0, s, a, G);
    else
      hipLaunchKernelGGL(k_interp_v4, dim3(grid1d(a.M, 4), ng), dim3(256), 0, s, a, G);
    TR_LAUNCH_CHECK();
    return kOk;
  }
  if (a.kgroups && kci == 64) {  // one 64-channel group per grid row, groups outermost
    hipLaunchKernelGGL((k_interp<64, 1>), dim3(blocks, nci), dim3(256), 0, s, a);
    TR_LAUNCH_CHECK();
    return kOk;
  }
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
  a.gate = gate_current();
  const int K = a.K;
  a.col_win = p->prune ? p->col_win : nullptr;
  const double G = a.col_win ? (double)p->support_points : (double)p->G;
  ProfScope ps(kKSpread, 8.0 * K * a.M + 16.0 * a.M + 8.0 * K * G, s);
  const int kci = K >= 64 ? 64 : pow2ceil(K);
  a.gg = geom(p);
  a.xcd_chunk = p->spread_xcd_chunk;
  for (int i = 0; i < 3; ++i) a.ncl[i] = p->ncl[i];
  a.NCL = p->NCL;
  a.row_ptr = b.row_ptr;
  a.pairs = b.pairs;
  a.seg = p->seg;
  a.ngrp = (K + kci - 1) / kci;
  if (a.NCL * a.ngrp >= ((int64_t)1 << 31)) TR_FAIL(kUnsupported, "too many grid clusters for the spreading");
  a.div_ngrp = make_fastdiv((unsigned)a.ngrp);
  a.div_plane = make_fastdiv((unsigned)(p->ncl[1] * p->ncl[2]));
  a.div_ncx = make_fastdiv((unsigned)p->ncl[2]);
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
#define TR_SPREAD_KCI(KERN, SRC, blocks)                                                           \
  do {                                                                                            \
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
#define TR_SPREAD_ALL(KERN, blocks)                          \
  do {                                                       \
    if (a.sorted_src) TR_SPREAD_KCI(KERN, 1, blocks);        \
    else TR_SPREAD_KCI(KERN, 0, blocks);                     \
  } while (0)
  // With support pruning and a visiting order only the ordered active clusters run, and the
  // heavy-cluster segments (the longest waves) lead them in the same launch.
  const bool ordered = a.col_win && b.corder;
  a.active = ordered ? b.corder : nullptr;
  a.nactive = ordered ? b.ncorder : 0;
  const int64_t nlight = ordered ? b.ncorder : a.NCL;
  const unsigned light_blocks = grid1d(nlight * a.ngrp, 4);
  {
    const unsigned S = (unsigned)a.xcd_chunk;
    a.div_xcd = make_fastdiv(S ? S : 1u);
    a.xcd_full = S ? light_blocks / (kNumXcd * S) * (kNumXcd * S) : 0u;
  }
  a.seg_blocks = ordered ? grid1d(b.nseg * a.ngrp, 4) : 0u;
  if (a.seg_blocks + light_blocks > 0) TR_SPREAD_ALL(k_spread, a.seg_blocks + light_blocks);
  if (b.nseg > 0) {
    if (!ordered) TR_SPREAD_ALL(k_spread_segments, grid1d(b.nseg * a.ngrp, 4));
    hipLaunchKernelGGL((k_spread_combine<kClZ, kClY, kClX>), dim3(grid1d(b.nheavy * kClPts * K, 256)), dim3(256),
                       0, s, a);
    TR_LAUNCH_CHECK();
  }
#undef TR_SPREAD_ALL
#undef TR_SPREAD_KCI
  return kOk;
}

}  // namespace tr
