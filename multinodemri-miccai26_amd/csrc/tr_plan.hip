// NUFFT plan: coordinate scaling (SigPy float32 semantics), tile sort of the
// samples, per-sample taps, the (cluster, sample) lists of the deterministic
// adjoint, tables and workspace management.

#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>

#include "tr_internal.hpp"

namespace tr {

// ---------------------------------------------------------------------------
// Plan construction kernels
// ---------------------------------------------------------------------------
__global__ void k_scale_keys(const float* coord, int64_t M, float3 s, float3 sh, int3 g, int3 nt, int tile_edge,
                             float4* kp, unsigned long long* key, int* idx) {
  // k' = fl32(fl32(k * s) + shift), as SigPy's _scale_coord on a float32 trajectory:
  // the product must be rounded before the add (no FMA), or a coordinate that
  // SigPy puts exactly on a grid line (5 taps) lands 1 ulp off it (4 taps).
#pragma clang fp contract(off)
  const int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= M) return;
  const float pz = coord[3 * m + 0] * s.x;
  const float py = coord[3 * m + 1] * s.y;
  const float px = coord[3 * m + 2] * s.z;
  const float kz = pz + sh.x;
  const float ky = py + sh.y;
  const float kx = px + sh.z;
  kp[m] = make_float4(kz, ky, kx, 0.f);
  // Key = (tile_edge^3 tile of the first tap, first tap within the tile, 5-tap flags), so samples
  // with the same taps (same first tap and tap counts, as axis_taps computes them) are adjacent
  // in the sorted order and the interpolation can gather their taps once.
  unsigned long long tile = 0;
  unsigned local = 0, five = 0;
  unsigned tq[3] = {0u, 0u, 0u};
  const float kv[3] = {kz, ky, kx};
  const int ntv[3] = {nt.x, nt.y, nt.z};
  for (int ax = 0; ax < 3; ++ax) {
    int t0 = 0, n = 0;
    if (isfinite(kv[ax]) && fabsf(kv[ax]) <= 1.0e6f) {
      const double kd = (double)kv[ax];
      t0 = (int)ceil(kd - 2.0);
      n = (int)floor(kd + 2.0) - t0 + 1;
    }
    int t = (int)floor((t0 + 4) / (double)tile_edge);
    t = t < 0 ? 0 : (t >= ntv[ax] ? ntv[ax] - 1 : t);
    const int off = t0 + 4 - t * tile_edge;  // in [0, tile_edge) unless clamped
    tq[ax] = (unsigned)t;
    local = local * (unsigned)(tile_edge + 2) + (unsigned)(off < 0 ? 0 : (off > tile_edge ? tile_edge + 1 : off));
    five = five * 2u + (n == 5 ? 1u : 0u);
  }
  // tiles in Z order (bits of z, y, x interleaved), so runs of consecutive tiles cover compact
  // blocks of the grid rather than rows
  for (int bit = 9; bit >= 0; --bit)
    tile = (tile << 3) | (((tq[0] >> bit) & 1u) << 2) | (((tq[1] >> bit) & 1u) << 1) | ((tq[2] >> bit) & 1u);
  const unsigned long long lw = (unsigned long long)(tile_edge + 2);
  key[m] = (tile * (lw * lw * lw) + local) * 8ull + five;
  idx[m] = (int)m;
}

__global__ void k_gather_kp(const float4* kp_in, const int* perm, int64_t M, float4* kp_out) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j < M) kp_out[j] = kp_in[perm[j]];
}

// Per-sample taps in sorted order; cnt[j] = number of clusters the sample's
// taps touch.
__global__ void k_tap_records(const float4* kp, int64_t M, int3 g, float beta, TapRec* tap, int* cnt) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= M) return;
  const float4 k = kp[j];
  const float kk[3] = {k.x, k.y, k.z};
  const int gg[3] = {g.x, g.y, g.z};
  const int csz[3] = {kClZ, kClY, kClX};
  TapRec r;
  int count = 1;
  int packed = 0;
  for (int ax = 0; ax < 3; ++ax) {
    float w[5];
    int pos[5];
    const int n = axis_taps(kk[ax], gg[ax], beta, ax == 2 ? 1.0f / (kWidth * kWidth * kWidth) : 1.0f, w, pos);
    r.p0[ax] = pos[0];
    packed |= n << (4 * ax);
    for (int d = 0; d < 5; ++d) r.w[5 * ax + d] = w[d];
    // taps sit at consecutive positions (mod g): count the distinct clusters
    int cl[5];
    int nc = 0;
    for (int d = 0; d < n; ++d) {
      const int c = pos[d] / csz[ax];
      bool seen = false;
      for (int e = 0; e < nc; ++e) seen |= cl[e] == c;
      if (!seen) cl[nc++] = c;
    }
    count *= nc;
  }
  r.w[15] = 0.0f;
  r.n = packed;
  tap[j] = r;
  cnt[j] = count;
}

// Extends a group's box [P, P + E) (mod g) along one axis to hold taps
// [p, p + n); false if the box would exceed 5 positions.
static bool box_extend(int P, int E, int p, int n, int g, int& nP, int& nE) {
  int o = p - P;
  if (o < 0) o += g;
  if (o + n <= 5) {
    nP = P;
    nE = std::max(E, o + n);
    return true;
  }
  int o2 = P - p;
  if (o2 < 0) o2 += g;
  if (o2 + E <= 5) {
    nP = p;
    nE = std::max(E + o2, n);
    return true;
  }
  return false;
}

// Every point of the box [P, P + E) (natural positions, mod g per axis) lies
// in the plan's k-space support (col_win): the pruned forward passes write
// exactly those points, so a box inside it never reads unwritten workspace.
static bool box_in_support(const int P[3], const int E[3], const int gdim[3], const std::vector<int2>& win) {
  for (int dy = 0; dy < E[1]; ++dy) {
    const int y = (P[1] + dy) % gdim[1];
    for (int dx = 0; dx < E[2]; ++dx) {
      const int2 w = win[(size_t)y * gdim[2] + (P[2] + dx) % gdim[2]];
      if (w.y <= 0) return false;
      for (int dz = 0; dz < E[0]; ++dz) {
        int d = (P[0] + dz) % gdim[0] - w.x;
        if (d < 0) d += gdim[0];
        if (d >= w.y) return false;
      }
    }
  }
  return true;
}

// Tap groups of the grouped interpolation: runs of at most kTapGroup
// consecutive sorted samples whose taps fit in one box of at most 5 positions
// per axis that lies inside the plan's k-space support (samples with no taps
// stay alone), greedily in sorted order; and b.tapg, each member's record
// rewritten to its group's box.  Shifting a member's weights into the box
// keeps the order of its nonzero terms, and the added terms have weight 0, so
// the gathered sums equal the per-sample ones; the support test keeps every
// box point one that the pruned forward passes write (0 * stale workspace
// could be NaN otherwise).
int build_tap_groups(tr_plan* p, BinData& b, hipStream_t s) {
  const int64_t M = b.M;
  if (M == 0 || b.grp) return kOk;
  std::vector<TapRec> h((size_t)M);
  TR_HIP(hipMemcpyAsync(h.data(), b.tap, sizeof(TapRec) * M, hipMemcpyDeviceToHost, s));
  TR_HIP(hipStreamSynchronize(s));
  const int gdim[3] = {p->g[0], p->g[1], p->g[2]};
  const std::vector<int2>& win = p->col_win_host;
  auto ntaps = [](const TapRec& r, int ax) { return (r.n >> (4 * ax)) & 15; };
  std::vector<int> starts;
  starts.reserve((size_t)M / 2 + 2);
  std::vector<TapRec> u(h);
  int64_t j = 0;
  while (j < M) {
    int P[3], E[3];
    bool empty = false;
    for (int ax = 0; ax < 3; ++ax) {
      P[ax] = h[j].p0[ax];
      E[ax] = ntaps(h[j], ax);
      empty |= E[ax] == 0;
    }
    int64_t k = j + 1;
    for (; !empty && k < M && k - j < kTapGroup; ++k) {
      int nP[3], nE[3];
      bool fits = true;
      for (int ax = 0; ax < 3 && fits; ++ax) {
        const int n = ntaps(h[k], ax);
        fits = n > 0 && box_extend(P[ax], E[ax], h[k].p0[ax], n, gdim[ax], nP[ax], nE[ax]);
      }
      if (!fits || (!win.empty() && !box_in_support(nP, nE, gdim, win))) break;
      for (int ax = 0; ax < 3; ++ax) {
        P[ax] = nP[ax];
        E[ax] = nE[ax];
      }
    }
    starts.push_back((int)j);
    if (!empty && k - j > 1) {
      for (int64_t i = j; i < k; ++i) {
        TapRec& r = u[(size_t)i];
        for (int ax = 0; ax < 3; ++ax) {
          const int n = ntaps(h[i], ax);
          int sh = h[i].p0[ax] - P[ax];
          if (sh < 0) sh += gdim[ax];
          for (int d = 0; d < 5; ++d)
            r.w[5 * ax + d] = (d >= sh && d < sh + n) ? h[i].w[5 * ax + d - sh] : 0.0f;
          r.p0[ax] = P[ax];
        }
        r.n = E[0] | (E[1] << 4) | (E[2] << 8);
      }
    }
    j = k;
  }
  starts.push_back((int)M);
  b.ngroups = (int64_t)starts.size() - 1;
  TR_HIP(hipMalloc(&b.grp, sizeof(int) * starts.size()));
  TR_HIP(hipMemcpy(b.grp, starts.data(), sizeof(int) * starts.size(), hipMemcpyHostToDevice));
  TR_HIP(hipMalloc(&b.tapg, sizeof(TapRec) * M));
  TR_HIP(hipMemcpy(b.tapg, u.data(), sizeof(TapRec) * M, hipMemcpyHostToDevice));
  return kOk;
}

// Distinct clusters touched along one axis, with the weight of every cluster
// point (a point hit by several taps, possible only when g < 5, gets their sum).
template <int CS>
__device__ int axis_clusters(const TapRec& r, int ax, int g, int cl[5], float cw[5][CS]) {
  const int n = (r.n >> (4 * ax)) & 15;
  const int p0 = r.p0[ax];
  int nc = 0;
  for (int d = 0; d < n; ++d) {
    const int c = ((p0 + d) % g) / CS;
    bool seen = false;
    for (int e = 0; e < nc; ++e) seen |= cl[e] == c;
    if (!seen) cl[nc++] = c;
  }
  for (int e = 0; e < nc; ++e)
    for (int i = 0; i < CS; ++i) {
      const int q = cl[e] * CS + i;
      float sum = 0.0f;
      for (int d = 0; d < n; ++d)
        if ((p0 + d) % g == q) sum += r.w[5 * ax + d];
      cw[e][i] = sum;
    }
  return nc;
}

__global__ void k_fill_cluster_pairs(const TapRec* tap, int64_t M, const int* off, int3 g, int3 ncl,
                                     unsigned* key, int* idx, ClusterPair* rec) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= M) return;
  const TapRec r = tap[j];
  int cz[5], cy[5], cx[5];
  float wz[5][kClZ], wy[5][kClY], wx[5][kClX];
  const int nz = axis_clusters<kClZ>(r, 0, g.x, cz, wz);
  const int ny = axis_clusters<kClY>(r, 1, g.y, cy, wy);
  const int nx = axis_clusters<kClX>(r, 2, g.z, cx, wx);
  int p = off[j];
  for (int a = 0; a < nz; ++a)
    for (int b = 0; b < ny; ++b)
      for (int c = 0; c < nx; ++c) {
        ClusterPair o;
        o.m = (int)j;
        for (int i = 0; i < kClZ; ++i) o.wz[i] = wz[a][i];
        for (int i = 0; i < kClY; ++i) o.wy[i] = wy[b][i];
        for (int i = 0; i < kClX; ++i) o.wx[i] = wx[c][i];
        o.rows = 0;
        for (int i = 0; i < kClZ; ++i)
          for (int k = 0; k < kClY; ++k)
            if (o.wz[i] * o.wy[k] != 0.0f) o.rows |= 1 << (i * kClY + k);
        key[p] = (unsigned)(((int64_t)cz[a] * ncl.y + cy[b]) * ncl.z + cx[c]);
        idx[p] = p;
        rec[p] = o;
        ++p;
      }
}

__global__ void k_gather_pairs(const ClusterPair* in, const int* idx, int64_t nnz, ClusterPair* out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nnz) out[i] = in[idx[i]];
}

// Per (y, x) column of the oversampled grid (natural order), the smallest and
// largest SigPy grid index along z that any tap of any sample touches.
__global__ void k_support_minmax(const float4* kp, int64_t M, int3 g, int* lo, int* hi) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= M) return;
  const float4 k = kp[j];
  const float kk[3] = {k.x, k.y, k.z};
  int t0[3], n[3];
  for (int ax = 0; ax < 3; ++ax) {
    if (!isfinite(kk[ax]) || fabsf(kk[ax]) > 1.0e6f) return;  // no taps (as axis_taps)
    const double kd = (double)kk[ax];
    t0[ax] = (int)ceil(kd - 2.0);
    n[ax] = (int)floor(kd + 2.0) - t0[ax] + 1;
  }
  for (int dy = 0; dy < n[1]; ++dy) {
    const int py = pmod(t0[1] + dy - g.y / 2, g.y);
    for (int dx = 0; dx < n[2]; ++dx) {
      const int col = py * g.z + pmod(t0[2] + dx - g.z / 2, g.z);
      atomicMin(&lo[col], t0[0]);
      atomicMax(&hi[col], t0[0] + n[0] - 1);
    }
  }
}

__global__ void k_fill_int(int* a, int64_t n, int v) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = v;
}

__global__ void k_row_ptr(const unsigned* key, int64_t nnz, int64_t R, int* row_ptr) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t > R) return;
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

void use_slot(tr_plan* p, int slot) {
  if (slot == p->cur) return;
  Workspace& c = p->slots[p->cur];
  c.work = p->work;
  c.work_k = p->work_k;
  c.partial = p->partial;
  c.partial_n = p->partial_n;
  const Workspace& n = p->slots[slot];
  p->work = n.work;
  p->work_k = n.work_k;
  p->partial = n.partial;
  p->partial_n = n.partial_n;
  p->cur = slot;
}

int ensure_partial(tr_plan* p, int64_t n) {
  if (p->partial_n >= n) return kOk;
  if (p->partial) {
    TR_HIP(hipDeviceSynchronize());
    TR_HIP(hipFree(p->partial));
    p->partial = nullptr;
    p->partial_n = 0;
  }
  TR_HIP(hipMalloc(&p->partial, sizeof(float2) * n));
  p->partial_n = n;
  return kOk;
}

int ensure_work(tr_plan* p, int64_t K) {
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

int build_bin(tr_plan* p, BinData& b, const float* coord, int64_t M, hipStream_t s) {
  b.M = M;
  TR_HIP(hipMalloc(&b.row_ptr, sizeof(int) * (p->NCL + 1)));
  if (M == 0) {
    TR_HIP(hipMemsetAsync(b.row_ptr, 0, sizeof(int) * (p->NCL + 1), s));
    TR_HIP(hipStreamSynchronize(s));
    return kOk;
  }
  // up to 4 x 3 x 3 clusters per sample must index an int32 pair list
  if (M > (int64_t)17000000) TR_FAIL(kUnsupported, "more than 17M samples in one bin");
  float4* kp_raw = nullptr;
  unsigned long long *key_in = nullptr, *key_out = nullptr;
  int* idx = nullptr;
  TR_HIP(hipMalloc(&kp_raw, sizeof(float4) * M));
  TR_HIP(hipMalloc(&key_in, sizeof(unsigned long long) * M));
  TR_HIP(hipMalloc(&key_out, sizeof(unsigned long long) * M));
  TR_HIP(hipMalloc(&idx, sizeof(int) * M));
  TR_HIP(hipMalloc(&b.perm, sizeof(int) * M));
  TR_HIP(hipMalloc(&b.kp, sizeof(float4) * M));
  TR_HIP(hipMalloc(&b.tap, sizeof(TapRec) * M));
  int3 nt;
  const int te = p->tile;
  nt.x = p->g[0] / te + 2;
  nt.y = p->g[1] / te + 2;
  nt.z = p->g[2] / te + 2;
  const int3 g3 = make_int3(p->g[0], p->g[1], p->g[2]);
  const int3 ncl3 = make_int3(p->ncl[0], p->ncl[1], p->ncl[2]);
  hipLaunchKernelGGL(k_scale_keys, dim3(grid1d(M, 256)), dim3(256), 0, s, coord, M,
                     make_float3(p->coord_scale[0], p->coord_scale[1], p->coord_scale[2]),
                     make_float3(p->coord_shift[0], p->coord_shift[1], p->coord_shift[2]), g3, nt, te,
                     kp_raw, key_in, idx);
  TR_LAUNCH_CHECK();
  int key_bits = 1;
  const int64_t zspan = (int64_t)pow2ceil(std::max(nt.x, std::max(nt.y, nt.z)));  // Z-order tile range per axis
  if (zspan > 1024) TR_FAIL(kUnsupported, "grid too large for the sample sort key");
  // the key's range, zspan^3 tiles x (te + 2)^3 offsets x 8 flags, in bits (a product of powers)
  int zbits = 0, lbits = 0;
  while (((int64_t)1 << zbits) < zspan) ++zbits;
  while (((int64_t)1 << lbits) < (int64_t)(te + 2) * (te + 2) * (te + 2)) ++lbits;
  key_bits = 3 * zbits + lbits + 3;
  if (key_bits > 64) TR_FAIL(kUnsupported, "grid too large for the sample sort key");
  size_t tmp_bytes = 0;
  TR_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp_bytes, key_in, key_out, idx, b.perm, (int)M, 0,
                                            key_bits, s));
  void* tmp = nullptr;
  TR_HIP(hipMalloc(&tmp, tmp_bytes));
  TR_HIP(hipcub::DeviceRadixSort::SortPairs(tmp, tmp_bytes, key_in, key_out, idx, b.perm, (int)M, 0,
                                            key_bits, s));
  hipLaunchKernelGGL(k_gather_kp, dim3(grid1d(M, 256)), dim3(256), 0, s, kp_raw, b.perm, M, b.kp);
  TR_LAUNCH_CHECK();
  // taps and cluster counts -> pair offsets
  int* cnt = nullptr;
  int* off = nullptr;
  TR_HIP(hipMalloc(&cnt, sizeof(int) * (M + 1)));
  TR_HIP(hipMalloc(&off, sizeof(int) * (M + 1)));
  TR_HIP(hipMemsetAsync(cnt + M, 0, sizeof(int), s));
  hipLaunchKernelGGL(k_tap_records, dim3(grid1d(M, 256)), dim3(256), 0, s, b.kp, M, g3, p->beta, b.tap, cnt);
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
    TR_HIP(hipMemsetAsync(b.row_ptr, 0, sizeof(int) * (p->NCL + 1), s));
    TR_HIP(hipStreamSynchronize(s));
    void* frees0[] = {kp_raw, key_in, key_out, idx, tmp, cnt, off, scan_tmp};
    for (void* f : frees0) TR_HIP(hipFree(f));
    return kOk;
  }
  unsigned *pk_in = nullptr, *pk_out = nullptr;
  int *pi_in = nullptr, *pi_out = nullptr;
  ClusterPair* rec = nullptr;
  TR_HIP(hipMalloc(&pk_in, sizeof(unsigned) * nnz));
  TR_HIP(hipMalloc(&pk_out, sizeof(unsigned) * nnz));
  TR_HIP(hipMalloc(&pi_in, sizeof(int) * nnz));
  TR_HIP(hipMalloc(&pi_out, sizeof(int) * nnz));
  TR_HIP(hipMalloc(&rec, sizeof(ClusterPair) * nnz));
  hipLaunchKernelGGL(k_fill_cluster_pairs, dim3(grid1d(M, 256)), dim3(256), 0, s, b.tap, M, off, g3, ncl3,
                     pk_in, pi_in, rec);
  TR_LAUNCH_CHECK();
  int cbits = 1;
  while (((int64_t)1 << cbits) < p->NCL) ++cbits;
  // radix sort is stable: within a cluster the pairs keep increasing sample order
  size_t tmp2_bytes = 0;
  TR_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp2_bytes, pk_in, pk_out, pi_in, pi_out, (int)nnz,
                                            0, cbits, s));
  void* tmp2 = nullptr;
  TR_HIP(hipMalloc(&tmp2, tmp2_bytes));
  TR_HIP(hipcub::DeviceRadixSort::SortPairs(tmp2, tmp2_bytes, pk_in, pk_out, pi_in, pi_out, (int)nnz, 0,
                                            cbits, s));
  TR_HIP(hipMalloc(&b.pairs, sizeof(ClusterPair) * nnz));
  hipLaunchKernelGGL(k_gather_pairs, dim3(grid1d(nnz, 256)), dim3(256), 0, s, rec, pi_out, nnz, b.pairs);
  TR_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_row_ptr, dim3(grid1d(p->NCL + 1, 256)), dim3(256), 0, s, pk_out, nnz, p->NCL,
                     b.row_ptr);
  TR_LAUNCH_CHECK();
  TR_HIP(hipStreamSynchronize(s));
  {  // heavy clusters -> fixed segments of p->seg pairs
    std::vector<int> rp(p->NCL + 1);
    TR_HIP(hipMemcpy(rp.data(), b.row_ptr, sizeof(int) * (p->NCL + 1), hipMemcpyDeviceToHost));
    std::vector<int> seg_row, seg_j0, heavy_row, heavy_ptr;
    for (int64_t t = 0; t < p->NCL; ++t) {
      const int len = rp[t + 1] - rp[t];
      if (len <= p->seg) continue;
      heavy_row.push_back((int)t);
      heavy_ptr.push_back((int)seg_row.size());
      for (int j = rp[t]; j < rp[t + 1]; j += p->seg) {
        seg_row.push_back((int)t);
        seg_j0.push_back(j);
      }
    }
    heavy_ptr.push_back((int)seg_row.size());
    b.nseg = (int64_t)seg_row.size();
    b.nheavy = (int64_t)heavy_row.size();
    if (b.nseg > 0) {
      TR_HIP(hipMalloc(&b.seg_row, sizeof(int) * b.nseg));
      TR_HIP(hipMalloc(&b.seg_j0, sizeof(int) * b.nseg));
      TR_HIP(hipMalloc(&b.heavy_row, sizeof(int) * b.nheavy));
      TR_HIP(hipMalloc(&b.heavy_seg_ptr, sizeof(int) * (b.nheavy + 1)));
      TR_HIP(hipMemcpy(b.seg_row, seg_row.data(), sizeof(int) * b.nseg, hipMemcpyHostToDevice));
      TR_HIP(hipMemcpy(b.seg_j0, seg_j0.data(), sizeof(int) * b.nseg, hipMemcpyHostToDevice));
      TR_HIP(hipMemcpy(b.heavy_row, heavy_row.data(), sizeof(int) * b.nheavy, hipMemcpyHostToDevice));
      TR_HIP(hipMemcpy(b.heavy_seg_ptr, heavy_ptr.data(), sizeof(int) * (b.nheavy + 1),
                       hipMemcpyHostToDevice));
    }
  }
  void* frees[] = {kp_raw, key_in, key_out, idx, tmp, cnt, off, scan_tmp, pk_in, pk_out, pi_in, pi_out, rec, tmp2};
  for (void* f : frees) TR_HIP(hipFree(f));
  return kOk;
}

// Sample support of all bins -> p->col_win (natural z window per column).
int build_support(tr_plan* p, hipStream_t s) {
  const int gz = p->g[0], gy = p->g[1], gx = p->g[2];
  const int64_t ncol = (int64_t)gy * gx;
  int *lo = nullptr, *hi = nullptr;
  TR_HIP(hipMalloc(&lo, sizeof(int) * ncol));
  TR_HIP(hipMalloc(&hi, sizeof(int) * ncol));
  hipLaunchKernelGGL(k_fill_int, dim3(grid1d(ncol, 256)), dim3(256), 0, s, lo, ncol, 0x7fffffff);
  hipLaunchKernelGGL(k_fill_int, dim3(grid1d(ncol, 256)), dim3(256), 0, s, hi, ncol, (int)0x80000000);
  TR_LAUNCH_CHECK();
  const int3 g3 = make_int3(gz, gy, gx);
  for (const BinData& b : p->bins) {
    if (b.M == 0) continue;
    hipLaunchKernelGGL(k_support_minmax, dim3(grid1d(b.M, 256)), dim3(256), 0, s, b.kp, b.M, g3, lo, hi);
    TR_LAUNCH_CHECK();
  }
  std::vector<int> hlo(ncol), hhi(ncol);
  TR_HIP(hipMemcpyAsync(hlo.data(), lo, sizeof(int) * ncol, hipMemcpyDeviceToHost, s));
  TR_HIP(hipMemcpyAsync(hhi.data(), hi, sizeof(int) * ncol, hipMemcpyDeviceToHost, s));
  TR_HIP(hipStreamSynchronize(s));
  TR_HIP(hipFree(lo));
  TR_HIP(hipFree(hi));
  std::vector<int2> win(ncol);
  for (int64_t c = 0; c < ncol; ++c) {
    if (hhi[c] < hlo[c]) {
      win[c] = make_int2(0, 0);
      continue;
    }
    const int64_t len = (int64_t)hhi[c] - hlo[c] + 1;
    win[c] = len >= gz ? make_int2(0, gz) : make_int2(pmod(hlo[c] - gz / 2, gz), (int)len);
  }
  // The y passes take one window per x: the shortest circular window holding
  // every live column (y, x) (the complement of the longest run of dead ones).
  // A dead column inside it gets a one-point z window, which the spreading
  // fills with zero, so the inverse y pass only reads written points.
  std::vector<int2> yw(gx);
  for (int x = 0; x < gx; ++x) {
    int nlive = 0;
    for (int y = 0; y < gy; ++y) nlive += win[(int64_t)y * gx + x].y > 0;
    if (nlive == 0) {
      yw[x] = make_int2(0, 0);
      continue;
    }
    if (nlive == gy) {
      yw[x] = make_int2(0, gy);
      continue;
    }
    int best_len = 0, best_end = 0, run = 0;
    for (int i = 0; i < 2 * gy; ++i) {  // longest circular run of dead columns
      if (win[(int64_t)(i % gy) * gx + x].y == 0) {
        if (++run > best_len && run <= gy) {
          best_len = run;
          best_end = i % gy;
        }
      } else {
        run = 0;
      }
    }
    yw[x] = make_int2(pmod(best_end + 1, gy), gy - best_len);
    for (int i = 0; i < yw[x].y; ++i) {
      int2& w = win[(int64_t)((yw[x].x + i) % gy) * gx + x];
      if (w.y == 0) w = make_int2(0, 1);
    }
  }
  p->live_cols = 0;
  p->support_points = 0;
  for (int64_t c = 0; c < ncol; ++c) {
    p->live_cols += win[c].y > 0;
    p->support_points += win[c].y;
  }
  TR_HIP(hipMalloc(&p->ywin, sizeof(int2) * gx));
  TR_HIP(hipMemcpy(p->ywin, yw.data(), sizeof(int2) * gx, hipMemcpyHostToDevice));
  TR_HIP(hipMalloc(&p->col_win, sizeof(int2) * ncol));
  TR_HIP(hipMemcpy(p->col_win, win.data(), sizeof(int2) * ncol, hipMemcpyHostToDevice));
  p->col_win_host = std::move(win);
  return kOk;
}

// Visiting order of the cluster kernel: the light clusters that hold at least one support
// point (heavy ones are written by the segment combine, clusters without support points by
// nobody), grouped in bricks of bz x by x bx clusters (code = bz | by << 4 | bx << 8; 1 x 1 x 1 =
// index order), bricks and the clusters inside a brick in index order.  A sample's taps touch
// 2 x 2 x 3 neighbouring clusters, so a brick keeps most of them in one XCD's run of blocks.
int build_cluster_order(tr_plan* p, BinData& b, int code) {
  (void)hipFree(b.corder);
  b.corder = nullptr;
  b.ncorder = 0;
  const bool lpt = code == kSpreadOrderLongestFirst;
  const bool brick_lpt = (code & kSpreadOrderBrickLongestFirst) != 0;
  const int bz = lpt ? 1 : std::max(1, code & 15), by = lpt ? 1 : std::max(1, (code >> 4) & 15),
            bx = lpt ? 1 : std::max(1, (code >> 8) & 255);
  std::vector<int64_t> brick_end;  // end of each brick's run in `order`
  std::vector<int> rp(p->NCL + 1);
  TR_HIP(hipMemcpy(rp.data(), b.row_ptr, sizeof(int) * (p->NCL + 1), hipMemcpyDeviceToHost));
  const int ncz = p->ncl[0], ncy = p->ncl[1], ncx = p->ncl[2];
  const int gz = p->g[0], gy = p->g[1], gx = p->g[2];
  const std::vector<int2>& win = p->col_win_host;
  std::vector<int> order;
  order.reserve((size_t)p->NCL);
  for (int z0 = 0; z0 < ncz; z0 += bz)
    for (int y0 = 0; y0 < ncy; y0 += by)
      for (int x0 = 0; x0 < ncx; x0 += bx) {
        for (int cz = z0; cz < std::min(z0 + bz, ncz); ++cz)
          for (int cy = y0; cy < std::min(y0 + by, ncy); ++cy)
            for (int cx = x0; cx < std::min(x0 + bx, ncx); ++cx) {
              const int64_t t = ((int64_t)cz * ncy + cy) * ncx + cx;
              const int len = rp[t + 1] - rp[t];
              if (len > p->seg) continue;
              bool live = len > 0 || win.empty();
              for (int iy = 0; iy < kClY && !live; ++iy)
                for (int ix = 0; ix < kClX && !live; ++ix) {
                  const int y = cy * kClY + iy, x = cx * kClX + ix;
                  if (y >= gy || x >= gx) continue;
                  const int2 w = win[(size_t)y * gx + x];
                  for (int iz = 0; iz < kClZ && !live; ++iz) {
                    const int zz = cz * kClZ + iz;
                    if (zz >= gz || w.y <= 0) continue;
                    int d = zz - w.x;
                    if (d < 0) d += gz;
                    live = d < w.y;
                  }
                }
              if (live) order.push_back((int)t);
            }
          if (brick_lpt && (brick_end.empty() ? !order.empty() : brick_end.back() < (int64_t)order.size()))
            brick_end.push_back((int64_t)order.size());
        }
  if (code == kSpreadOrderLongestFirst)  // longest pair lists first (index order among equal lengths)
    std::stable_sort(order.begin(), order.end(),
                     [&rp](int x, int y) { return rp[x + 1] - rp[x] > rp[y + 1] - rp[y]; });
  if (brick_lpt) {
    // whole bricks, the brick holding the longest list first (brick order among equal maxima);
    // inside a brick, index order, so clusters that share samples run close together
    std::vector<std::pair<int, int64_t>> key;  // (longest list, brick index)
    int64_t s0 = 0;
    for (size_t i = 0; i < brick_end.size(); ++i) {
      int m = 0;
      for (int64_t j = s0; j < brick_end[i]; ++j) m = std::max(m, rp[order[j] + 1] - rp[order[j]]);
      key.push_back({m, (int64_t)i});
      s0 = brick_end[i];
    }
    std::stable_sort(key.begin(), key.end(), [](const std::pair<int, int64_t>& a,
                                                const std::pair<int, int64_t>& b) { return a.first > b.first; });
    std::vector<int> sorted;
    sorted.reserve(order.size());
    for (const auto& k : key) {
      const int64_t i = k.second;
      for (int64_t j = i == 0 ? 0 : brick_end[i - 1]; j < brick_end[i]; ++j) sorted.push_back(order[j]);
    }
    order.swap(sorted);
  }
  b.ncorder = (int64_t)order.size();
  if (!order.empty()) {
    TR_HIP(hipMalloc(&b.corder, sizeof(int) * order.size()));
    TR_HIP(hipMemcpy(b.corder, order.data(), sizeof(int) * order.size(), hipMemcpyHostToDevice));
  }
  return kOk;
}

void free_plan(tr_plan* p) {
  if (!p) return;
  for (int i = 0; i < 3; ++i) {
    (void)hipFree(p->chirp[i]);
    (void)hipFree(p->hhat[i]);
  }
  for (auto& b : p->bins) {
    (void)hipFree(b.perm);
    (void)hipFree(b.kp);
    (void)hipFree(b.tap);
    (void)hipFree(b.grp);
    (void)hipFree(b.tapg);
    (void)hipFree(b.corder);
    (void)hipFree(b.row_ptr);
    (void)hipFree(b.pairs);
    (void)hipFree(b.seg_row);
    (void)hipFree(b.seg_j0);
    (void)hipFree(b.heavy_row);
    (void)hipFree(b.heavy_seg_ptr);
  }
  for (int i = 0; i < 3; ++i) {
    (void)hipFree(p->tw[i]);
    (void)hipFree(p->apod[i]);
  }
  use_slot(p, 0);
  (void)hipFree(p->work);
  (void)hipFree(p->partial);
  for (int i = 1; i < kMaxSlots; ++i) {
    (void)hipFree(p->slots[i].work);
    (void)hipFree(p->slots[i].partial);
  }
  (void)hipFree(p->col_win);
  (void)hipFree(p->ywin);
  delete p;
}

int check_plan_bin(const tr_plan* p, int bin) {
  if (!p) TR_FAIL(kInvalidArgument, "null plan");
  if (bin < 0 || bin >= (int)p->bins.size())
    TR_FAIL(kInvalidArgument, "bin_index " + std::to_string(bin) + " is outside [0, " +
                                  std::to_string(p->bins.size()) + ").");
  return kOk;
}

}  // namespace tr

using namespace tr;

// Bluestein tables of axis i (length L, transform length M): chirp[n] = w_n =
// exp(-i pi n^2 / L) for n < L, with n^2 reduced mod 2L in integers so the angle
// is exact, and hhat = FFT_M(b) / M for b_m = conj(w_m), b_{M-m} = conj(w_m),
// evaluated as a direct float64 DFT with integer-reduced angles.
static int build_bluestein(tr_plan* p, int i) {
  const int L = p->ax[i].L, M = p->ax[i].M;
  std::vector<double> wr(L), wi(L);
  for (int n = 0; n < L; ++n) {
    const int64_t q = ((int64_t)n * n) % (2 * (int64_t)L);
    const double a = -M_PI * (double)q / L;
    wr[n] = std::cos(a);
    wi[n] = std::sin(a);
  }
  std::vector<double> br(M, 0.0), bi(M, 0.0);
  for (int m = 0; m < L; ++m) {
    br[m] = wr[m];
    bi[m] = -wi[m];
    if (m > 0) {
      br[M - m] = wr[m];
      bi[M - m] = -wi[m];
    }
  }
  std::vector<double> cr(M), ci(M);
  for (int t = 0; t < M; ++t) {
    cr[t] = std::cos(2.0 * M_PI * (double)t / M);
    ci[t] = -std::sin(2.0 * M_PI * (double)t / M);
  }
  std::vector<float2> chirp(L), hhat(M);
  for (int n = 0; n < L; ++n) chirp[n] = make_float2((float)wr[n], (float)wi[n]);
  for (int k = 0; k < M; ++k) {
    double sr = 0.0, si = 0.0;
    for (int m = 0; m < M; ++m) {
      if (br[m] == 0.0 && bi[m] == 0.0) continue;
      const int t = (int)(((int64_t)m * k) % M);
      sr += br[m] * cr[t] - bi[m] * ci[t];
      si += br[m] * ci[t] + bi[m] * cr[t];
    }
    hhat[k] = make_float2((float)(sr / M), (float)(si / M));
  }
  TR_HIP(hipMalloc(&p->chirp[i], sizeof(float2) * L));
  TR_HIP(hipMalloc(&p->hhat[i], sizeof(float2) * M));
  TR_HIP(hipMemcpy(p->chirp[i], chirp.data(), sizeof(float2) * L, hipMemcpyHostToDevice));
  TR_HIP(hipMemcpy(p->hhat[i], hhat.data(), sizeof(float2) * M, hipMemcpyHostToDevice));
  p->ax[i].chirp = p->chirp[i];
  p->ax[i].hhat = p->hhat[i];
  return kOk;
}

extern "C" {

int tr_plan_create(tr_plan** out, int nz, int ny, int nx, int nbins, const float* const* coords,
                   const int64_t* samples_per_bin, void* stream) {
  return tr_plan_create_geom(out, nz, ny, nx, nbins, coords, samples_per_bin, kTile, kSpreadSeg, stream);
}

int tr_plan_create_geom(tr_plan** out, int nz, int ny, int nx, int nbins, const float* const* coords,
                        const int64_t* samples_per_bin, int sort_tile, int spread_segment, void* stream) {
  if (!out) TR_FAIL(kInvalidArgument, "null output pointer");
  *out = nullptr;
  if (nz < 1 || ny < 1 || nx < 1) TR_FAIL(kInvalidArgument, "img_shape must be three positive values");
  if (nbins < 1) TR_FAIL(kInvalidArgument, "At least one trajectory bin is required.");
  if (sort_tile < kMinTile || sort_tile > kMaxTile)
    TR_FAIL(kInvalidArgument, "sort tile edge outside [" + std::to_string(kMinTile) + ", " +
                                  std::to_string(kMaxTile) + "]");
  if (spread_segment < kMinSpreadSeg || spread_segment > kMaxSpreadSeg)
    TR_FAIL(kInvalidArgument, "spreading segment length outside [" + std::to_string(kMinSpreadSeg) + ", " +
                                  std::to_string(kMaxSpreadSeg) + "]");
  hipStream_t s = (hipStream_t)stream;
  tr_plan* p = new tr_plan();
  p->tile = sort_tile;
  p->seg = spread_segment;
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
      TR_FAIL(kUnsupported, "oversampled grid length " + std::to_string(p->g[i]) + " is not supported (at most " +
                                std::to_string(kMaxTransform) + " points with prime factors up to 13, or " +
                                std::to_string(kMaxTransform / 2) + " otherwise)");
    }
  }
  p->G = (int64_t)p->g[0] * p->g[1] * p->g[2];
  p->ncl[0] = (p->g[0] + kClZ - 1) / kClZ;
  p->ncl[1] = (p->g[1] + kClY - 1) / kClY;
  p->ncl[2] = (p->g[2] + kClX - 1) / kClX;
  p->NCL = (int64_t)p->ncl[0] * p->ncl[1] * p->ncl[2];
  for (int i = 0; i < 3; ++i) {
    const int L = p->ax[i].M;  // the transform's twiddles
    std::vector<float2> tw(L);
    for (int t = 0; t < L; ++t) {
      const double a = -2.0 * M_PI * (double)t / L;
      tw[t] = make_float2((float)std::cos(a), (float)std::sin(a));
    }
    if (p->ax[i].bluestein && build_bluestein(p, i)) {
      std::string msg = last_error();
      free_plan(p);
      set_error(msg);
      return kOutOfMemory;
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
    int r = build_bin(p, p->bins[b], coords[b], samples_per_bin[b], s);
    if (!r) r = count_touched(p, p->bins[b], s);
    if (r) {
      std::string msg = last_error();
      free_plan(p);
      set_error(msg);
      return r;
    }
  }
  int r = build_support(p, s);
  for (int b = 0; b < nbins && !r; ++b) r = build_tap_groups(p, p->bins[b], s);
  for (int b = 0; b < nbins && !r; ++b)
    if (p->bins[b].M > 0) r = build_cluster_order(p, p->bins[b], kSpreadOrderLongestFirst);
  if (!r) p->spread_order = kSpreadOrderLongestFirst;
  if (r) {
    std::string msg = last_error();
    free_plan(p);
    set_error(msg);
    return r;
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
  info[10] = (bin >= 0 && bin < (int)p->bins.size()) ? p->bins[bin].touched : -1;
  info[11] = p->support_points;
  return kOk;
}

int tr_plan_release_workspace(tr_plan* p) {
  if (!p) TR_FAIL(kInvalidArgument, "null plan");
  TR_HIP(hipDeviceSynchronize());
  use_slot(p, 0);
  TR_HIP(hipFree(p->work));
  p->work = nullptr;
  p->work_k = 0;
  for (int i = 1; i < kMaxSlots; ++i) {
    TR_HIP(hipFree(p->slots[i].work));
    p->slots[i].work = nullptr;
    p->slots[i].work_k = 0;
  }
  return kOk;
}

int tr_plan_set_param(tr_plan* p, int param, int64_t value) {
  if (!p) TR_FAIL(kInvalidArgument, "null plan");
  switch (param) {
    case TR_PARAM_INTERP_XCD_CHUNK:
    case TR_PARAM_SPREAD_XCD_CHUNK:
      if (value < 0 || value > 4096) TR_FAIL(kInvalidArgument, "parameter value outside [0, 4096]");
      (param == TR_PARAM_INTERP_XCD_CHUNK ? p->interp_xcd_chunk : p->spread_xcd_chunk) = (int)value;
      return kOk;
    case TR_PARAM_PRUNE:
      if (value != 0 && value != 1) TR_FAIL(kInvalidArgument, "prune is 0 or 1");
      p->prune = (int)value;
      return kOk;
    case TR_PARAM_INTERP_KGROUPS:
      if (value < 0 || value > 3) TR_FAIL(kInvalidArgument, "interp channel groups is 0 .. 3");
      p->interp_kgroups = (int)value;
      return kOk;
    case TR_PARAM_WORKSPACE_SLOTS:
      if (value < 1 || value > kMaxSlots) TR_FAIL(kInvalidArgument, "workspace slots outside [1, 8]");
      p->nslots = (int)value;
      return kOk;
    case TR_PARAM_LINES_XCD_LINES:
      if (value < 0 || value > 64) TR_FAIL(kInvalidArgument, "lines per XCD run outside [0, 64]");
      p->lines_xcd_lines = (int)value;
      return kOk;
    case TR_PARAM_SPREAD_ORDER: {
      if (value < 0 || value > (kSpreadOrderBrickLongestFirst | 0xffff) ||
          (value > 0xffff && (value & 0xffff) == 0xffff))
        TR_FAIL(kInvalidArgument, "spreading order code outside [0, 65535] and 65536 + [1, 65534]");
      TR_HIP(hipDeviceSynchronize());
      for (BinData& b : p->bins) {
        if (value == 0 || b.M == 0) {
          (void)hipFree(b.corder);
          b.corder = nullptr;
          b.ncorder = 0;
        } else if (int r = build_cluster_order(p, b, (int)value)) {
          return r;
        }
      }
      p->spread_order = (int)value;
      return kOk;
    }
    case TR_PARAM_SUPPORT_POINTS:
    case TR_PARAM_LIVE_COLUMNS:
    case TR_PARAM_SORT_TILE:
    case TR_PARAM_SPREAD_SEGMENT:
      TR_FAIL(kInvalidArgument, "read-only plan parameter " + std::to_string(param));
    default: TR_FAIL(kInvalidArgument, "unknown plan parameter " + std::to_string(param));
  }
}

int tr_plan_get_param(const tr_plan* p, int param, int64_t* value) {
  if (!p || !value) TR_FAIL(kInvalidArgument, "null argument");
  switch (param) {
    case TR_PARAM_INTERP_XCD_CHUNK: *value = p->interp_xcd_chunk; return kOk;
    case TR_PARAM_SPREAD_XCD_CHUNK: *value = p->spread_xcd_chunk; return kOk;
    case TR_PARAM_PRUNE: *value = p->prune; return kOk;
    case TR_PARAM_INTERP_KGROUPS: *value = p->interp_kgroups; return kOk;
    case TR_PARAM_WORKSPACE_SLOTS: *value = p->nslots; return kOk;
    case TR_PARAM_LINES_XCD_LINES: *value = p->lines_xcd_lines; return kOk;
    case TR_PARAM_SUPPORT_POINTS: *value = p->support_points; return kOk;
    case TR_PARAM_LIVE_COLUMNS: *value = p->live_cols; return kOk;
    case TR_PARAM_SPREAD_ORDER: *value = p->spread_order; return kOk;
    case TR_PARAM_SORT_TILE: *value = p->tile; return kOk;
    case TR_PARAM_SPREAD_SEGMENT: *value = p->seg; return kOk;
    default: TR_FAIL(kInvalidArgument, "unknown plan parameter " + std::to_string(param));
  }
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

int tr_plan_scatter_weights(const tr_plan* p, int bin, const float* src, float* dst, void* stream) {
  if (int r = check_plan_bin(p, bin)) return r;
  const BinData& b = p->bins[bin];
  if (b.M == 0) return kOk;
  hipLaunchKernelGGL(k_scatter_samples<float>, dim3(grid1d(b.M, 256)), dim3(256), 0,
                     (hipStream_t)stream, src, b.perm, b.M, 1, dst);
  TR_LAUNCH_CHECK();
  return kOk;
}

}  // extern "C"
