// Internal declarations shared by the toporecon HIP sources (not part of the C ABI).
//
// Grid layout in HBM for one motion bin: [gz][gy][gx][K] complex64 with
// K = E * C channels (k = e * C + c), in FFT-natural order: image voxel m maps
// to position (m - n//2) mod g and SigPy grid index t to (t - g//2) mod g, which
// folds the centred resize / ifftshift / fftshift of sp.nufft into indexing.
// Samples of a bin are kept sorted by 8^3 grid tile ("sorted order");
// per-sample data of the fused solver path lives as [M_b][K].
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "toporecon_hip.h"
#include "tr_common.hpp"
#include "tr_fft.hpp"

namespace tr {

constexpr int kWidth = 4;
constexpr int kTile = 8;  // default sample sort tile edge (grid points)
constexpr int kMinTile = 2, kMaxTile = 32;
constexpr int kMaxKc = 16;  // channels per FFT workgroup
#ifndef TR_SPREAD_SEG
#define TR_SPREAD_SEG 1024
#endif
constexpr int kSpreadSeg = TR_SPREAD_SEG;  // default pairs per heavy-cluster segment
constexpr int kMinSpreadSeg = 64, kMaxSpreadSeg = 1 << 16;
constexpr int kSpreadOrderLongestFirst = 0xffff;  // TR_PARAM_SPREAD_ORDER code of the default order
constexpr int kSpreadOrderBrickLongestFirst = 0x10000;  // + brick code: bricks, longest list's brick first
#ifndef TR_SPREAD_UNROLL
#define TR_SPREAD_UNROLL 4
#endif
constexpr int kSpreadUnroll = TR_SPREAD_UNROLL;  // pairs whose sample loads are in flight together

// Per-sample taps (sorted order): natural positions of the first tap per axis,
// tap counts (4 or 5) and float32 Kaiser-Bessel weights (x weights already
// divided by width^3), evaluated once at plan time.
struct __align__(16) TapRec {
  int p0[3];   // (t0 - g//2) mod g per axis (z, y, x)
  int n;       // nz | ny << 4 | nx << 8
  float w[16];  // z: 0..4, y: 5..9, x: 10..14
};

// The adjoint gathers over clusters of kClZ x kClY x kClX grid points: a wave
// keeps the cluster's sums for 64 channels in registers and walks the sorted
// samples touching the cluster once.
constexpr int kClZ = 4, kClY = 4, kClX = 2;
constexpr int kClPts = kClZ * kClY * kClX;

// One (cluster, sample) pair: the sample's sorted index and its weights on the
// cluster's points per axis (zero off the taps or past the grid edge).
struct __align__(16) ClusterPair {
  int m;
  float wz[kClZ];
  float wy[kClY];
  float wx[kClX];  // includes 1 / width^3
  int rows;        // bit (z * kClY + y): the (z, y) point row has a nonzero weight
};

struct BinData {
  int64_t M = 0;
  int64_t nnz = 0;          // (cluster, sample) pairs
  int* perm = nullptr;      // sorted position -> original sample index
  float4* kp = nullptr;     // sorted scaled coordinates (kz', ky', kx', 0)
  TapRec* tap = nullptr;    // sorted per-sample taps
  int* row_ptr = nullptr;   // NCL + 1, over clusters
  ClusterPair* pairs = nullptr;  // nnz, sample order within a cluster
  // clusters with more than kSpreadSeg samples, cut into segments of kSpreadSeg
  int64_t nseg = 0;
  int64_t nheavy = 0;
  int* seg_row = nullptr;        // cluster of each segment
  int* seg_j0 = nullptr;         // first pair of each segment
  int* heavy_row = nullptr;      // heavy cluster ids
  int* heavy_seg_ptr = nullptr;  // nheavy + 1
  int64_t touched = 0;           // distinct grid points under the bin's taps
  // groups of at most kTapGroup consecutive sorted samples whose taps fit in
  // one box of at most 5 positions per axis, gathered once per group; tapg
  // holds every member's taps in its group's box (the box origin and extent,
  // the member's weights shifted into it, zero where it has no tap)
  int64_t ngroups = 0;
  int* grp = nullptr;            // ngroups + 1 offsets into the sorted samples
  TapRec* tapg = nullptr;
  // light clusters holding at least one support point, in the visiting order of the cluster
  // kernel (TR_PARAM_SPREAD_ORDER); empty = every cluster in index order
  int64_t ncorder = 0;
  int* corder = nullptr;
};

}  // namespace tr

namespace tr {
constexpr int kMaxSlots = 8;
struct Workspace {
  float2* work = nullptr;
  int64_t work_k = 0;
  float2* partial = nullptr;
  int64_t partial_n = 0;
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
  float2* chirp[3] = {nullptr, nullptr, nullptr};  // Bluestein axes only
  float2* hhat[3] = {nullptr, nullptr, nullptr};
  float* apod[3] = {nullptr, nullptr, nullptr};
  std::vector<tr::BinData> bins;
  float2* work = nullptr;  // grid workspace, G * work_k complex
  int64_t work_k = 0;
  float2* partial = nullptr;  // heavy-cluster segment partial sums
  int64_t partial_n = 0;
  // work / work_k / partial / partial_n above belong to workspace slot `cur`; a call for
  // bin b uses slot b % nslots, so calls for bins in different slots may run concurrently
  // on different streams
  tr::Workspace slots[tr::kMaxSlots];
  int nslots = 1;
  int cur = 0;
  int ncl[3];                 // clusters per axis
  int64_t NCL;
  // launch parameters (tr_plan_set_param): runs of consecutive blocks kept on
  // one XCD by the gather kernels (0 = hardware round-robin order)
  int interp_xcd_chunk = 64;
  int spread_xcd_chunk = 0;
  int lines_xcd_lines = 4;  // y / z passes: whole lines per XCD run (0 = round-robin)
  int interp_kgroups = 0;  // 1: one 64-channel group per launch row (group-major order)
  int spread_order = 0;    // 0: all clusters in index order; else the code of the active-cluster order
  // sample geometry, fixed when the plan is built: the sort tile edge (samples are ordered by
  // tile, then by first tap inside the tile) and the heavy-cluster segment length (pair lists
  // longer than it are cut into segments of that many pairs, reduced in segment order)
  int tile = tr::kTile;
  int seg = tr::kSpreadSeg;
  // k-space support of all bins' taps: per (y, x) column of the oversampled
  // grid (natural order) the window (start, len) of z positions any tap
  // touches; len 0 = empty column.  Grid points outside it are never read by
  // the interpolation and never receive a spread contribution, so the line
  // passes skip them.
  int2* col_win = nullptr;
  std::vector<int2> col_win_host;  // host copy of col_win (tap groups are checked against it)
  int2* ywin = nullptr;  // per x: the y window covering every live column (y, x)
  int prune = 1;
  int64_t live_cols = 0;
  int64_t support_points = 0;
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
  int K;  // grid channel count (stride of a grid point's channel record)
  int E;  // echoes of this launch: img holds E images, grid channels e * C + c
  int eg;  // echoes whose image rows the block stages in LDS together (1 .. E)
  float2* grid;
  const int* gate = nullptr;  // batched solves: skip the launch while *gate != 0
};

// Division of unsigned values below 2^31 by a divisor fixed at launch: one
// high multiply, an add and a shift (Granlund and Montgomery, "Division by
// invariant integers using multiplication", PLDI 1994, fig. 4.1), which for
// uniform operands stays on the scalar unit.
struct FastDiv {
  unsigned d, m, s;
  __host__ __device__ unsigned div(unsigned x) const {
#if defined(__HIP_DEVICE_COMPILE__)
    const unsigned hi = __umulhi(x, m);
#else
    const unsigned hi = (unsigned)(((unsigned long long)x * m) >> 32);
#endif
    return (unsigned)(((unsigned long long)hi + x) >> s);
  }
};

inline FastDiv make_fastdiv(unsigned d) {  // d >= 1
  FastDiv f;
  f.d = d;
  f.s = 0;
  while ((1ull << f.s) < d) ++f.s;
  f.m = (unsigned)((((1ull << 32) * ((1ull << f.s) - d)) / d) + 1);
  return f;
}

struct LinesArgs {
  FftAxis ax;
  unsigned xcd_chunk;  // workgroups per XCD run (whole lines of channel chunks), 0 = round-robin
  // derived by launch_lines: block-index decoding without runtime divisions,
  // and whether every position of a line lies within 2^31 bytes of the line's
  // first position plus 15 position strides (32-bit buffer offsets)
  FastDiv div_nchunk, div_inner, div_xcd;
  unsigned xcd_full;  // blocks below this are remapped in XCD runs, the tail keeps its order
  int off32;
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
  // support pruning (null = off): line_win[o * win_so + in * win_si] is the
  // line's window along the FFT axis; only positions inside it are stored
  // (forward) or loaded (inverse; others are zero), and a line with an empty
  // window is skipped.
  const int2* line_win;
  int win_so, win_si;
  int zero_empty;  // an empty line still stores zeros (its outputs are read later)
  const int* gate = nullptr;  // batched solves: skip the launch while *gate != 0
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
  int K;  // channel stride of the grid
  int E;  // echoes of this launch (at most kAdjXEchoes): grid channels e * C + c, out (E, Z, Y, X)
  float2* out;
  int accumulate;
  const int* gate = nullptr;  // batched solves: skip the launch while *gate != 0
};

constexpr int kAdjXEchoes = 8;  // echo accumulators of the inverse x pass

enum InterpMode : int { kInterpPlain = 0, kInterpDual = 1, kInterpSim = 2, kInterpResid = 3 };

constexpr int kTapGroup = 4;  // samples sharing one tap gather in the grouped interpolation
// y rows whose 4 x taps the grouped interpolation loads together (1, 2 or 4): one row keeps the kernel
// at 51 VGPRs, which hides the gathers' latency better than 16 loads in flight per wave at 94 VGPRs
// (5 waves per SIMD): 2.80 -> 2.71 ms per launch at 192 channels
#ifndef TR_INTERP_ROWS
#define TR_INTERP_ROWS 1
#endif
constexpr int kInterpRows = TR_INTERP_ROWS;
static_assert(kInterpRows == 1 || kInterpRows == 2 || kInterpRows == 4, "y rows per tap batch: 1, 2 or 4");

struct InterpArgs {
  GridGeom gg;
  const int* grp;   // tap groups (grouped mode)
  const TapRec* tapg;  // per-sample taps in the group's box (grouped mode)
  int64_t ngroups;
  int64_t touched;  // distinct grid points read (algorithmic byte count only)
  int xcd_chunk;
  int kgroups;
  const TapRec* tap;
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
  const int* gate = nullptr;  // batched solves: skip the launch while *gate != 0
};

struct SpreadArgs {
  GridGeom gg;
  int xcd_chunk;
  int ncl[3];
  int64_t NCL;
  int K;
  int64_t M;
  const int* row_ptr;
  const ClusterPair* pairs;
  int sorted_src;  // 1: q (sorted, [M][K]); 0: ksp (original order, [K][M])
  // sorted source: q[m * K + k] (optionally times dcf_sorted[m])
  const float2* q;
  const float* dcf_sorted;
  // plain source: ksp[k * M + perm[m]] (optionally times dcf_orig[perm[m]])
  const float2* ksp;
  const float* dcf_orig;
  const int* perm;
  float2* grid;
  // heavy-cluster segments
  int64_t nseg;
  const int* seg_row;
  const int* seg_j0;
  int64_t nheavy;
  const int* heavy_row;
  const int* heavy_seg_ptr;
  float2* partial;  // nseg * kClPts * K
  int ngrp;         // channel groups per cluster
  // derived by launch_spread for the light-cluster kernel: block and cluster decoding with
  // precomputed multiplicative inverses (uniform values stay on the scalar unit)
  FastDiv div_ngrp, div_plane, div_ncx, div_xcd;
  unsigned xcd_full;  // blocks below this are remapped in XCD runs of xcd_chunk
  const int2* col_win;  // support windows (null = write every point)
  // visiting order of the light clusters (null: every cluster in index order)
  const int* active;
  int64_t nactive;
  unsigned seg_blocks;  // leading workgroups of the cluster kernel that run the heavy segments
  int seg;              // pairs per heavy-cluster segment (the plan's)
  const int* gate = nullptr;  // batched solves: skip the launch while *gate != 0
};

// Raw buffer accesses with a lane offset at or past the record count read
// zero and write nothing, so masked lanes need no per-lane branch and make no
// memory request; a wave with no live lane skips the instruction altogether
// (one scalar branch), which matters where support pruning masks whole waves.
// The base is a uniform pointer (scalar registers); each lane adds one 32-bit
// byte offset, and a per-instruction scalar offset can step between rows.
constexpr int kBufRecords = 0x7fffffff;
constexpr unsigned kBufOob = 0x80000000u;
constexpr int kBufWord3 = 0x00020000;  // gfx9 raw buffer: 32-bit data format, no swizzle

__device__ __forceinline__ __amdgpu_buffer_rsrc_t buf_rsrc(char* p) {
  return __builtin_amdgcn_make_buffer_rsrc(p, 0, kBufRecords, kBufWord3);
}

// Position p of an axis of length g lies in the window (start, len) of that axis.
__device__ __forceinline__ bool in_window(int p, int2 w, int g) {
  int d = p - w.x;
  d = d < 0 ? d + g : d;
  return d < w.y;
}

// Loads of a line tile are staged in registers (all of a thread's loads are
// issued before the first LDS write) so that every wave keeps up to
// kMaxFftLen * KC / kBlock loads in flight.
template <int KC, int LC>
struct Tile {
  static constexpr int ld = lds_ld(KC);
  static constexpr int W = kBlock / KC;  // workers per channel
  static constexpr int PER = ((LC > 0 ? LC : max_line(KC)) + W - 1) / W;  // positions per thread
};

// ---------------------------------------------------------------------------
// Host helpers
// ---------------------------------------------------------------------------
inline unsigned grid1d(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }


inline int pow2ceil(int v) {
  int p = 1;
  while (p < v) p <<= 1;
  return p;
}

// Channel chunk of the FFT passes: at most 16 channels per workgroup keeps the
// Stockham stages under ~130 VGPRs and the LDS tile at L * 17 * 8 bytes.
// ... and at most max_line(kc) >= M points per line, so a chunk's tile stays within
// kMaxLine * 8 bytes (about 34 KiB with the twiddles) whatever the length.
inline int fft_kc(int K, int M) {
  int kc = K >= kMaxKc ? kMaxKc : pow2ceil(K);
  while (kc > 1 && M > max_line(kc)) kc >>= 1;
  return kc;
}

inline size_t fft_smem(int M, int kc) { return sizeof(float2) * (size_t)M * (1 + lds_ld(kc)); }


inline GridGeom geom(const tr_plan* p) {
  GridGeom g;
  for (int i = 0; i < 3; ++i) {
    g.n[i] = p->n[i];
    g.g[i] = p->g[i];
  }
  return g;
}


int ensure_partial(tr_plan* p, int64_t n);
int ensure_work(tr_plan* p, int64_t K);
int build_support(tr_plan* p, hipStream_t s);
int build_tap_groups(tr_plan* p, BinData& b, hipStream_t s);
int build_cluster_order(tr_plan* p, BinData& b, int code);
void use_slot(tr_plan* p, int slot);
int build_bin(tr_plan* p, BinData& b, const float* coord, int64_t M, hipStream_t s);
int count_touched(const tr_plan* p, BinData& b, hipStream_t s);
void free_plan(tr_plan* p);
int check_plan_bin(const tr_plan* p, int bin);

// Kernel launchers (one source file per kernel family).
int launch_fwd_x(int kc, int L, dim3 grid, size_t smem, hipStream_t s, const FwdXArgs& a);
// The inverse x pass without a coil sum (C == 1, no maps): grid (image rows, channel chunks of kc),
// out[k] = crop(IFFT_x(channel k)) * apodisation * scale for every channel k.
int launch_adj_x_channels(int kc, int L, dim3 grid, size_t smem, hipStream_t s, const AdjXArgs& a);
int launch_adj_x(int kc, int L, dim3 grid, size_t smem, hipStream_t s, const AdjXArgs& a);
int launch_lines(int kc, int L, dim3 grid, size_t smem, hipStream_t s, const LinesArgs& a);
int launch_interp(const InterpArgs& a, hipStream_t s);
int launch_spread(SpreadArgs a, const BinData& b, tr_plan* p, hipStream_t s);

// Oversampled-grid lengths with a compile-time FFT plan for channel chunks of
// 8 and 16; other (length, chunk) pairs run the runtime-radix kernels.
#define TR_FAST_LENGTHS(X) X(40) X(50) X(60) X(64) X(80) X(96) X(100) X(120) X(128) X(160) X(200) X(240)

template <class Launch, int KC, class Args>
inline void launch_len(int L, dim3 g, size_t sm, hipStream_t s, const Args& a) {
  switch (L) {
#define TR_CASE_(n)                                 \
  case n:                                           \
    Launch::template go<KC, n>(g, sm, s, a);        \
    return;
    TR_FAST_LENGTHS(TR_CASE_)
#undef TR_CASE_
    default:
      Launch::template go<KC, 0>(g, sm, s, a);
  }
}

template <class Launch, class Args>
inline int launch_fft_kernel(int kc, int L, dim3 g, size_t sm, hipStream_t s, const Args& a) {
  switch (kc) {
    case 1: Launch::template go<1, 0>(g, sm, s, a); break;
    case 2: Launch::template go<2, 0>(g, sm, s, a); break;
    case 4: Launch::template go<4, 0>(g, sm, s, a); break;
    case 8: launch_len<Launch, 8>(L, g, sm, s, a); break;
    case 16: launch_len<Launch, 16>(L, g, sm, s, a); break;
    default: TR_FAIL(kUnsupported, "channel chunk must be a power of two <= 16");
  }
  TR_LAUNCH_CHECK();
  return kOk;
}

}  // namespace tr
