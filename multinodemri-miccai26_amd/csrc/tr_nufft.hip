// 3-D Kaiser-Bessel NUFFT (SigPy 0.1.27 semantics, width 4, oversampling 1.25)
// for MI355X: channel-innermost grids, LDS Stockham FFT passes with pruned
// zero-padding, a gather interpolation fused with the PDHG dual update, and a
// deterministic CSR gather for the adjoint spreading.  Host orchestration and
// the NUFFT / data-term C entry points.

#include <algorithm>

#include "tr_internal.hpp"

namespace tr {

// Forward passes x (with pad), y, z into p->work for K = E*C channels.  With
// prune, grid points outside the plan's sample support are left unwritten
// (nothing reads them).
static int forward_to_grid(tr_plan* p, const float2* img, const float2* maps, int E, int C,
                           hipStream_t s, bool prune = true) {
  prune = prune && p->prune && p->col_win;
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
    const int kc = fft_kc(C, p->ax[2].M);
    dim3 grid((unsigned)(nz * ny * ((C + kc - 1) / kc)));
    ProfScope ps(kKFwdXPad, 8.0 * (E + C) * p->N + 8.0 * K * nz * ny * gx, s);
    // LDS: twiddles + FFT tile + the image rows of a group of echoes (staged for the block's coils,
    // the coils' maps staying in registers across all echoes); the group is sized so that four
    // blocks share a CU's 160 KiB (four waves per SIMD), else as many rows as 64 KiB holds
    const size_t tile = fft_smem(p->ax[2].M, kc), row = sizeof(float2) * (size_t)p->n[2];
    if (tile + row > 65536) TR_FAIL(kUnsupported, "image rows too long for the forward x pass (LDS)");
    const size_t budget = tile + row <= 40960 ? 40960 : 65536;
    a.E = E;
    a.eg = (int)std::max<size_t>(1, std::min<size_t>((size_t)E, (budget - tile) / row));
    a.img = img;
    a.grid = p->work;
    if (int r = launch_fwd_x(kc, gx, grid, tile + row * a.eg, s, a)) return r;
  }
  {  // y pass over the nz occupied z-planes, windowed load along y
    const int kc = fft_kc(K, p->ax[1].M);
    const int nchunk = (K + kc - 1) / kc;
    LinesArgs a;
    a.xcd_chunk = (unsigned)(p->lines_xcd_lines * nchunk);
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
    a.line_win = prune ? p->ywin : nullptr;
    a.win_so = 0;
    a.win_si = 1;
    a.zero_empty = 0;
    const double cols = prune ? (double)p->live_cols : (double)gy * gx;
    dim3 grid((unsigned)(nz * gx * nchunk));
    ProfScope ps(kKFftYFwd, 8.0 * K * nz * ((double)gx * ny + cols), s);
    if (int r = launch_lines(kc, gy, grid, fft_smem(p->ax[1].M, kc), s, a)) return r;
  }
  {  // z pass over every (y, x) line, windowed load along z
    const int kc = fft_kc(K, p->ax[0].M);
    const int nchunk = (K + kc - 1) / kc;
    LinesArgs a;
    a.xcd_chunk = (unsigned)(p->lines_xcd_lines * nchunk);
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
    a.line_win = prune ? p->col_win : nullptr;
    a.win_so = gx;
    a.win_si = 1;
    a.zero_empty = 0;  // the y pass reads only columns inside its windows
    const double cols = prune ? (double)p->live_cols : (double)gy * gx;
    const double pts = prune ? (double)p->support_points : (double)p->G;
    dim3 grid((unsigned)(gy * gx * nchunk));
    ProfScope ps(kKFftZFwd, 8.0 * K * (cols * nz + pts), s);
    if (int r = launch_lines(kc, gz, grid, fft_smem(p->ax[0].M, kc), s, a)) return r;
  }
  return kOk;
}

// Inverse passes z, y (cropping) on p->work, then x with crop/coil sum into out.
static int grid_to_image(tr_plan* p, const float2* maps, int E, int C, float2* out, int accumulate,
                         hipStream_t s, bool prune = true) {
  prune = prune && p->prune && p->col_win;
  const int K = E * C;
  const int nz = p->n[0], ny = p->n[1];
  const int gz = p->g[0], gy = p->g[1], gx = p->g[2];
  {
    const int kc = fft_kc(K, p->ax[0].M);
    const int nchunk = (K + kc - 1) / kc;
    LinesArgs a;
    a.xcd_chunk = (unsigned)(p->lines_xcd_lines * nchunk);
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
    a.line_win = prune ? p->col_win : nullptr;
    a.win_so = gx;
    a.win_si = 1;
    a.zero_empty = 0;  // the y pass reads only columns inside its windows
    const double cols = prune ? (double)p->live_cols : (double)gy * gx;
    const double pts = prune ? (double)p->support_points : (double)p->G;
    dim3 grid((unsigned)(gy * gx * nchunk));
    ProfScope ps(kKFftZInv, 8.0 * K * (cols * nz + pts), s);
    if (int r = launch_lines(kc, gz, grid, fft_smem(p->ax[0].M, kc), s, a)) return r;
  }
  {
    const int kc = fft_kc(K, p->ax[1].M);
    const int nchunk = (K + kc - 1) / kc;
    LinesArgs a;
    a.xcd_chunk = (unsigned)(p->lines_xcd_lines * nchunk);
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
    a.line_win = prune ? p->ywin : nullptr;
    a.win_so = 0;
    a.win_si = 1;
    a.zero_empty = 1;  // the x pass reads every (z, y, kx) it leaves
    const double cols = prune ? (double)p->live_cols : (double)gy * gx;
    dim3 grid((unsigned)(nz * gx * nchunk));
    ProfScope ps(kKFftYInv, 8.0 * K * nz * ((double)gx * ny + cols), s);
    if (int r = launch_lines(kc, gy, grid, fft_smem(p->ax[1].M, kc), s, a)) return r;
  }
  {
    AdjXArgs a;
    a.gg = geom(p);
    a.ax = p->ax[2];
    a.tw = p->tw[2];
    a.maps = maps;
    a.apod_z = p->apod[0];
    a.apod_y = p->apod[1];
    a.apod_x = p->apod[2];
    a.scale = p->pad_scale;
    a.C = C;
    a.K = K;
    a.accumulate = accumulate;
    ProfScope ps(kKAdjXCrop, 8.0 * K * nz * ny * gx + 8.0 * C * p->N + 8.0 * E * p->N * (accumulate ? 2 : 1), s);
    if (!maps && C == 1) {  // no coil sum: every channel is its own image, FFTs run many channels wide
      const int kc = fft_kc(K, p->ax[2].M);
      a.E = K;
      a.grid = p->work;
      a.out = out;
      dim3 grid((unsigned)(nz * ny), (unsigned)((K + kc - 1) / kc));
      return launch_adj_x_channels(kc, gx, grid, fft_smem(p->ax[2].M, kc), s, a);
    }
    const int kcc = fft_kc(C, p->ax[2].M);
    dim3 grid((unsigned)(nz * ny));
    for (int e0 = 0; e0 < E; e0 += kAdjXEchoes) {  // at most kAdjXEchoes per launch
      a.E = std::min(kAdjXEchoes, E - e0);
      a.grid = p->work + (int64_t)e0 * C;
      a.out = out + (int64_t)e0 * p->N;
      if (int r = launch_adj_x(kcc, gx, grid, fft_smem(p->ax[2].M, kcc), s, a)) return r;
    }
  }
  return kOk;
}

}  // namespace tr

using namespace tr;

extern "C" {

int tr_nufft_forward(tr_plan* p, int bin, const tr_c64* img, const tr_c64* maps, int E, int C,
                     tr_c64* out, void* stream) {
  if (int r = check_plan_bin(p, bin)) return r;
  use_slot(p, bin % p->nslots);
  if (E < 1 || C < 1 || (!maps && C != 1)) TR_FAIL(kInvalidArgument, "need E >= 1 and C == 1 without maps");
  hipStream_t s = (hipStream_t)stream;
  const BinData& b = p->bins[bin];
  if (b.M == 0) return kOk;
  if (int r = forward_to_grid(p, F2(img), F2(maps), E, C, s)) return r;
  InterpArgs a = {};
  a.gg = geom(p);
  a.xcd_chunk = p->interp_xcd_chunk;
  a.kgroups = p->interp_kgroups;
  a.tap = b.tap;
  a.touched = b.touched;
  a.grp = b.grp;
  a.tapg = b.tapg;
  a.ngroups = b.ngroups;
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
  use_slot(p, bin % p->nslots);
  if (E < 1 || C < 1 || (!maps && C != 1)) TR_FAIL(kInvalidArgument, "need E >= 1 and C == 1 without maps");
  hipStream_t s = (hipStream_t)stream;
  const BinData& b = p->bins[bin];
  const int K = E * C;
  if (int r = ensure_work(p, K)) return r;
  SpreadArgs a = {};
  a.K = K;
  a.M = b.M;
  a.sorted_src = 0;
  a.ksp = F2(ksp);
  a.dcf_orig = dcf;
  a.perm = b.perm;
  a.grid = p->work;
  if (int r = launch_spread(a, b, p, s)) return r;
  return grid_to_image(p, F2(maps), E, C, F2(out), accumulate, s);
}

int tr_nufft_forward_grid(tr_plan* p, const tr_c64* img, const tr_c64* maps, int E, int C, tr_c64* grid,
                          void* stream) {
  if (!p) TR_FAIL(kInvalidArgument, "null plan");
  use_slot(p, 0);
  if (E < 1 || C < 1 || (!maps && C != 1)) TR_FAIL(kInvalidArgument, "need E >= 1 and C == 1 without maps");
  hipStream_t s = (hipStream_t)stream;
  if (int r = forward_to_grid(p, F2(img), F2(maps), E, C, s, false)) return r;
  TR_HIP(hipMemcpyAsync(grid, p->work, sizeof(float2) * p->G * E * C, hipMemcpyDeviceToDevice, s));
  return kOk;
}

int tr_nufft_adjoint_grid(tr_plan* p, const tr_c64* grid, const tr_c64* maps, int E, int C, tr_c64* out,
                          void* stream) {
  if (!p) TR_FAIL(kInvalidArgument, "null plan");
  use_slot(p, 0);
  if (E < 1 || C < 1 || (!maps && C != 1)) TR_FAIL(kInvalidArgument, "need E >= 1 and C == 1 without maps");
  hipStream_t s = (hipStream_t)stream;
  if (int r = ensure_work(p, (int64_t)E * C)) return r;
  TR_HIP(hipMemcpyAsync(p->work, grid, sizeof(float2) * p->G * E * C, hipMemcpyDeviceToDevice, s));
  return grid_to_image(p, F2(maps), E, C, F2(out), 0, s, false);
}

int tr_data_dual_update(tr_plan* p, int bin, const tr_c64* ubar, const tr_c64* maps, int E, int C,
                        const tr_c64* y_sorted, tr_c64* q_sorted, float sigma, float one_plus_sigma,
                        void* stream) {
  if (int r = check_plan_bin(p, bin)) return r;
  use_slot(p, bin % p->nslots);
  hipStream_t s = (hipStream_t)stream;
  const BinData& b = p->bins[bin];
  if (b.M == 0) return kOk;
  if (int r = forward_to_grid(p, F2(ubar), F2(maps), E, C, s)) return r;
  InterpArgs a = {};
  a.gg = geom(p);
  a.xcd_chunk = p->interp_xcd_chunk;
  a.kgroups = p->interp_kgroups;
  a.tap = b.tap;
  a.touched = b.touched;
  a.grp = b.grp;
  a.tapg = b.tapg;
  a.ngroups = b.ngroups;
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
  use_slot(p, bin % p->nslots);
  hipStream_t s = (hipStream_t)stream;
  const BinData& b = p->bins[bin];
  if (b.M == 0) return kOk;
  if (int r = forward_to_grid(p, F2(x), F2(maps), E, C, s)) return r;
  InterpArgs a = {};
  a.gg = geom(p);
  a.xcd_chunk = p->interp_xcd_chunk;
  a.kgroups = p->interp_kgroups;
  a.tap = b.tap;
  a.touched = b.touched;
  a.grp = b.grp;
  a.tapg = b.tapg;
  a.ngroups = b.ngroups;
  a.M = b.M;
  a.K = E * C;
  a.grid = p->work;
  a.mode = kInterpSim;
  a.q = F2(y_sorted);
  return launch_interp(a, s);
}

int tr_data_residual(tr_plan* p, int bin, const tr_c64* x, const tr_c64* maps, int E, int C,
                     const tr_c64* y_sorted, tr_c64* q_sorted, void* stream) {
  if (int r = check_plan_bin(p, bin)) return r;
  use_slot(p, bin % p->nslots);
  hipStream_t s = (hipStream_t)stream;
  const BinData& b = p->bins[bin];
  if (b.M == 0) return kOk;
  if (int r = forward_to_grid(p, F2(x), F2(maps), E, C, s)) return r;
  InterpArgs a = {};
  a.gg = geom(p);
  a.xcd_chunk = p->interp_xcd_chunk;
  a.kgroups = p->interp_kgroups;
  a.tap = b.tap;
  a.touched = b.touched;
  a.grp = b.grp;
  a.tapg = b.tapg;
  a.ngroups = b.ngroups;
  a.M = b.M;
  a.K = E * C;
  a.grid = p->work;
  a.mode = kInterpResid;
  a.y = F2(y_sorted);
  a.q = F2(q_sorted);
  return launch_interp(a, s);
}

int tr_data_adjoint(tr_plan* p, int bin, const tr_c64* q_sorted, const float* dcf_sorted,
                    const tr_c64* maps, int E, int C, tr_c64* g, int accumulate, void* stream) {
  if (int r = check_plan_bin(p, bin)) return r;
  use_slot(p, bin % p->nslots);
  hipStream_t s = (hipStream_t)stream;
  const BinData& b = p->bins[bin];
  const int K = E * C;
  if (int r = ensure_work(p, K)) return r;
  SpreadArgs a = {};
  a.K = K;
  a.M = b.M;
  a.sorted_src = 1;
  a.q = F2(q_sorted);
  a.dcf_sorted = dcf_sorted;
  a.grid = p->work;
  if (int r = launch_spread(a, b, p, s)) return r;
  return grid_to_image(p, F2(maps), E, C, F2(g), accumulate, s);
}

}  // extern "C"
