// 3-D db6 wavelet of the TVMW spatial regulariser on a batch of complex images.
//
// Matches ptwt.wavedec3(x, "db6", mode="zero", level) / ptwt.waverec3 as called
// at src/toporecon/algorithms/tvmw.py:196-206 (real and imaginary parts see the
// same real filters): one analysis level is a full convolution with the
// decomposition filter along z, then y, then x, keeping odd output indices
// (length n -> (n + 11) / 2); synthesis is its exact transpose.  Coefficients
// of one image are stored flat in the order cA_L, level-L details, ...,
// level-1 details, each level's details as aad, ada, add, daa, dad, dda, ddd
// (letters = z, y, x; a = low pass, d = high pass).

#include "tr_internal.hpp"

namespace tr {

// db6 decomposition filters (Daubechies' 12-tap extremal-phase scaling filter,
// time-reversed, in PyWavelets' orientation; high pass by the QMF relation).
__constant__ float c_dec_lo[12] = {
    -0.0010773010853084792f, 0.0047772575109455064f, 0.0005538422011614997f, -0.031582039317486009f,
    0.027522865530305706f,   0.097501605587322862f,  -0.12976686756726172f,  -0.22626469396543944f,
    0.31525035170919768f,    0.75113390802109525f,   0.49462389039845273f,   0.1115407433501094f};
__constant__ float c_dec_hi[12] = {
    -0.1115407433501094f,    0.49462389039845273f,   -0.75113390802109525f, 0.31525035170919768f,
    0.22626469396543944f,    -0.12976686756726172f,  -0.097501605587322862f, 0.027522865530305706f,
    0.031582039317486009f,   0.0005538422011614997f, -0.0047772575109455064f, -0.0010773010853084792f};

constexpr int kFilt = 12;

struct AxisArgs {
  int64_t rows;   // images * rows per image
  int rows_img;   // rows per image
  int n;          // signal length along the axis
  int c;          // coefficient length (n + 11) / 2
  int64_t Q;      // contiguous trailing extent
  int64_t sig_stride;  // complex elements per image of the signal array
  int64_t lo_stride, hi_stride;
  int64_t lo_off, hi_off;
  const int* gate = nullptr;  // batched solves: the launch returns at once while *gate != 0
};

// Analysis along one axis: lo[k] = sum_j h[j] x[2k + 1 - j], zeros outside [0, n).  One thread
// computes the coefficient pair k = 2p, 2p + 1, whose taps x[4p - 10 .. 4p + 3] overlap in 10 of 12:
// 14 loads for two coefficients instead of 24.  Every tap is loaded from a clamped, in-bounds index
// before any is used (no load sits behind a per-lane branch); taps outside [0, n) are then skipped,
// and each sum runs over j in increasing order as before.
__global__ void k_dwt_axis(AxisArgs a, const float2* __restrict__ x, float2* __restrict__ lo,
                           float2* __restrict__ hi) {
  if (gate_closed(a.gate)) return;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int cp = (a.c + 1) / 2;  // coefficient pairs per row
  const int64_t total = a.rows * cp * a.Q;
  if (t >= total) return;
  const int64_t q = t % a.Q;
  const int p = (int)((t / a.Q) % cp);
  const int64_t row = t / (a.Q * cp);
  const int64_t img = row / a.rows_img;
  const int64_t r = row - img * a.rows_img;
  const float2* xs = x + img * a.sig_stride + r * a.n * a.Q + q;
  const int m0 = 4 * p - 10;  // first tap of the pair
  float2 v[kFilt + 2];
#pragma unroll
  for (int i = 0; i < kFilt + 2; ++i) v[i] = xs[(int64_t)min(max(m0 + i, 0), a.n - 1) * a.Q];
  const int k0 = 2 * p, k1 = 2 * p + 1;
  float2 sl0 = make_float2(0.f, 0.f), sh0 = sl0, sl1 = sl0, sh1 = sl0;
#pragma unroll
  for (int j = 0; j < kFilt; ++j) {
    const int ma = 2 * k0 + 1 - j, mb = 2 * k1 + 1 - j;  // v[11 - j], v[13 - j]
    if (ma >= 0 && ma < a.n) {
      sl0.x += c_dec_lo[j] * v[11 - j].x;
      sl0.y += c_dec_lo[j] * v[11 - j].y;
      sh0.x += c_dec_hi[j] * v[11 - j].x;
      sh0.y += c_dec_hi[j] * v[11 - j].y;
    }
    if (mb >= 0 && mb < a.n) {
      sl1.x += c_dec_lo[j] * v[13 - j].x;
      sl1.y += c_dec_lo[j] * v[13 - j].y;
      sh1.x += c_dec_hi[j] * v[13 - j].x;
      sh1.y += c_dec_hi[j] * v[13 - j].y;
    }
  }
  const int64_t o = (r * a.c + k0) * a.Q + q;
  lo[img * a.lo_stride + a.lo_off + o] = sl0;
  hi[img * a.hi_stride + a.hi_off + o] = sh0;
  if (k1 < a.c) {
    lo[img * a.lo_stride + a.lo_off + o + a.Q] = sl1;
    hi[img * a.hi_stride + a.hi_off + o + a.Q] = sh1;
  }
}

// Synthesis along one axis (transpose of k_dwt_axis):
// x[m] = sum_k lo[k] h_lo[2k + 1 - m] + hi[k] h_hi[2k + 1 - m].
// The taps of x[m] are j = 2k + 1 - m in [0, 11], i.e. k = m / 2 + i, i = 0 .. 5 (j = 2i + 1 for
// even m, 2i for odd m), k <= c - 1: the pair x[2p], x[2p + 1] reads the same 6 + 6 coefficients, so
// one thread computes both from one set of loads.  All 12 values are loaded (clamped, in bounds)
// before any is used; each sum runs over increasing k as before.
__global__ void k_idwt_axis(AxisArgs a, const float2* __restrict__ lo, const float2* __restrict__ hi,
                            float2* __restrict__ x) {
  if (gate_closed(a.gate)) return;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int np = (a.n + 1) / 2;  // output pairs per row
  const int64_t total = a.rows * np * a.Q;
  if (t >= total) return;
  const int64_t q = t % a.Q;
  const int p = (int)((t / a.Q) % np);
  const int64_t row = t / (a.Q * np);
  const int64_t img = row / a.rows_img;
  const int64_t r = row - img * a.rows_img;
  const float2* ls = lo + img * a.lo_stride + a.lo_off + r * a.c * a.Q + q;
  const float2* hs = hi + img * a.hi_stride + a.hi_off + r * a.c * a.Q + q;
  const int kb = p;
  float2 vl[6], vh[6];
#pragma unroll
  for (int i = 0; i < 6; ++i) {
    const int64_t kc = min(kb + i, a.c - 1);
    vl[i] = ls[kc * a.Q];
    vh[i] = hs[kc * a.Q];
  }
  float2 se = make_float2(0.f, 0.f), so = se;
#pragma unroll
  for (int i = 0; i < 6; ++i) {
    if (kb + i <= a.c - 1) {
      se.x += c_dec_lo[2 * i + 1] * vl[i].x + c_dec_hi[2 * i + 1] * vh[i].x;
      se.y += c_dec_lo[2 * i + 1] * vl[i].y + c_dec_hi[2 * i + 1] * vh[i].y;
      so.x += c_dec_lo[2 * i] * vl[i].x + c_dec_hi[2 * i] * vh[i].x;
      so.y += c_dec_lo[2 * i] * vl[i].y + c_dec_hi[2 * i] * vh[i].y;
    }
  }
  float2* xo = x + img * a.sig_stride + (r * a.n + 2 * p) * a.Q + q;
  xo[0] = se;
  if (2 * p + 1 < a.n) xo[a.Q] = so;
}

// p <- proj(p + sigma w): t * min(1, lam / sqrt(re^2 + im^2 + 1e-12)); 0 if lam <= 0
// (tvmw.py:162-193 tree axpy + L-inf projection of every sub-band).
__global__ void k_w2_dual(float2* p, const float2* w, int64_t n, float sigma, float lam, const int* gate) {
  if (gate_closed(gate)) return;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (lam <= 0.f) {
    p[i] = make_float2(0.f, 0.f);
    return;
  }
  const float2 a = p[i], b = w[i];
  const float2 t = make_float2(a.x + sigma * b.x, a.y + sigma * b.y);
  const float mag = sqrtf(t.x * t.x + t.y * t.y + 1e-12f);
  const float s = fminf(lam / mag, 1.0f);
  p[i] = make_float2(t.x * s, t.y * s);
}

// ---- host -------------------------------------------------------------------
struct Level {
  int n[3];  // input (signal) shape of this level
  int c[3];  // coefficient shape
};

static int coeff_len(int n) { return (n + kFilt - 1) / 2; }

static int make_levels(int Z, int Y, int X, int level, std::vector<Level>& lv, int64_t* ncoef,
                       std::vector<int64_t>& det_off, int64_t* work) {
  if (level < 0 || level > 8) TR_FAIL(kInvalidArgument, "level must be in [0, 8]");
  int n[3] = {Z, Y, X};
  lv.clear();
  for (int l = 0; l < level; ++l) {
    Level L;
    for (int i = 0; i < 3; ++i) {
      L.n[i] = n[i];
      L.c[i] = coeff_len(n[i]);
      n[i] = L.c[i];
    }
    lv.push_back(L);
  }
  if (level == 0) {
    *ncoef = (int64_t)Z * Y * X;
    det_off.clear();
    *work = 0;
    return kOk;
  }
  const Level& top = lv.back();
  const int64_t SL = (int64_t)top.c[0] * top.c[1] * top.c[2];
  // offsets of the detail blocks: level L first, level 1 last
  det_off.assign(level, 0);
  int64_t off = SL;
  for (int l = level - 1; l >= 0; --l) {
    det_off[l] = off;
    off += 7 * (int64_t)lv[l].c[0] * lv[l].c[1] * lv[l].c[2];
  }
  *ncoef = off;
  const Level& f = lv[0];
  const int64_t sz = (int64_t)f.c[0] * f.n[1] * f.n[2];
  const int64_t sy = (int64_t)f.c[0] * f.c[1] * f.n[2];
  const int64_t s1 = (int64_t)f.c[0] * f.c[1] * f.c[2];
  *work = 2 * sz + 4 * sy + 2 * s1;
  return kOk;
}

static int launch_axis(const AxisArgs& a0, const float2* x, float2* lo, float2* hi, hipStream_t s) {
  AxisArgs a = a0;
  a.gate = gate_current();
  const int64_t total = a.rows * ((a.c + 1) / 2) * a.Q;  // one thread per coefficient pair
  if (total == 0) return kOk;
  hipLaunchKernelGGL(k_dwt_axis, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, a, x, lo, hi);
  TR_LAUNCH_CHECK();
  return kOk;
}

static int launch_iaxis(const AxisArgs& a0, const float2* lo, const float2* hi, float2* x, hipStream_t s) {
  AxisArgs a = a0;
  a.gate = gate_current();
  const int64_t total = a.rows * ((a.n + 1) / 2) * a.Q;  // one thread per output pair
  if (total == 0) return kOk;
  hipLaunchKernelGGL(k_idwt_axis, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, a, lo, hi, x);
  TR_LAUNCH_CHECK();
  return kOk;
}

static AxisArgs axis(int64_t nimg, int rows_img, int n, int c, int64_t Q, int64_t sig_stride,
                     int64_t lo_stride, int64_t lo_off, int64_t hi_stride, int64_t hi_off) {
  AxisArgs a;
  a.rows = nimg * rows_img;
  a.rows_img = rows_img;
  a.n = n;
  a.c = c;
  a.Q = Q;
  a.sig_stride = sig_stride;
  a.lo_stride = lo_stride;
  a.hi_stride = hi_stride;
  a.lo_off = lo_off;
  a.hi_off = hi_off;
  return a;
}

}  // namespace tr

using namespace tr;

extern "C" {

int64_t tr_w2_coeff_count(int Z, int Y, int X, int level) {
  std::vector<Level> lv;
  std::vector<int64_t> off;
  int64_t n = -1, w = 0;
  if (make_levels(Z, Y, X, level, lv, &n, off, &w)) return -1;
  return n;
}

int64_t tr_w2_work_count(int Z, int Y, int X, int level) {
  std::vector<Level> lv;
  std::vector<int64_t> off;
  int64_t n = 0, w = -1;
  if (make_levels(Z, Y, X, level, lv, &n, off, &w)) return -1;
  return w;
}

int tr_w2_forward(const tr_c64* img, int64_t nimg, int Z, int Y, int X, int level, tr_c64* coeffs,
                  tr_c64* work, void* stream) {
  std::vector<Level> lv;
  std::vector<int64_t> det;
  int64_t ncoef = 0, wn = 0;
  if (int r = make_levels(Z, Y, X, level, lv, &ncoef, det, &wn)) return r;
  hipStream_t s = (hipStream_t)stream;
  ProfScope ps(kKW2Forward, 2.0 * 8.0 * nimg * (double)Z * Y * X + 8.0 * nimg * ncoef, s);
  if (level == 0) {
    TR_HIP(hipMemcpyAsync(coeffs, img, sizeof(float2) * nimg * ncoef, hipMemcpyDeviceToDevice, s));
    return kOk;
  }
  const Level& f = lv[0];
  const int64_t sz = (int64_t)f.c[0] * f.n[1] * f.n[2];
  const int64_t sy = (int64_t)f.c[0] * f.c[1] * f.n[2];
  const int64_t s1 = (int64_t)f.c[0] * f.c[1] * f.c[2];
  float2* W = F2(work);
  float2* zl = W;                        // nimg * sz each (per-level compact layouts)
  float2* zh = zl + nimg * sz;
  float2* yb[4] = {zh + nimg * sz, nullptr, nullptr, nullptr};
  for (int i = 1; i < 4; ++i) yb[i] = yb[i - 1] + nimg * sy;
  float2* abuf[2] = {yb[3] + nimg * sy, yb[3] + nimg * sy + nimg * s1};
  float2* C = F2(coeffs);
  const float2* a_in = F2(img);
  int64_t a_stride = (int64_t)Z * Y * X;
  for (int l = 0; l < level; ++l) {
    const Level& L = lv[l];
    const int n0 = L.n[0], n1 = L.n[1], n2 = L.n[2];
    const int c0 = L.c[0], c1 = L.c[1], c2 = L.c[2];
    const int64_t lz = (int64_t)c0 * n1 * n2, ly = (int64_t)c0 * c1 * n2, lc = (int64_t)c0 * c1 * c2;
    // z
    if (int r = launch_axis(axis(nimg, 1, n0, c0, (int64_t)n1 * n2, a_stride, lz, 0, lz, 0), a_in, zl, zh, s))
      return r;
    // y
    if (int r = launch_axis(axis(nimg, c0, n1, c1, n2, lz, ly, 0, ly, 0), zl, yb[0], yb[1], s)) return r;
    if (int r = launch_axis(axis(nimg, c0, n1, c1, n2, lz, ly, 0, ly, 0), zh, yb[2], yb[3], s)) return r;
    // x: LL -> (aaa, aad), LH -> (ada, add), HL -> (daa, dad), HH -> (dda, ddd)
    const bool last = (l == level - 1);
    float2* a_out = last ? C : abuf[l & 1];
    const int64_t a_out_stride = last ? ncoef : lc;
    const int64_t d = det[l];
    if (int r = launch_axis(axis(nimg, c0 * c1, n2, c2, 1, ly, a_out_stride, 0, ncoef, d + 0 * lc), yb[0],
                            a_out, C, s))
      return r;
    if (int r = launch_axis(axis(nimg, c0 * c1, n2, c2, 1, ly, ncoef, d + 1 * lc, ncoef, d + 2 * lc), yb[1],
                            C, C, s))
      return r;
    if (int r = launch_axis(axis(nimg, c0 * c1, n2, c2, 1, ly, ncoef, d + 3 * lc, ncoef, d + 4 * lc), yb[2],
                            C, C, s))
      return r;
    if (int r = launch_axis(axis(nimg, c0 * c1, n2, c2, 1, ly, ncoef, d + 5 * lc, ncoef, d + 6 * lc), yb[3],
                            C, C, s))
      return r;
    a_in = a_out;
    a_stride = lc;
  }
  (void)sz;
  (void)sy;
  (void)s1;
  return kOk;
}

int tr_w2_adjoint(const tr_c64* coeffs, int64_t nimg, int Z, int Y, int X, int level, tr_c64* out,
                  tr_c64* work, void* stream) {
  std::vector<Level> lv;
  std::vector<int64_t> det;
  int64_t ncoef = 0, wn = 0;
  if (int r = make_levels(Z, Y, X, level, lv, &ncoef, det, &wn)) return r;
  hipStream_t s = (hipStream_t)stream;
  ProfScope ps(kKW2Adjoint, 2.0 * 8.0 * nimg * (double)Z * Y * X + 8.0 * nimg * ncoef, s);
  if (level == 0) {
    TR_HIP(hipMemcpyAsync(out, coeffs, sizeof(float2) * nimg * ncoef, hipMemcpyDeviceToDevice, s));
    return kOk;
  }
  const Level& f = lv[0];
  const int64_t sz = (int64_t)f.c[0] * f.n[1] * f.n[2];
  const int64_t sy = (int64_t)f.c[0] * f.c[1] * f.n[2];
  const int64_t s1 = (int64_t)f.c[0] * f.c[1] * f.c[2];
  float2* W = F2(work);
  float2* zl = W;
  float2* zh = zl + nimg * sz;
  float2* yb[4] = {zh + nimg * sz, nullptr, nullptr, nullptr};
  for (int i = 1; i < 4; ++i) yb[i] = yb[i - 1] + nimg * sy;
  float2* abuf[2] = {yb[3] + nimg * sy, yb[3] + nimg * sy + nimg * s1};
  const float2* C = F2(coeffs);
  // a_L lives in the coefficient buffer (offset 0, stride ncoef)
  const float2* a_in = C;
  int64_t a_in_stride = ncoef;
  for (int l = level - 1; l >= 0; --l) {
    const Level& L = lv[l];
    const int n0 = L.n[0], n1 = L.n[1], n2 = L.n[2];
    const int c0 = L.c[0], c1 = L.c[1], c2 = L.c[2];
    const int64_t lz = (int64_t)c0 * n1 * n2, ly = (int64_t)c0 * c1 * n2, lc = (int64_t)c0 * c1 * c2;
    const int64_t d = det[l];
    // x: (aaa, aad) -> LL, (ada, add) -> LH, (daa, dad) -> HL, (dda, ddd) -> HH
    // The aaa input has its own stride, so LL is assembled with an a-stride axis call.
    {
      AxisArgs a = axis(nimg, c0 * c1, n2, c2, 1, ly, a_in_stride, 0, ncoef, d + 0 * lc);
      if (int r = launch_iaxis(a, a_in, C, yb[0], s)) return r;
    }
    if (int r = launch_iaxis(axis(nimg, c0 * c1, n2, c2, 1, ly, ncoef, d + 1 * lc, ncoef, d + 2 * lc), C, C,
                             yb[1], s))
      return r;
    if (int r = launch_iaxis(axis(nimg, c0 * c1, n2, c2, 1, ly, ncoef, d + 3 * lc, ncoef, d + 4 * lc), C, C,
                             yb[2], s))
      return r;
    if (int r = launch_iaxis(axis(nimg, c0 * c1, n2, c2, 1, ly, ncoef, d + 5 * lc, ncoef, d + 6 * lc), C, C,
                             yb[3], s))
      return r;
    // y: (LL, LH) -> L, (HL, HH) -> H
    if (int r = launch_iaxis(axis(nimg, c0, n1, c1, n2, lz, ly, 0, ly, 0), yb[0], yb[1], zl, s)) return r;
    if (int r = launch_iaxis(axis(nimg, c0, n1, c1, n2, lz, ly, 0, ly, 0), yb[2], yb[3], zh, s)) return r;
    // z: (L, H) -> a_{l-1}
    const bool first = (l == 0);
    float2* a_out = first ? F2(out) : abuf[l & 1];
    const int64_t a_out_stride = first ? (int64_t)Z * Y * X : (int64_t)n0 * n1 * n2;
    if (int r = launch_iaxis(axis(nimg, 1, n0, c0, (int64_t)n1 * n2, a_out_stride, lz, 0, lz, 0), zl, zh,
                             a_out, s))
      return r;
    a_in = a_out;
    a_in_stride = a_out_stride;
  }
  (void)sz;
  (void)sy;
  (void)s1;
  return kOk;
}

int tr_w2_dual_update(tr_c64* p, const tr_c64* w, int64_t n, float sigma, float lam, void* stream) {
  if (n <= 0) return kOk;
  hipStream_t s = (hipStream_t)stream;
  ProfScope ps(kKW2Dual, 24.0 * n, s);
  hipLaunchKernelGGL(k_w2_dual, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, F2(p), F2(w), n, sigma,
                     lam, gate_current());
  TR_LAUNCH_CHECK();
  return kOk;
}

}  // extern "C"
