// Fused element-wise PDHG kernels for the motion/echo TV solvers, single GPU.
//
// Arrays are (B, E, Z, Y, X) complex64, N = Z*Y*X voxels.  One thread owns a
// (bin, voxel) column and walks the echoes, so the L2 coupling over echoes and
// the echo divergence stay in registers.  Boundary rules are those of one
// process (all MPI peers PROC_NULL, ghost slabs zero).  Stopping-rule sums are
// accumulated in float64 per block and reduced in a fixed order.

#include <mutex>
#include <string>

#include "toporecon_hip.h"
#include "tr_common.hpp"

namespace tr {

constexpr int kEwBlock = 256;

struct VolGeom {
  int B, E, Z, Y, X;
  int64_t N;
  const int* gate = nullptr;  // batched solves: the launch returns at once while *gate != 0
};

__device__ __forceinline__ float2 ld(const float2* p, int64_t i) { return p[i]; }

// Block-wide float64 pair reduction to partial[blockIdx.x].
__device__ __forceinline__ void block_reduce2(double a, double b, double2* partial) {
  __shared__ double sa[kEwBlock / 64], sb[kEwBlock / 64];
  for (int off = 32; off > 0; off >>= 1) {
    a += __shfl_down(a, off, 64);
    b += __shfl_down(b, off, 64);
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) {
    sa[wave] = a;
    sb[wave] = b;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double ta = 0.0, tb = 0.0;
    for (int w = 0; w < kEwBlock / 64; ++w) {
      ta += sa[w];
      tb += sb[w];
    }
    partial[blockIdx.x] = make_double2(ta, tb);
  }
}

// p_m update with the motion forward difference and its prox, for one (b, v).
// L2 coupling scales by max(1, sqrt(sum_e |p|^2) / lambda) (tvm.py:276-306);
// L1 by max(1, |p| / lambda) per echo (tvm.py:307-309).
__device__ __forceinline__ void motion_dual(const VolGeom& g, int b, int64_t v, const float2* ubar,
                                            float2* p_m, float sigma, float lam, int l2) {
  const int64_t EN = (int64_t)g.E * g.N;
  const int64_t base = (int64_t)b * EN + v;
  constexpr int kRegE = 8;
  if (l2 && g.E <= kRegE) {
    // the echoes' ascended values stay in registers until the coupled norm is known, so p_m is
    // read and written once
    float2 pe[kRegE];
    float ss = 0.f;
#pragma unroll
    for (int e = 0; e < kRegE; ++e) {
      if (e < g.E) {
        const int64_t i = base + (int64_t)e * g.N;
        const float2 u0 = ubar[i];
        float2 d = make_float2(0.f, 0.f);
        if (b < g.B - 1) d = csub(ubar[i + EN], u0);
        const float2 p = p_m[i];
        pe[e] = make_float2(p.x + sigma * d.x, p.y + sigma * d.y);
        ss += cabs2(pe[e]);
      }
    }
    const float den = fmaxf(1.0f, sqrtf(ss) / lam);
#pragma unroll
    for (int e = 0; e < kRegE; ++e)
      if (e < g.E) p_m[base + (int64_t)e * g.N] = make_float2(pe[e].x / den, pe[e].y / den);
    return;
  }
  float ss = 0.f;
  for (int e = 0; e < g.E; ++e) {
    const int64_t i = base + (int64_t)e * g.N;
    const float2 u0 = ubar[i];
    float2 d = make_float2(0.f, 0.f);
    if (b < g.B - 1) d = csub(ubar[i + EN], u0);
    float2 p = p_m[i];
    p = make_float2(p.x + sigma * d.x, p.y + sigma * d.y);
    if (l2) {
      ss += cabs2(p);
      p_m[i] = p;
    } else {
      const float den = fmaxf(1.0f, sqrtf(cabs2(p)) / lam);
      p_m[i] = make_float2(p.x / den, p.y / den);
    }
  }
  if (l2) {
    const float den = fmaxf(1.0f, sqrtf(ss) / lam);
    for (int e = 0; e < g.E; ++e) {
      const int64_t i = base + (int64_t)e * g.N;
      const float2 p = p_m[i];
      p_m[i] = make_float2(p.x / den, p.y / den);
    }
  }
}

__global__ void __launch_bounds__(kEwBlock) k_tvm_dual(VolGeom g, const float2* ubar, float2* p_m,
                                                       float sigma, float lam, int l2) {
  if (gate_closed(g.gate)) return;
  const int64_t t = (int64_t)blockIdx.x * kEwBlock + threadIdx.x;
  if (t >= (int64_t)g.B * g.N) return;
  const int b = (int)(t / g.N);
  const int64_t v = t - (int64_t)b * g.N;
  motion_dual(g, b, v, ubar, p_m, sigma, lam, l2);
}

// TVME dual step (tvme.py:233-337): p_m as TVM, then the echo difference
// d_e = ubar[e+1] - ubar[e] (last echo 0), forward spatial differences of d_e
// (last index 0), ascent and the isotropic prox over (p_ex, p_ey, p_ez).
__global__ void __launch_bounds__(kEwBlock) k_tvme_dual(VolGeom g, const float2* ubar, float2* p_m,
                                                        float2* p_ex, float2* p_ey, float2* p_ez,
                                                        float sigma, float lam_m, float lam_e, int l2) {
  if (gate_closed(g.gate)) return;
  const int64_t t = (int64_t)blockIdx.x * kEwBlock + threadIdx.x;
  if (t >= (int64_t)g.B * g.N) return;
  const int b = (int)(t / g.N);
  const int64_t v = t - (int64_t)b * g.N;
  motion_dual(g, b, v, ubar, p_m, sigma, lam_m, l2);
  const int x = (int)(v % g.X);
  const int y = (int)((v / g.X) % g.Y);
  const int z = (int)(v / ((int64_t)g.X * g.Y));
  const int64_t sx = 1, sy = g.X, sz = (int64_t)g.X * g.Y;
  const bool hx = x < g.X - 1, hy = y < g.Y - 1, hz = z < g.Z - 1;
  for (int e = 0; e < g.E; ++e) {
    const int64_t i = ((int64_t)b * g.E + e) * g.N + v;
    float2 dx = make_float2(0.f, 0.f), dy = dx, dz = dx;
    if (e < g.E - 1) {
      const float2 d0 = csub(ubar[i + g.N], ubar[i]);
      if (hx) dx = csub(csub(ubar[i + g.N + sx], ubar[i + sx]), d0);
      if (hy) dy = csub(csub(ubar[i + g.N + sy], ubar[i + sy]), d0);
      if (hz) dz = csub(csub(ubar[i + g.N + sz], ubar[i + sz]), d0);
    }
    float2 px = p_ex[i], py = p_ey[i], pz = p_ez[i];
    px = make_float2(px.x + sigma * dx.x, px.y + sigma * dx.y);
    py = make_float2(py.x + sigma * dy.x, py.y + sigma * dy.y);
    pz = make_float2(pz.x + sigma * dz.x, pz.y + sigma * dz.y);
    const float mag = sqrtf(cabs2(px) + cabs2(py) + cabs2(pz));
    const float den = fmaxf(1.0f, mag / lam_e);
    p_ex[i] = make_float2(px.x / den, px.y / den);
    p_ey[i] = make_float2(py.x / den, py.y / den);
    p_ez[i] = make_float2(pz.x / den, pz.y / den);
  }
}

// Adjoint of the forward difference along one axis at index i of n
// (dxbc/dybc/dzbc, tvme.py:171-187): -q[0], q[i-1]-q[i], q[n-2]; 0 if n == 1.
__device__ __forceinline__ float2 bdiff(const float2* q, int64_t idx, int pos, int n, int64_t stride) {
  if (n == 1) return make_float2(0.f, 0.f);
  if (pos == 0) {
    const float2 a = q[idx];
    return make_float2(-a.x, -a.y);
  }
  if (pos == n - 1) return q[idx - stride];
  return csub(q[idx - stride], q[idx]);
}

// Motion divergence terms for voxel index i of bin b.
// TVM 'div' = -D_m^T p (tvm.py:342-357); TVME divp_m = +D_m^T p (tvme.py:421-439).
__device__ __forceinline__ float2 dmt(const VolGeom& g, const float2* p_m, int b, int64_t i) {
  const int64_t EN = (int64_t)g.E * g.N;
  if (g.B == 1) return make_float2(0.f, 0.f);
  if (b == 0) {
    const float2 a = p_m[i];
    return make_float2(-a.x, -a.y);
  }
  if (b == g.B - 1) return p_m[i - EN];
  return csub(p_m[i - EN], p_m[i]);
}

// TVM primal step: u <- u - tau (g - div), ubar <- 2u - u_old (tvm.py:342-379),
// with float64 partial sums of |u - u_old|^2 and |u|^2 (tvm.py:416-424).
__global__ void __launch_bounds__(kEwBlock) k_tvm_primal(VolGeom g, float2* u, float2* ubar,
                                                         const float2* grad, const float2* p_m,
                                                         float tau, double2* partial) {
  if (gate_closed(g.gate)) return;
  const int64_t t = (int64_t)blockIdx.x * kEwBlock + threadIdx.x;
  double num = 0.0, den = 0.0;
  if (t < (int64_t)g.B * g.N) {
    const int b = (int)(t / g.N);
    const int64_t v = t - (int64_t)b * g.N;
    for (int e = 0; e < g.E; ++e) {
      const int64_t i = ((int64_t)b * g.E + e) * g.N + v;
      const float2 dm = dmt(g, p_m, b, i);
      const float2 div = make_float2(-dm.x, -dm.y);
      const float2 x = csub(grad[i], div);
      const float2 uo = u[i];
      const float2 un = make_float2(uo.x - tau * x.x, uo.y - tau * x.y);
      u[i] = un;
      ubar[i] = make_float2(2.f * un.x - uo.x, 2.f * un.y - uo.y);
      const double ddx = (double)un.x - (double)uo.x, ddy = (double)un.y - (double)uo.y;
      num += ddx * ddx + ddy * ddy;
      den += (double)un.x * un.x + (double)un.y * un.y;
    }
  }
  block_reduce2(num, den, partial);
}

// TVME primal step (tvme.py:356-480): divp = sum of backward differences of
// (p_ex, p_ey, p_ez); divp_e = D_e^T divp; u -= tau divp_m; u -= tau divp_e;
// u -= tau g; ubar = 2u - u_old.  Float64 stopping partial sums.
__global__ void __launch_bounds__(kEwBlock) k_tvme_primal(VolGeom g, float2* u, float2* ubar,
                                                          const float2* grad, const float2* p_m,
                                                          const float2* p_ex, const float2* p_ey,
                                                          const float2* p_ez, float tau,
                                                          double2* partial) {
  if (gate_closed(g.gate)) return;
  const int64_t t = (int64_t)blockIdx.x * kEwBlock + threadIdx.x;
  double num = 0.0, den = 0.0;
  if (t < (int64_t)g.B * g.N) {
    const int b = (int)(t / g.N);
    const int64_t v = t - (int64_t)b * g.N;
    const int x = (int)(v % g.X);
    const int y = (int)((v / g.X) % g.Y);
    const int z = (int)(v / ((int64_t)g.X * g.Y));
    const int64_t sy = g.X, sz = (int64_t)g.X * g.Y;
    float2 prev = make_float2(0.f, 0.f);  // divp at echo e-1
    for (int e = 0; e < g.E; ++e) {
      const int64_t i = ((int64_t)b * g.E + e) * g.N + v;
      const float2 div = cadd(cadd(bdiff(p_ex, i, x, g.X, 1), bdiff(p_ey, i, y, g.Y, sy)),
                              bdiff(p_ez, i, z, g.Z, sz));
      float2 de;
      if (g.E == 1 || e == 0) {
        de = make_float2(-div.x, -div.y);
      } else if (e == g.E - 1) {
        de = prev;
      } else {
        de = csub(prev, div);
      }
      prev = div;
      const float2 dm = dmt(g, p_m, b, i);
      const float2 uo = u[i];
      float2 un = make_float2(uo.x - tau * dm.x, uo.y - tau * dm.y);
      un = make_float2(un.x - tau * de.x, un.y - tau * de.y);
      const float2 gr = grad[i];
      un = make_float2(un.x - tau * gr.x, un.y - tau * gr.y);
      u[i] = un;
      ubar[i] = make_float2(2.f * un.x - uo.x, 2.f * un.y - uo.y);
      const double ddx = (double)un.x - (double)uo.x, ddy = (double)un.y - (double)uo.y;
      num += ddx * ddx + ddy * ddy;
      den += (double)un.x * un.x + (double)un.y * un.y;
    }
  }
  block_reduce2(num, den, partial);
}

// TVMW dual step (tvmw.py:468-520): p_m as TVM (lambda1), then the db1 echo
// wavelet of ubar with global echo index (even e: (x_e + x_{e+1})/sqrt2 with
// x_E = 0, odd e: (x_e - x_{e-1})/sqrt2), ascent and p_w1 /= max(1, |p_w1|/lambda2)
// (p_w1 = 0 when lambda2 <= 0).
__global__ void __launch_bounds__(kEwBlock) k_tvmw_dual(VolGeom g, const float2* ubar, float2* p_m,
                                                        float2* p_w1, float sigma, float lam1,
                                                        float lam2, int l2) {
  if (gate_closed(g.gate)) return;
  const int64_t t = (int64_t)blockIdx.x * kEwBlock + threadIdx.x;
  if (t >= (int64_t)g.B * g.N) return;
  const int b = (int)(t / g.N);
  const int64_t v = t - (int64_t)b * g.N;
  motion_dual(g, b, v, ubar, p_m, sigma, lam1, l2);
  const float s = 0.70710678118654752440f;
  for (int e = 0; e < g.E; ++e) {
    const int64_t i = ((int64_t)b * g.E + e) * g.N + v;
    float2 w;
    if ((e & 1) == 0) {
      const float2 x1 = (e + 1 < g.E) ? ubar[i + g.N] : make_float2(0.f, 0.f);
      w = cscale(cadd(ubar[i], x1), s);
    } else {
      w = cscale(csub(ubar[i], ubar[i - g.N]), s);
    }
    float2 p = p_w1[i];
    p = make_float2(p.x + sigma * w.x, p.y + sigma * w.y);
    if (lam2 > 0.f) {
      const float den = fmaxf(1.0f, sqrtf(cabs2(p)) / lam2);
      p = make_float2(p.x / den, p.y / den);
    } else {
      p = make_float2(0.f, 0.f);
    }
    p_w1[i] = p;
  }
}

// TVMW primal step (tvmw.py:661-684): divp_m[b] = p[b] - p[b-1], divp_m[0] = p[0];
// W1^H p_w1 (even e: (a_e - d_{e+1})/sqrt2 with d_E = 0, odd e: (a_{e-1} + d_e)/sqrt2);
// u <- u - tau ((g - divp_m + divp_w1) + divp_w2); ubar <- 2u - u_old.  Stopping
// sums (tvmw.py:723-769) in float64: sum |ubar_prev - ubar|^2 and sum |ubar_prev|^2.
__global__ void __launch_bounds__(kEwBlock) k_tvmw_primal(VolGeom g, float2* u, float2* ubar,
                                                          const float2* grad, const float2* p_m,
                                                          const float2* p_w1, const float2* divw2,
                                                          float tau, double2* partial) {
  if (gate_closed(g.gate)) return;
  const int64_t t = (int64_t)blockIdx.x * kEwBlock + threadIdx.x;
  double num = 0.0, den = 0.0;
  if (t < (int64_t)g.B * g.N) {
    const int b = (int)(t / g.N);
    const int64_t v = t - (int64_t)b * g.N;
    const int64_t EN = (int64_t)g.E * g.N;
    const float s = 0.70710678118654752440f;
    for (int e = 0; e < g.E; ++e) {
      const int64_t i = ((int64_t)b * g.E + e) * g.N + v;
      const float2 dm = (b == 0) ? p_m[i] : csub(p_m[i], p_m[i - EN]);
      float2 dw1;
      if ((e & 1) == 0) {
        const float2 dn = (e + 1 < g.E) ? p_w1[i + g.N] : make_float2(0.f, 0.f);
        dw1 = cscale(csub(p_w1[i], dn), s);
      } else {
        dw1 = cscale(cadd(p_w1[i - g.N], p_w1[i]), s);
      }
      const float2 x = cadd(cadd(csub(grad[i], dm), dw1), divw2[i]);
      const float2 uo = u[i];
      const float2 un = make_float2(uo.x - tau * x.x, uo.y - tau * x.y);
      const float2 bo = ubar[i];
      const float2 bn = make_float2(2.f * un.x - uo.x, 2.f * un.y - uo.y);
      u[i] = un;
      ubar[i] = bn;
      const double ddx = (double)bo.x - (double)bn.x, ddy = (double)bo.y - (double)bn.y;
      num += ddx * ddx + ddy * ddy;
      den += (double)bo.x * bo.x + (double)bo.y * bo.y;
    }
  }
  block_reduce2(num, den, partial);
}

// Fixed-order final reduction of the per-block partial sums (one block).
__global__ void __launch_bounds__(kEwBlock) k_cscale(const float2* x, int64_t n, float s, float2* out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = make_float2(s * x[i].x, s * x[i].y);
}

// Conjugate-gradient vector steps for the data-consistency refinement of a warm
// start: x += alpha p, r -= alpha Ap with per-block float64 sums of |r|^2; p = r +
// beta p; per-block float64 sums of Re(conj(a) b).  Block sums are added in block
// order on the host, so every result is reproducible.
__device__ __forceinline__ void block_sum1(double v, double* partial) {
  __shared__ double sh[kEwBlock / 64];
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int w = 0; w < kEwBlock / 64; ++w) t += sh[w];
    partial[blockIdx.x] = t;
  }
}

__global__ void __launch_bounds__(kEwBlock) k_cg_update(float2* x, float2* r, const float2* p, const float2* ap,
                                                        int64_t n, float alpha, double* partial) {
  double acc = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const float2 pv = p[i], av = ap[i];
    float2 xv = x[i], rv = r[i];
    xv = make_float2(xv.x + alpha * pv.x, xv.y + alpha * pv.y);
    rv = make_float2(rv.x - alpha * av.x, rv.y - alpha * av.y);
    x[i] = xv;
    r[i] = rv;
    acc += (double)rv.x * rv.x + (double)rv.y * rv.y;
  }
  block_sum1(acc, partial);
}

__global__ void __launch_bounds__(kEwBlock) k_cg_direction(float2* p, const float2* r, int64_t n, float beta) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    const float2 pv = p[i], rv = r[i];
    p[i] = make_float2(rv.x + beta * pv.x, rv.y + beta * pv.y);
  }
}

__global__ void __launch_bounds__(kEwBlock) k_cdot_re(const float2* a, const float2* b, int64_t n, double* partial) {
  double acc = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    acc += (double)a[i].x * b[i].x + (double)a[i].y * b[i].y;
  block_sum1(acc, partial);
}

__global__ void __launch_bounds__(kEwBlock) k_reduce_partials(const double2* partial, int n, double* out,
                                                              const int* gate) {
  if (gate_closed(gate)) return;
  double a = 0.0, b = 0.0;
  // 8 partials per thread requested together; the additions keep their order (same bits)
#pragma unroll 8
  for (int i = threadIdx.x; i < n; i += kEwBlock) {
    a += partial[i].x;
    b += partial[i].y;
  }
  __shared__ double sa[kEwBlock], sb[kEwBlock];
  sa[threadIdx.x] = a;
  sb[threadIdx.x] = b;
  __syncthreads();
  for (int s = kEwBlock / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
      sa[threadIdx.x] += sa[threadIdx.x + s];
      sb[threadIdx.x] += sb[threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out[0] = sa[0];
    out[1] = sb[0];
  }
}

static int check_geom(int B, int E, int Z, int Y, int X, VolGeom* g) {
  if (B < 1 || E < 1 || Z < 1 || Y < 1 || X < 1) TR_FAIL(kInvalidArgument, "shape must be positive");
  g->B = B;
  g->E = E;
  g->Z = Z;
  g->Y = Y;
  g->X = X;
  g->N = (int64_t)Z * Y * X;
  g->gate = gate_current();
  return kOk;
}

static unsigned ew_blocks(const VolGeom& g) {
  return (unsigned)(((int64_t)g.B * g.N + kEwBlock - 1) / kEwBlock);
}

// Stopping-sum partials of tr_tvme_primal_noalloc: one float64 pair per 256
// (bin, voxel) columns, so 2^18 pairs (4 MiB per GPU, held until the process
// ends) cover B * Z * Y * X up to 64 Mi.  Every partial the reduction reads is
// written by the primal kernel of the same call, so nothing needs clearing.
constexpr int64_t kScratchPairs = (int64_t)1 << 18;
constexpr int kMaxDevices = 64;
static __device__ __align__(256) double2 scratch[kScratchPairs];

// Device address of `scratch` on the current device, looked up once per device
// (the symbol's host-side name is not a device address).
static int scratch_address(double2** out) {
  static std::mutex mu;
  static double2* addr[kMaxDevices] = {};
  int dev = 0;
  TR_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) TR_FAIL(kUnsupported, "device index out of range");
  std::lock_guard<std::mutex> lk(mu);
  if (!addr[dev]) {
    void* p = nullptr;
    TR_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    addr[dev] = (double2*)p;
  }
  *out = addr[dev];
  return kOk;
}

}  // namespace tr

using namespace tr;

extern "C" {

int64_t tr_partial_count(int B, int Z, int Y, int X) {
  return ((int64_t)B * Z * Y * X + kEwBlock - 1) / kEwBlock;
}

int tr_tvm_dual(int B, int E, int Z, int Y, int X, const tr_c64* ubar, tr_c64* p_m, float sigma,
                float lam, int l2_coupling, void* stream) {
  VolGeom g;
  if (int r = check_geom(B, E, Z, Y, X, &g)) return r;
  ProfScope ps(kKTvmDual, 3.0 * 8.0 * g.B * g.E * g.N, (hipStream_t)stream);
  hipLaunchKernelGGL(k_tvm_dual, dim3(ew_blocks(g)), dim3(kEwBlock), 0, (hipStream_t)stream, g, F2(ubar),
                     F2(p_m), sigma, lam, l2_coupling);
  TR_LAUNCH_CHECK();
  return kOk;
}

int tr_tvm_primal(int B, int E, int Z, int Y, int X, tr_c64* u, tr_c64* ubar, const tr_c64* grad,
                  const tr_c64* p_m, float tau, double* partial, void* stream) {
  VolGeom g;
  if (int r = check_geom(B, E, Z, Y, X, &g)) return r;
  ProfScope ps(kKTvmPrimal, 5.0 * 8.0 * g.B * g.E * g.N, (hipStream_t)stream);
  hipLaunchKernelGGL(k_tvm_primal, dim3(ew_blocks(g)), dim3(kEwBlock), 0, (hipStream_t)stream, g, F2(u),
                     F2(ubar), F2(grad), F2(p_m), tau, (double2*)partial);
  TR_LAUNCH_CHECK();
  return kOk;
}

int tr_tvme_dual(int B, int E, int Z, int Y, int X, const tr_c64* ubar, tr_c64* p_m, tr_c64* p_ex,
                 tr_c64* p_ey, tr_c64* p_ez, float sigma, float lam_m, float lam_e, int l2_coupling,
                 void* stream) {
  VolGeom g;
  if (int r = check_geom(B, E, Z, Y, X, &g)) return r;
  ProfScope ps(kKTvmeDual, 9.0 * 8.0 * g.B * g.E * g.N, (hipStream_t)stream);
  hipLaunchKernelGGL(k_tvme_dual, dim3(ew_blocks(g)), dim3(kEwBlock), 0, (hipStream_t)stream, g, F2(ubar),
                     F2(p_m), F2(p_ex), F2(p_ey), F2(p_ez), sigma, lam_m, lam_e, l2_coupling);
  TR_LAUNCH_CHECK();
  return kOk;
}

int tr_tvme_primal(int B, int E, int Z, int Y, int X, tr_c64* u, tr_c64* ubar, const tr_c64* grad,
                   const tr_c64* p_m, const tr_c64* p_ex, const tr_c64* p_ey, const tr_c64* p_ez,
                   float tau, double* partial, void* stream) {
  VolGeom g;
  if (int r = check_geom(B, E, Z, Y, X, &g)) return r;
  ProfScope ps(kKTvmePrimal, 8.0 * 8.0 * g.B * g.E * g.N, (hipStream_t)stream);
  hipLaunchKernelGGL(k_tvme_primal, dim3(ew_blocks(g)), dim3(kEwBlock), 0, (hipStream_t)stream, g, F2(u),
                     F2(ubar), F2(grad), F2(p_m), F2(p_ex), F2(p_ey), F2(p_ez), tau, (double2*)partial);
  TR_LAUNCH_CHECK();
  return kOk;
}

int tr_tvmw_dual(int B, int E, int Z, int Y, int X, const tr_c64* ubar, tr_c64* p_m, tr_c64* p_w1,
                 float sigma, float lam1, float lam2, int l2_coupling, void* stream) {
  VolGeom g;
  if (int r = check_geom(B, E, Z, Y, X, &g)) return r;
  ProfScope ps(kKTvmwDual, 5.0 * 8.0 * g.B * g.E * g.N, (hipStream_t)stream);
  hipLaunchKernelGGL(k_tvmw_dual, dim3(ew_blocks(g)), dim3(kEwBlock), 0, (hipStream_t)stream, g, F2(ubar),
                     F2(p_m), F2(p_w1), sigma, lam1, lam2, l2_coupling);
  TR_LAUNCH_CHECK();
  return kOk;
}

int tr_tvmw_primal(int B, int E, int Z, int Y, int X, tr_c64* u, tr_c64* ubar, const tr_c64* grad,
                   const tr_c64* p_m, const tr_c64* p_w1, const tr_c64* divw2, float tau, double* partial,
                   void* stream) {
  VolGeom g;
  if (int r = check_geom(B, E, Z, Y, X, &g)) return r;
  ProfScope ps(kKTvmwPrimal, 9.0 * 8.0 * g.B * g.E * g.N, (hipStream_t)stream);
  hipLaunchKernelGGL(k_tvmw_primal, dim3(ew_blocks(g)), dim3(kEwBlock), 0, (hipStream_t)stream, g, F2(u),
                     F2(ubar), F2(grad), F2(p_m), F2(p_w1), F2(divw2), tau, (double2*)partial);
  TR_LAUNCH_CHECK();
  return kOk;
}

int tr_cg_update(tr_c64* x, tr_c64* r, const tr_c64* p, const tr_c64* ap, int64_t n, double alpha,
                 double* partial, int nblocks, void* stream) {
  if (!x || !r || !p || !ap || !partial) TR_FAIL(kInvalidArgument, "null argument");
  if (nblocks < 1 || nblocks > 65536) TR_FAIL(kInvalidArgument, "nblocks outside [1, 65536]");
  hipLaunchKernelGGL(k_cg_update, dim3(nblocks), dim3(kEwBlock), 0, (hipStream_t)stream, (float2*)x, (float2*)r,
                     (const float2*)p, (const float2*)ap, n, (float)alpha, partial);
  TR_LAUNCH_CHECK();
  return kOk;
}

int tr_cg_direction(tr_c64* p, const tr_c64* r, int64_t n, double beta, void* stream) {
  if ((!p || !r) && n > 0) TR_FAIL(kInvalidArgument, "null argument");
  if (n == 0) return kOk;
  hipLaunchKernelGGL(k_cg_direction, dim3((unsigned)((n + kEwBlock - 1) / kEwBlock)), dim3(kEwBlock), 0,
                     (hipStream_t)stream, (float2*)p, (const float2*)r, n, (float)beta);
  TR_LAUNCH_CHECK();
  return kOk;
}

int tr_cdot_re(const tr_c64* a, const tr_c64* b, int64_t n, double* partial, int nblocks, void* stream) {
  if (!a || !b || !partial) TR_FAIL(kInvalidArgument, "null argument");
  if (nblocks < 1 || nblocks > 65536) TR_FAIL(kInvalidArgument, "nblocks outside [1, 65536]");
  hipLaunchKernelGGL(k_cdot_re, dim3(nblocks), dim3(kEwBlock), 0, (hipStream_t)stream, (const float2*)a,
                     (const float2*)b, n, partial);
  TR_LAUNCH_CHECK();
  return kOk;
}

int tr_cscale(const tr_c64* x, int64_t n, double s, tr_c64* out, void* stream) {
  if ((!x || !out) && n > 0) TR_FAIL(kInvalidArgument, "null argument");
  if (n == 0) return kOk;
  hipLaunchKernelGGL(k_cscale, dim3((unsigned)((n + kEwBlock - 1) / kEwBlock)), dim3(kEwBlock), 0,
                     (hipStream_t)stream, (const float2*)x, n, (float)s, (float2*)out);
  TR_LAUNCH_CHECK();
  return kOk;
}

int tr_tvme_primal_noalloc(int B, int E, int Z, int Y, int X, tr_c64* u, tr_c64* ubar, const tr_c64* grad,
                           const tr_c64* p_m, const tr_c64* p_ex, const tr_c64* p_ey, const tr_c64* p_ez,
                           float tau, double* sums, void* stream) {
  VolGeom g;
  if (int r = check_geom(B, E, Z, Y, X, &g)) return r;
  const int64_t n = ((int64_t)g.B * g.N + kEwBlock - 1) / kEwBlock;
  if (n > kScratchPairs)
    TR_FAIL(kInvalidArgument, "B*Z*Y*X = " + std::to_string((int64_t)g.B * g.N) +
                                  " exceeds the no-allocation limit of " +
                                  std::to_string(kScratchPairs * kEwBlock) + " (bin, voxel) columns");
  if (!sums) TR_FAIL(kInvalidArgument, "null sums");
  hipStream_t s = (hipStream_t)stream;
  int sdev = 0, cur = 0;
  TR_HIP(hipGetDevice(&cur));
  if (s) TR_HIP(hipStreamGetDevice(s, &sdev));
  if (s && sdev != cur) TR_FAIL(kInvalidArgument, "the stream's device must be the current device");
  double2* part = nullptr;
  if (int r = scratch_address(&part)) return r;
  {
    ProfScope ps(kKTvmePrimal, 8.0 * 8.0 * g.B * g.E * g.N, s);
    hipLaunchKernelGGL(k_tvme_primal, dim3(ew_blocks(g)), dim3(kEwBlock), 0, s, g, F2(u), F2(ubar), F2(grad),
                       F2(p_m), F2(p_ex), F2(p_ey), F2(p_ez), tau, part);
    TR_LAUNCH_CHECK();
  }
  ProfScope ps(kKReduce, 16.0 * n + 16.0, s);
  hipLaunchKernelGGL(k_reduce_partials, dim3(1), dim3(kEwBlock), 0, s, (const double2*)part, (int)n, sums,
                     gate_current());
  TR_LAUNCH_CHECK();
  return kOk;
}

int tr_reduce_partials(const double* partial, int64_t n, double* out2, void* stream) {
  if (n < 1 || n > INT32_MAX) TR_FAIL(kInvalidArgument, "bad partial count");
  ProfScope ps(kKReduce, 16.0 * n + 16.0, (hipStream_t)stream);
  hipLaunchKernelGGL(k_reduce_partials, dim3(1), dim3(kEwBlock), 0, (hipStream_t)stream,
                     (const double2*)partial, (int)n, out2, gate_current());
  TR_LAUNCH_CHECK();
  return kOk;
}

}  // extern "C"
