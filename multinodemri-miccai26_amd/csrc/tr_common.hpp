// Shared host/device helpers for the toporecon HIP library (gfx950).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>

namespace tr {

// ---------------------------------------------------------------------------
// Error reporting: every C entry point returns 0 or a negative status and
// leaves a message for tr_last_error().
// ---------------------------------------------------------------------------
enum Status : int {
  kOk = 0,
  kInvalidArgument = -1,
  kHipError = -2,
  kUnsupported = -3,
  kOutOfMemory = -4,
};

void set_error(const std::string& msg);
const char* last_error();

#define TR_FAIL(code, msg)                                    \
  do {                                                        \
    ::tr::set_error(std::string(__func__) + ": " + (msg));    \
    return (code);                                            \
  } while (0)

#define TR_HIP(call)                                                              \
  do {                                                                            \
    hipError_t _e = (call);                                                       \
    if (_e != hipSuccess) {                                                       \
      ::tr::set_error(std::string(__func__) + ": " #call " failed: " +            \
                      hipGetErrorString(_e));                                     \
      return (_e == hipErrorOutOfMemory) ? ::tr::kOutOfMemory : ::tr::kHipError;  \
    }                                                                             \
  } while (0)

#define TR_LAUNCH_CHECK() TR_HIP(hipGetLastError())

// ---------------------------------------------------------------------------
// Gates of batched solves: a host thread may name an int in device memory
// (tr_gate_set); every kernel of the solver family that thread launches
// afterwards returns at once while that int is nonzero, so a captured graph
// of several iterations leaves an instance's state untouched once its own
// stopping rule has closed the gate.  Null = no gate.
// ---------------------------------------------------------------------------
// Workgroups are handed to the 8 XCDs round-robin (block b runs on XCD b % 8).
// With chunk S > 0 the logical block order is permuted so that every XCD
// receives runs of S consecutive logical blocks, which then share its L2; the
// tail that does not fill whole 8 * S groups keeps the identity order.
constexpr unsigned kNumXcd = 8;
__device__ __forceinline__ unsigned xcd_block(unsigned b, unsigned nb, unsigned S) {
  if (S == 0) return b;
  const unsigned full = nb / (kNumXcd * S) * (kNumXcd * S);
  if (b >= full) return b;
  const unsigned r = b / kNumXcd, x = b % kNumXcd;
  return (r / S) * (kNumXcd * S) + x * S + (r % S);
}

const int* gate_current();
__device__ __forceinline__ bool gate_closed(const int* g) { return g != nullptr && *g != 0; }

// ---------------------------------------------------------------------------
// Optional per-kernel timing: when enabled, every launch is bracketed by HIP
// events on its own stream and credited with its algorithmic bytes (each
// input read once, each output written once).
// ---------------------------------------------------------------------------
enum KernelId : int {
  kKFwdXPad = 0,
  kKFftYFwd,
  kKFftZFwd,
  kKInterp,
  kKSpread,
  kKFftZInv,
  kKFftYInv,
  kKAdjXCrop,
  kKTvmDual,
  kKTvmPrimal,
  kKTvmeDual,
  kKTvmePrimal,
  kKReduce,
  kKW2Forward,
  kKW2Adjoint,
  kKW2Dual,
  kKTvmwDual,
  kKTvmwPrimal,
  kKJsConv,
  kKJsConvAdjImg,
  kKJsConvAdjKer,
  kKJsAxis,
  kKJsRss,
  kNumKernelIds
};

const char* kernel_name(int id);

class ProfScope {
 public:
  ProfScope(int id, double bytes, hipStream_t s);
  ~ProfScope();

 private:
  int id_;
  double bytes_;
  hipStream_t s_;
  void* start_;
};

// ---------------------------------------------------------------------------
// Complex helpers (float2 = interleaved complex64).
// ---------------------------------------------------------------------------
__host__ __device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__host__ __device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__host__ __device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// conj(a) * b
__host__ __device__ __forceinline__ float2 cmulc(float2 a, float2 b) {
  return make_float2(a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x);
}
__host__ __device__ __forceinline__ float2 cscale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }
__host__ __device__ __forceinline__ float2 cconj(float2 a) { return make_float2(a.x, -a.y); }
__host__ __device__ __forceinline__ float cabs2(float2 a) { return a.x * a.x + a.y * a.y; }

// The C ABI passes complex64 as tr_c64 (same layout as float2).
template <typename T>
__host__ __device__ __forceinline__ const float2* F2(const T* p) { return reinterpret_cast<const float2*>(p); }
template <typename T>
__host__ __device__ __forceinline__ float2* F2(T* p) { return reinterpret_cast<float2*>(p); }

__host__ __device__ __forceinline__ int pmod(int a, int n) {
  int r = a % n;
  return r < 0 ? r + n : r;
}

// ---------------------------------------------------------------------------
// Kaiser-Bessel kernel of SigPy's nufft: I0(beta * sqrt(1 - x^2)) for |x| <= 1.
// I0 by the Abramowitz & Stegun 9.8.1 / 9.8.2 polynomials, float32.
// ---------------------------------------------------------------------------
__device__ __forceinline__ float kb_weight(float x, float beta) {
  if (fabsf(x) > 1.0f) return 0.0f;
  const float z = beta * sqrtf(1.0f - x * x);
  const float t = z * (1.0f / 3.75f);
  if (z < 3.75f) {
    const float t2 = t * t;
    return 1.0f + t2 * (3.5156229f + t2 * (3.0899424f + t2 * (1.2067492f +
           t2 * (0.2659732f + t2 * (0.0360768f + t2 * 0.0045813f)))));
  }
  const float it = 1.0f / t;
  const float p = 0.39894228f + it * (0.01328592f + it * (0.00225319f + it * (-0.00157565f +
                  it * (0.00916281f + it * (-0.02057706f + it * (0.02635537f +
                  it * (-0.01647633f + it * 0.00392377f)))))));
  return p * expf(z) * rsqrtf(z);
}

// Taps of one axis for a scaled coordinate kp (SigPy units already scaled to
// the oversampled grid): integer taps ceil(kp - 2) .. floor(kp + 2) evaluated
// in double like SigPy's kernels, weights in float, FFT-natural positions
// (t - g//2) mod g.  Returns the tap count (4, or 5 for an integral kp).
__device__ __forceinline__ int axis_taps(float kp, int g, float beta, float scale, float w[5], int pos[5]) {
  if (!isfinite(kp) || fabsf(kp) > 1.0e6f) {  // no taps for a non-finite / absurd coordinate
#pragma unroll
    for (int d = 0; d < 5; ++d) {
      w[d] = 0.0f;
      pos[d] = 0;
    }
    return 0;
  }
  const double kd = (double)kp;
  const int t0 = (int)ceil(kd - 2.0);
  const int t1 = (int)floor(kd + 2.0);
  const int n = t1 - t0 + 1;
  const int half = g / 2;
#pragma unroll
  for (int d = 0; d < 5; ++d) {
    const int t = t0 + d;
    const float x = ((float)t - kp) * 0.5f;
    w[d] = (d < n) ? kb_weight(x, beta) * scale : 0.0f;
    pos[d] = pmod(t - half, g);
  }
  return n;
}

}  // namespace tr
