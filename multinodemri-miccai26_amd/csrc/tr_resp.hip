// Respiratory self-gating: the zero-phase Butterworth band-pass of every
// (echo, coil) centre trace, in float64 with the operation order of SciPy's
// sosfiltfilt (odd extension of 3 * ntaps samples, a forward and a backward
// second-order-section pass started from the steady state zi scaled by the
// first sample), low-pass then high-pass, as the reference's respiratory.py:
// 86-92 applies them.  One thread per trace: the recurrence is sequential in
// time, the traces are independent.

#include "toporecon_hip.h"
#include "tr_common.hpp"

namespace tr {

constexpr int kRespMaxSections = 16;

struct RespFilter {
  const double* sos;  // (n, 6): b0 b1 b2 a0 a1 a2 with a0 = 1
  const double* zi;   // (n, 2): steady state of a unit step
  int n;
  int pad;            // samples of odd extension at each end
};

// One pass of the cascade over buf[0 .. L) in place, forwards or backwards, the state starting at
// zi * (first sample of the pass).
__device__ void sos_pass(const RespFilter& f, double* buf, int64_t L, bool backward) {
  // products and sums rounded one by one (no fused multiply-add), as in SciPy's compiled loop
#pragma clang fp contract(off)
  double z0[kRespMaxSections], z1[kRespMaxSections];
  const double x0 = backward ? buf[L - 1] : buf[0];
  for (int s = 0; s < f.n; ++s) {
    z0[s] = f.zi[2 * s] * x0;
    z1[s] = f.zi[2 * s + 1] * x0;
  }
  for (int64_t j = 0; j < L; ++j) {
    const int64_t n = backward ? L - 1 - j : j;
    double xc = buf[n];
    for (int s = 0; s < f.n; ++s) {
      const double* c = f.sos + 6 * s;
      const double xn = c[0] * xc + z0[s];
      z0[s] = c[1] * xc - c[4] * xn + z1[s];
      z1[s] = c[2] * xc - c[5] * xn;
      xc = xn;
    }
    buf[n] = xc;
  }
}

// ext = [2 x[0] - x[pad .. 1], x, 2 x[T-1] - x[T-2 .. T-1-pad]] in the input's precision T_IN,
// stored as float64, then the forward and backward passes; the filtered trace is ext[pad .. pad+T).
template <typename T_IN>
__device__ void filtfilt(const RespFilter& f, const T_IN* x, int64_t T, double* ext) {
#pragma clang fp contract(off)
  const int e = f.pad;
  const T_IN two = (T_IN)2;
  for (int j = 0; j < e; ++j) ext[j] = (double)(two * x[0] - x[e - j]);
  for (int64_t t = 0; t < T; ++t) ext[e + t] = (double)x[t];
  for (int j = 0; j < e; ++j) ext[e + T + j] = (double)(two * x[T - 1] - x[T - 2 - j]);
  const int64_t L = T + 2 * (int64_t)e;
  sos_pass(f, ext, L, false);
  sos_pass(f, ext, L, true);
}

__global__ void k_resp_bandpass(const float* traces, int64_t ntraces, int64_t T, RespFilter lo, RespFilter hi,
                                double* work, int64_t wlen, float* out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ntraces) return;
  double* a = work + i * 2 * wlen;
  double* b = a + wlen;
  filtfilt<float>(lo, traces + i * T, T, a);          // float32 input: its extension in float32
  filtfilt<double>(hi, a + lo.pad, T, b);              // float64 input
  float* o = out + i * T;
  for (int64_t t = 0; t < T; ++t) o[t] = (float)b[hi.pad + t];
}

static int check_filter(const RespFilter& f, int64_t T, const char* name) {
  if (!f.sos || !f.zi) TR_FAIL(kInvalidArgument, std::string(name) + ": null coefficients");
  if (f.n < 1 || f.n > kRespMaxSections)
    TR_FAIL(kInvalidArgument, std::string(name) + ": sections outside [1, " + std::to_string(kRespMaxSections) + "]");
  if (f.pad < 0 || f.pad >= T)
    TR_FAIL(kInvalidArgument, std::string(name) + ": the trace must be longer than the padding");
  return kOk;
}

}  // namespace tr

using namespace tr;

extern "C" {

int64_t tr_resp_work_count(int64_t ntraces, int64_t T, int pad_lo, int pad_hi) {
  const int64_t wlen = T + 2 * (int64_t)(pad_lo > pad_hi ? pad_lo : pad_hi);
  return ntraces * 2 * wlen;
}

int tr_resp_bandpass(const float* traces, int64_t ntraces, int64_t T, const double* sos_lo, const double* zi_lo,
                     int n_lo, int pad_lo, const double* sos_hi, const double* zi_hi, int n_hi, int pad_hi,
                     double* work, float* out, void* stream) {
  if (ntraces < 0 || T < 2) TR_FAIL(kInvalidArgument, "need traces of at least 2 samples");
  if (!traces || !work || !out) TR_FAIL(kInvalidArgument, "null argument");
  RespFilter lo{sos_lo, zi_lo, n_lo, pad_lo}, hi{sos_hi, zi_hi, n_hi, pad_hi};
  if (int r = check_filter(lo, T, "low-pass")) return r;
  if (int r = check_filter(hi, T, "high-pass")) return r;
  if (ntraces == 0) return kOk;
  const int64_t wlen = T + 2 * (int64_t)(pad_lo > pad_hi ? pad_lo : pad_hi);
  constexpr int kThreads = 64;
  hipLaunchKernelGGL(k_resp_bandpass, dim3((unsigned)((ntraces + kThreads - 1) / kThreads)), dim3(kThreads), 0,
                     (hipStream_t)stream, traces, ntraces, T, lo, hi, work, wlen, out);
  TR_LAUNCH_CHECK();
  return kOk;
}

}  // extern "C"
