// Error reporting, per-kernel timing and version entry points.

#include <mutex>

#include "tr_internal.hpp"

namespace tr {

static thread_local std::string g_last_error;
void set_error(const std::string& msg) { g_last_error = msg; }
const char* last_error() { return g_last_error.c_str(); }

static thread_local const int* g_gate = nullptr;
const int* gate_current() { return g_gate; }

// The stopping rule of one instance on the device (one thread).
__global__ void k_stop_rule(const double* sums, double tol, int max_iter, int rule, int* gate, int* its,
                            double* hist, int64_t* stamp) {
  if (*gate) return;
  const double num = sums[0], den = sums[1];
  double v = HUGE_VAL;
  if (rule == 0) {
    if (isfinite(num) && isfinite(den) && den > 0.0) v = sqrt(fmax(num, 0.0)) / sqrt(den);
  } else if (isfinite(num) && isfinite(den)) {
    v = sqrt(fmax(num, 0.0)) / fmax(sqrt(fmax(den, 0.0)), 1e-12);
  }
  const int c = *its + 1;
  *its = c;
  hist[c - 1] = v;
  if (v < tol || c >= max_iter) {
    *gate = 1;
    if (stamp) *stamp = (int64_t)wall_clock64();
  }
}

__global__ void k_clock_stamp(int64_t* out) { *out = (int64_t)wall_clock64(); }

// ---- kernel timing -----------------------------------------------------------
namespace {
struct ProfPending {
  hipEvent_t a, b;
  int id;
  double bytes;
};
std::mutex g_prof_mu;
bool g_prof_on = false;
std::vector<hipEvent_t> g_prof_pool;
std::vector<ProfPending> g_prof_pending;
double g_prof_ms[kNumKernelIds];
double g_prof_bytes[kNumKernelIds];
int64_t g_prof_count[kNumKernelIds];

hipEvent_t pool_get() {
  if (!g_prof_pool.empty()) {
    hipEvent_t e = g_prof_pool.back();
    g_prof_pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  (void)hipEventCreate(&e);
  return e;
}

void prof_drain() {
  for (auto& p : g_prof_pending) {
    float ms = 0.f;
    if (hipEventSynchronize(p.b) == hipSuccess && hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess) {
      g_prof_ms[p.id] += ms;
      g_prof_bytes[p.id] += p.bytes;
      g_prof_count[p.id] += 1;
    }
    g_prof_pool.push_back(p.a);
    g_prof_pool.push_back(p.b);
  }
  g_prof_pending.clear();
}
}  // namespace

const char* kernel_name(int id) {
  static const char* names[kNumKernelIds] = {
      "fwd_x_pad", "fft_y_fwd", "fft_z_fwd", "interp", "spread", "fft_z_inv", "fft_y_inv",
      "adj_x_crop", "tvm_dual", "tvm_primal", "tvme_dual", "tvme_primal", "reduce",
      "w2_forward", "w2_adjoint", "w2_dual", "tvmw_dual", "tvmw_primal", "js_conv", "js_conv_adj_img",
      "js_conv_adj_ker", "js_axis", "js_rss"};
  return (id >= 0 && id < kNumKernelIds) ? names[id] : "";
}

ProfScope::ProfScope(int id, double bytes, hipStream_t s) : id_(id), bytes_(bytes), s_(s), start_(nullptr) {
  if (!g_prof_on) return;
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing(s, &cs) != hipSuccess || cs != hipStreamCaptureStatusNone) return;  // not in graphs
  std::lock_guard<std::mutex> lk(g_prof_mu);
  hipEvent_t e = pool_get();
  if (e && hipEventRecord(e, s) == hipSuccess) start_ = e;
}

ProfScope::~ProfScope() {
  if (!start_) return;
  std::lock_guard<std::mutex> lk(g_prof_mu);
  hipEvent_t e = pool_get();
  if (e && hipEventRecord(e, s_) == hipSuccess) {
    g_prof_pending.push_back({(hipEvent_t)start_, e, id_, bytes_});
  }
}

}  // namespace tr

using namespace tr;

extern "C" {

const char* tr_last_error(void) { return tr::last_error(); }

int tr_version(void) { return TR_ABI_VERSION; }

int tr_profile_enable(int on) {
  std::lock_guard<std::mutex> lk(g_prof_mu);
  prof_drain();
  for (int i = 0; i < kNumKernelIds; ++i) {
    g_prof_ms[i] = 0.0;
    g_prof_bytes[i] = 0.0;
    g_prof_count[i] = 0;
  }
  g_prof_on = on != 0;
  return kOk;
}

int tr_profile_num_kernels(void) { return kNumKernelIds; }

const char* tr_profile_kernel_name(int id) { return kernel_name(id); }

int tr_profile_read(int id, double* ms_total, int64_t* launches, double* bytes_total) {
  if (id < 0 || id >= kNumKernelIds) TR_FAIL(kInvalidArgument, "bad kernel id");
  std::lock_guard<std::mutex> lk(g_prof_mu);
  prof_drain();
  if (ms_total) *ms_total = g_prof_ms[id];
  if (launches) *launches = g_prof_count[id];
  if (bytes_total) *bytes_total = g_prof_bytes[id];
  return kOk;
}

struct tr_graph {
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

int tr_graph_begin(void* stream) {
  TR_HIP(hipStreamBeginCapture((hipStream_t)stream, hipStreamCaptureModeThreadLocal));
  return kOk;
}

int tr_graph_end(void* stream, tr_graph** out) {
  if (!out) TR_FAIL(kInvalidArgument, "null output pointer");
  *out = nullptr;
  hipGraph_t g = nullptr;
  TR_HIP(hipStreamEndCapture((hipStream_t)stream, &g));
  tr_graph* t = new tr_graph();
  t->graph = g;
  const hipError_t e = hipGraphInstantiate(&t->exec, g, nullptr, nullptr, 0);
  if (e != hipSuccess) {
    (void)hipGraphDestroy(g);
    delete t;
    TR_FAIL(kHipError, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
  }
  *out = t;
  return kOk;
}

int tr_graph_launch(tr_graph* g, void* stream) {
  if (!g) TR_FAIL(kInvalidArgument, "null graph");
  TR_HIP(hipGraphLaunch(g->exec, (hipStream_t)stream));
  return kOk;
}

int tr_graph_destroy(tr_graph* g) {
  if (g) {
    if (g->exec) (void)hipGraphExecDestroy(g->exec);
    if (g->graph) (void)hipGraphDestroy(g->graph);
    delete g;
  }
  return kOk;
}

int tr_gate_set(const int* gate) {
  g_gate = gate;
  return kOk;
}

int tr_stop_rule(const double* sums, double tol, int max_iter, int rule, int* gate, int* its, double* hist,
                 int64_t* stamp, void* stream) {
  if (!sums || !gate || !its || !hist) TR_FAIL(kInvalidArgument, "null argument");
  if (max_iter < 1) TR_FAIL(kInvalidArgument, "max_iter must be positive");
  if (rule != 0 && rule != 1) TR_FAIL(kInvalidArgument, "rule is 0 (TVM, TVME) or 1 (TVMW)");
  hipLaunchKernelGGL(k_stop_rule, dim3(1), dim3(1), 0, (hipStream_t)stream, sums, tol, max_iter, rule, gate, its,
                     hist, stamp);
  TR_LAUNCH_CHECK();
  return kOk;
}

int tr_clock_stamp(int64_t* out, void* stream) {
  if (!out) TR_FAIL(kInvalidArgument, "null argument");
  hipLaunchKernelGGL(k_clock_stamp, dim3(1), dim3(1), 0, (hipStream_t)stream, out);
  TR_LAUNCH_CHECK();
  return kOk;
}

int tr_clock_rate(int64_t* hz) {
  if (!hz) TR_FAIL(kInvalidArgument, "null argument");
  int dev = 0, khz = 0;
  TR_HIP(hipGetDevice(&dev));
  TR_HIP(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev));
  *hz = (int64_t)khz * 1000;
  return kOk;
}

}  // extern "C"
