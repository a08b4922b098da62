/*
 * toporecon_hip: C ABI of the MI355X-native PDHG reconstruction hot path.
 *
 * All pointers are device pointers unless stated otherwise; complex values are
 * interleaved float32 pairs (complex64).  Every entry point returns 0 on
 * success or a negative status (-1 invalid argument, -2 HIP error,
 * -3 unsupported, -4 out of memory) and leaves a message for tr_last_error().
 * `stream` is a hipStream_t (NULL = legacy default stream).  A plan is not
 * thread-safe: calls on one plan must be serialised on one stream, because
 * they share the plan's grid workspace.
 *
 * Layouts: images (B, E, Z, Y, X) / maps (C, Z, Y, X) C-order; per-bin k-space
 * (E, C, M_b) = (K, M_b) with channel k = e * C + c; "sorted" per-sample data
 * (M_b, K) in the plan's tile order of the bin's samples.
 */
#ifndef TOPORECON_HIP_H_
#define TOPORECON_HIP_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TR_ABI_VERSION 1

typedef struct tr_plan tr_plan;
typedef struct tr_graph tr_graph;

/* complex64 as stored by NumPy / PyTorch: real then imaginary float32. */
typedef struct tr_c64 {
  float re;
  float im;
} tr_c64;

/* Message of the last failing call on this thread.
 * Replaces the Python exceptions of src/toporecon/nufft_backend.py:19-22,46-52,80-87. */
const char* tr_last_error(void);
int tr_version(void);

/* Per-kernel timing for measurement (no reference counterpart).  When enabled,
 * every launch is bracketed by HIP events on its own stream and credited with
 * its algorithmic bytes; tr_profile_read waits for the recorded events.
 * tr_profile_enable clears the counters. */
int tr_profile_enable(int on);
int tr_profile_num_kernels(void);
const char* tr_profile_kernel_name(int id);
int tr_profile_read(int id, double* ms_total, int64_t* launches, double* bytes_total);

/* ---- HIP graph capture (no reference counterpart: the reference launches every kernel from
 * Python each iteration).  Everything issued to `stream` (and to streams that join it through
 * events) between begin and end is recorded; launch replays it.  Only launches that allocate,
 * set memory or synchronise nothing may be captured: the solvers' iteration launches once
 * their workspaces exist. ---- */
int tr_graph_begin(void* stream);
int tr_graph_end(void* stream, tr_graph** out);
int tr_graph_launch(tr_graph* graph, void* stream);
int tr_graph_destroy(tr_graph* graph);

/* ---- Gated launches for many instances in one graph (no reference counterpart: the
 * reference runs one problem per invocation and stops on the host, tvm.py:416-471).
 * tr_gate_set(gate) makes every solver kernel this host thread launches afterwards (NUFFT
 * passes, PDHG element-wise kernels, wavelets, reductions) return at once while *gate != 0;
 * NULL removes the gate.  The gate pointer is read when a kernel is launched (or captured),
 * its value when the kernel runs.
 * tr_stop_rule applies the stopping rule on the device to the float64 sums (num, den) of
 * one iteration: unless *gate is already nonzero it increments *its, stores the stopping
 * value in hist[*its - 1] and sets *gate = 1 when the value is below tol or *its reaches
 * max_iter.  rule 0: sqrt(num) / sqrt(den), +inf unless both are finite and den > 0 (TVM,
 * TVME, tvm.py:416-463); rule 1: sqrt(num) / max(sqrt(den), 1e-12), +inf unless both are
 * finite (TVMW, tvmw.py:717-773).  hist holds max_iter values; stamp (NULL = none) receives the
 * device wall clock when the rule closes the gate.  tr_clock_stamp writes the device wall clock
 * to *out in stream order; tr_clock_rate gives its frequency (Hz) on the current device. ---- */
int tr_gate_set(const int* gate);
int tr_stop_rule(const double* sums, double tol, int max_iter, int rule, int* gate, int* its, double* hist,
                 int64_t* stamp, void* stream);
int tr_clock_stamp(int64_t* out, void* stream);
int tr_clock_rate(int64_t* hz);

/* ---- NUFFT plan (SigPy 0.1.27 nufft defaults: width 4, oversampling 1.25) ----
 * Replaces SigPyNufftBackend.__init__ (src/toporecon/nufft_backend.py:58-88):
 * coords[b] is a device (M_b, 3) float32 array in [z, y, x] SigPy units.  The
 * plan scales the coordinates like SigPy's _scale_coord, sorts each bin's
 * samples by grid tile, evaluates every sample's Kaiser-Bessel taps once and
 * builds the (grid cluster, sample) lists of the deterministic adjoint. */
int tr_plan_create(tr_plan** out, int nz, int ny, int nx, int nbins, const float* const* coords,
                   const int64_t* samples_per_bin, void* stream);
/* tr_plan_create with the sample geometry chosen by the caller (no reference counterpart: SigPy
 * has no gridding tiles; an autotuner picks them).  sort_tile (2 .. 32, default 8): edge of the
 * grid tiles samples are ordered by (then by first tap inside the tile), which sets which
 * samples the interpolation gathers together and the order of every cluster's pair list.
 * spread_segment (64 .. 65536, default 1024): pair lists of the spreading longer than this are
 * cut into segments of that many pairs whose partial sums are added in segment order.  Results
 * are deterministic for each geometry; two geometries add the same products in different
 * orders (float32 rounding). */
int tr_plan_create_geom(tr_plan** out, int nz, int ny, int nx, int nbins, const float* const* coords,
                        const int64_t* samples_per_bin, int sort_tile, int spread_segment, void* stream);
int tr_plan_destroy(tr_plan* plan);
/* info (12 values): [0..2] image dims, [3..5] grid dims, [6] bins, [7] M_b, [8] nnz_b, [9]
 * channels of the workspace last used, [10] distinct grid points under bin b's taps, [11] grid
 * points of the plan's pruned support (all bins). */
int tr_plan_info(const tr_plan* plan, int bin, int64_t* info);
int tr_plan_release_workspace(tr_plan* plan);

/* Launch parameters of a plan (no reference counterpart: SigPy has no tuning
 * knobs).  They change the block order only, never a result bit.
 * TR_PARAM_INTERP_XCD_CHUNK / TR_PARAM_SPREAD_XCD_CHUNK: runs of consecutive
 * workgroups kept on one XCD by the gather kernels, 0 .. 4096 (0 = hardware
 * round-robin; defaults 64 for the interpolation, 0 for the spreading). */
#define TR_PARAM_INTERP_XCD_CHUNK 1
#define TR_PARAM_SPREAD_XCD_CHUNK 2
/* TR_PARAM_PRUNE (default 1): the line passes and the spreading skip grid points outside the
 * k-space support of the plan's samples (never read by the interpolation, never written by
 * the spreading); results are identical either way.  TR_PARAM_SUPPORT_POINTS and
 * TR_PARAM_LIVE_COLUMNS are read-only: the support's size in grid points and (y, x) columns. */
#define TR_PARAM_PRUNE 3
#define TR_PARAM_SUPPORT_POINTS 4
#define TR_PARAM_LIVE_COLUMNS 5
/* TR_PARAM_INTERP_KGROUPS (default 0): 1 runs the interpolation one 64-channel group at a time
 * (group-major), shrinking the grid footprint each XCD's L2 must hold; 2 does the same with
 * 16-byte loads (two channels per lane, groups of up to 128 channels; needs an even channel
 * count >= 96 and 16-byte aligned buffers, else it runs as 1); 3 is 2 with one tap gather per
 * group of up to 4 consecutive sorted samples whose taps fit one box of 5 x 5 x 5 grid points
 * (dense k-space). */
#define TR_PARAM_INTERP_KGROUPS 6
/* TR_PARAM_WORKSPACE_SLOTS (1..8, default 1): calls for bin b use grid workspace b % slots, so
 * calls for bins in different slots may run concurrently on different streams (each slot holds
 * its own oversampled grids). */
#define TR_PARAM_WORKSPACE_SLOTS 7
/* TR_PARAM_LINES_XCD_LINES (0..64, default 4): the y and z FFT passes keep runs of this many
 * whole lines (all channel chunks of each) on one XCD, so the chunks of a grid point's channel
 * record are fetched and written close together in time (0 = hardware round-robin). */
#define TR_PARAM_LINES_XCD_LINES 8
/* TR_PARAM_SPREAD_ORDER (default 65535): order in which the spreading visits grid clusters.
 * 65535: the light clusters holding a support point, longest pair list first, after the heavy
 * clusters' segments; bz | by << 4 | bx << 8 (each >= 1): the same clusters in bricks of
 * bz x by x bx clusters; 65536 + (bz | by << 4 | bx << 8): those bricks, the brick holding the
 * longest list first, index order inside a brick; 0: every cluster in index order.  Results are
 * identical for every order. */
#define TR_PARAM_SPREAD_ORDER 10
/* TR_PARAM_SORT_TILE, TR_PARAM_SPREAD_SEGMENT: read-only, the geometry the plan was built with. */
#define TR_PARAM_SORT_TILE 11
#define TR_PARAM_SPREAD_SEGMENT 12
int tr_plan_set_param(tr_plan* plan, int param, int64_t value);
int tr_plan_get_param(const tr_plan* plan, int param, int64_t* value);

/* Sample-order helpers for the fused solver path (no reference counterpart:
 * the reference keeps samples in trajectory order).  gather: (K, M_b) original
 * order -> (M_b, K) sorted; scatter: inverse; gather_weights: (M_b) -> (M_b), scatter_weights its
 * inverse. */
int tr_plan_gather_samples(const tr_plan* plan, int bin, const tr_c64* src, int K, tr_c64* dst,
                           void* stream);
int tr_plan_scatter_samples(const tr_plan* plan, int bin, const tr_c64* src, int K, tr_c64* dst,
                            void* stream);
int tr_plan_gather_weights(const tr_plan* plan, int bin, const float* src, float* dst, void* stream);
int tr_plan_scatter_weights(const tr_plan* plan, int bin, const float* src, float* dst, void* stream);

/* Forward NUFFT of E*C images of one bin: out[(e*C + c), m] = nufft(img[e] * maps[c]).
 * Replaces SigPyNufftBackend.forward -> sp.nufft (src/toporecon/nufft_backend.py:101-107);
 * maps == NULL requires C == 1.  out is (E*C, M_b) in the trajectory's order. */
int tr_nufft_forward(tr_plan* plan, int bin, const tr_c64* img, const tr_c64* maps, int E, int C,
                     tr_c64* out, void* stream);

/* Adjoint NUFFT: out[e] (+)= sum_c conj(maps[c]) * nufft_adjoint(dcf * ksp[e*C + c]).
 * Replaces SigPyNufftBackend.adjoint -> sp.nufft_adjoint (src/toporecon/nufft_backend.py:109-116)
 * together with the coil loop of tvm.py:216-225; dcf may be NULL, maps == NULL requires C == 1. */
int tr_nufft_adjoint(tr_plan* plan, int bin, const tr_c64* ksp, const float* dcf, const tr_c64* maps,
                     int E, int C, tr_c64* out, int accumulate, void* stream);

/* FFT-pass access for tests and diagnostics (no reference counterpart):
 * forward_grid runs the apodise/pad/FFT passes of tr_nufft_forward and copies
 * the (gz, gy, gx, E*C) grid in FFT-natural order; adjoint_grid runs the
 * inverse FFT / crop / apodise / coil-sum passes of tr_nufft_adjoint on a
 * given grid. */
int tr_nufft_forward_grid(tr_plan* plan, const tr_c64* img, const tr_c64* maps, int E, int C,
                          tr_c64* grid, void* stream);
int tr_nufft_adjoint_grid(tr_plan* plan, const tr_c64* grid, const tr_c64* maps, int E, int C,
                          tr_c64* out, void* stream);

/* Fused data-term dual step of one bin on sorted samples:
 * q <- (q + sigma (A ubar_b - y)) / one_plus_sigma.
 * Replaces the forward triple loop + prox of tvm.py:246-260 / tvme.py:257-267 / tvmw.py:454-465. */
int tr_data_dual_update(tr_plan* plan, int bin, const tr_c64* ubar_b, const tr_c64* maps, int E, int C,
                        const tr_c64* y_sorted, tr_c64* q_sorted, float sigma, float one_plus_sigma,
                        void* stream);
/* y_sorted = A x_b (synthetic k-space; no reference counterpart). */
int tr_data_simulate(tr_plan* plan, int bin, const tr_c64* x_b, const tr_c64* maps, int E, int C,
                     tr_c64* y_sorted, void* stream);
/* q_b = A x_b - y_b on sorted samples: the data dual's PDHG fixed point for an initial
 * image x (no reference counterpart: the reference starts q at zero, tvme.py:180-186). */
int tr_data_residual(tr_plan* plan, int bin, const tr_c64* x, const tr_c64* maps, int E, int C,
                     const tr_c64* y_sorted, tr_c64* q_sorted, void* stream);
/* g_b (+)= A^H W q: sum_c conj(S_c) NUFFT^H(dcf * q_c) on sorted samples.
 * Replaces tvm.py:216-225 (+ coil all-reduce tvm.py:188-214), tvme.py:383-419, tvmw.py:586-629. */
int tr_data_adjoint(tr_plan* plan, int bin, const tr_c64* q_sorted, const float* dcf_sorted,
                    const tr_c64* maps, int E, int C, tr_c64* g_b, int accumulate, void* stream);

/* Conjugate-gradient vector steps (the warm start's data-consistency refinement; no reference
 * counterpart).  partial[b] receives block b's float64 sum (nblocks blocks, added in order by the
 * caller): update x += alpha p, r -= alpha Ap, sum |r|^2; direction p = r + beta p;
 * dot sum Re(conj(a) b). */
int tr_cg_update(tr_c64* x, tr_c64* r, const tr_c64* p, const tr_c64* ap, int64_t n, double alpha,
                 double* partial, int nblocks, void* stream);
int tr_cg_direction(tr_c64* p, const tr_c64* r, int64_t n, double beta, void* stream);
int tr_cdot_re(const tr_c64* a, const tr_c64* b, int64_t n, double* partial, int nblocks, void* stream);
/* out = s * x (complex64, n values): the power iteration's normalisation. */
int tr_cscale(const tr_c64* x, int64_t n, double s, tr_c64* out, void* stream);
/* ---- element-wise PDHG steps on (B, E, Z, Y, X) arrays ----
 * `partial` holds tr_partial_count() float64 pairs (sum |u - u_old|^2, sum |u|^2)
 * reduced by tr_reduce_partials into out2[2]. */
int64_t tr_partial_count(int B, int Z, int Y, int X);
int tr_reduce_partials(const double* partial, int64_t n, double* out2, void* stream);

/* TVM motion dual + prox (tvm.py:262-309). */
int tr_tvm_dual(int B, int E, int Z, int Y, int X, const tr_c64* ubar, tr_c64* p_m, float sigma,
                float lam, int l2_coupling, void* stream);
/* TVM primal + extrapolation + stopping sums (tvm.py:342-379, 416-424). */
int tr_tvm_primal(int B, int E, int Z, int Y, int X, tr_c64* u, tr_c64* ubar, const tr_c64* grad,
                  const tr_c64* p_m, float tau, double* partial, void* stream);
/* TVME motion + echo/spatial TV duals and prox (tvme.py:233-337). */
int tr_tvme_dual(int B, int E, int Z, int Y, int X, const tr_c64* ubar, tr_c64* p_m, tr_c64* p_ex,
                 tr_c64* p_ey, tr_c64* p_ez, float sigma, float lam_m, float lam_e, int l2_coupling,
                 void* stream);
/* TVME divergences, primal update, extrapolation, stopping sums (tvme.py:356-480, 520-523). */
int tr_tvme_primal(int B, int E, int Z, int Y, int X, tr_c64* u, tr_c64* ubar, const tr_c64* grad,
                   const tr_c64* p_m, const tr_c64* p_ex, const tr_c64* p_ey, const tr_c64* p_ez,
                   float tau, double* partial, void* stream);

/* No counterpart in the reference: tr_tvme_primal followed by tr_reduce_partials, allocating
 * nothing, for callers that cannot allocate device memory (code under HIP graph capture, serving
 * loops or harnesses that forbid hipMalloc).  The per-block partial sums live in a 4 MiB device
 * array owned by the library (one per GPU, held until the process ends), so B*Z*Y*X may be at most
 * 2^26 (else -1, nothing launched).  sums: device double[2] receiving (sum |u - u_old|^2,
 * sum |u|^2), bit-identical to the two-call path.  `stream` must belong to the current device.
 * Calls that overlap on one GPU share that array and race: serialise them. */
int tr_tvme_primal_noalloc(int B, int E, int Z, int Y, int X, tr_c64* u, tr_c64* ubar, const tr_c64* grad,
                           const tr_c64* p_m, const tr_c64* p_ex, const tr_c64* p_ey, const tr_c64* p_ez,
                           float tau, double* sums, void* stream);

/* TVMW motion dual (lambda1) + db1 echo-wavelet dual (lambda2) (tvmw.py:468-520). */
int tr_tvmw_dual(int B, int E, int Z, int Y, int X, const tr_c64* ubar, tr_c64* p_m, tr_c64* p_w1,
                 float sigma, float lam1, float lam2, int l2_coupling, void* stream);
/* TVMW primal: u -= tau ((g - divp_m + W1^H p_w1) + W2^H p_w2); ubar = 2u - u_old; stopping
 * sums of |ubar_prev - ubar|^2 and |ubar_prev|^2 (tvmw.py:661-684, 723-769). */
int tr_tvmw_primal(int B, int E, int Z, int Y, int X, tr_c64* u, tr_c64* ubar, const tr_c64* grad,
                   const tr_c64* p_m, const tr_c64* p_w1, const tr_c64* divw2, float tau, double* partial,
                   void* stream);

/* ---- db6 zero-mode 3-D wavelet (ptwt.wavedec3 / waverec3, tvmw.py:196-206) ----
 * Coefficients per image, flat: cA_L, level-L details, ..., level-1 details, each level's
 * details ordered aad, ada, add, daa, dad, dda, ddd (letters = z, y, x).  `work` holds
 * nimg * tr_w2_work_count() complex values.  level 0 is the identity. */
int64_t tr_w2_coeff_count(int Z, int Y, int X, int level);
int64_t tr_w2_work_count(int Z, int Y, int X, int level);
int tr_w2_forward(const tr_c64* img, int64_t nimg, int Z, int Y, int X, int level, tr_c64* coeffs,
                  tr_c64* work, void* stream);
int tr_w2_adjoint(const tr_c64* coeffs, int64_t nimg, int Z, int Y, int X, int level, tr_c64* out,
                  tr_c64* work, void* stream);
/* p <- p + sigma w projected onto |p| <= lam per coefficient with sqrt(re^2 + im^2 + 1e-12)
 * (0 when lam <= 0) (tvmw.py:162-193). */
int tr_w2_dual_update(tr_c64* p, const tr_c64* w, int64_t n, float sigma, float lam, void* stream);

/* ---- learned warm start (no counterpart in the reference, whose solvers start from
 * zeros: tvme.py:180-186).  A 3-D CNN of 3x3x3 'same' convolutions maps a (bin, echo)
 * image to an estimate of the converged one.  Activations are channels-last bf16
 * volumes [n][Z][Y][X][C] (C = 8 for the input: re, im, 6 zeros; 32 otherwise);
 * master weights fp32 [27][Cin][32] with tap t = 9(dz+1) + 3(dy+1) + (dx+1).
 * bf16 values are passed as uint16_t bit patterns. ---- */
/* wfrag: tr_cnn_pack_weights output.  mode 0: + bias, ReLU -> bf16 [..][32]; mode 1: + bias
 * -> fp32 [..][cout_keep]; mode 2 (data gradient with the flipped kernel): zero where
 * mask <= 0 -> bf16 [..][32]. */
int tr_cnn_conv(const uint16_t* x, int n, int Z, int Y, int X, int cin, const uint16_t* wfrag,
                const float* bias, int mode, const uint16_t* mask, void* out, int cout_keep, void* stream);
/* MFMA fragments of w (27 * cin * 32 bf16); flip = 1 packs W'[t][co][ci] = W[26-t][ci][co]. */
int tr_cnn_pack_weights(const float* w, int cin, int flip, uint16_t* frag, void* stream);
/* dw[27][cin][32] = sum_v x[v + off(t)][ci] dy[v][co], db[32] = sum_v dy[v][co]; nblocks
 * workgroups write partial sums to work (tr_cnn_wgrad_work_size floats), added in block order. */
int64_t tr_cnn_wgrad_work_size(int nblocks);
int tr_cnn_wgrad(const uint16_t* x, const uint16_t* dy, int n, int Z, int Y, int X, int cin, int nblocks,
                 float* work, float* dw, float* db, void* stream);
/* dy[v] = (scale (out - target)[v][0:2], 0 x 30) bf16; partial[2 * b] = block b's sum of squares. */
int tr_cnn_mse(const float* out, const float* target, int64_t V, float scale, uint16_t* dy, double* partial,
               int nblocks, void* stream);
int tr_cnn_adam(float* p, const float* g, float* m, float* v, int64_t n, float lr, float b1, float b2,
                float eps, int step, void* stream);
/* x[i][v] = bf16(img[i][v] / scale[i]) as (re, im, 0 x 6); img = scale (x[0:2] + out[0:2]). */
int tr_cnn_pack_input(const tr_c64* img, const float* scale, int64_t V, int n, uint16_t* x, void* stream);
int tr_cnn_unpack(const uint16_t* x, const float* out, const float* scale, int64_t V, int n, tr_c64* img,
                  void* stream);
/* out[i] = sum_v |img[i][v]|^2 (float64, fixed order). */
int tr_cnn_sumsq(const tr_c64* img, int64_t V, int n, double* out, void* stream);
/* out[k] = (sum_m dcf_m Re(conj(a_mk) y_mk), sum_m dcf_m |a_mk|^2) over sorted samples (float64;
 * work holds 2 * nblocks * K doubles): the least-squares scale of a predicted image. */
int tr_data_fit_sums(const tr_c64* a_sorted, const tr_c64* y_sorted, const float* dcf_sorted, int64_t M, int K,
                     double* work, int nblocks, double* out, void* stream);
/* q[m, k] *= scale[k] on sorted samples (M, K), scale a device array of K floats: A (c v) from A v
 * for a per-(bin, echo) scale c, so the warm start's CG needs no second forward NUFFT (no reference
 * counterpart). */
int tr_data_scale_channels(tr_c64* q_sorted, int64_t M, int K, const float* scale, void* stream);

/* ---- JSENSE coil-sensitivity estimation ----
 * Replace the linear operators of SigPy 0.1.27's sigpy.mri.app.JsenseRecon, which the reference
 * runs on the first echo (src/toporecon/preprocessing/jsense.py:91).  Its operators are
 * sqrt(w) NUFFT(IFFT(mps_ker_c (*) img_ker)) (sigpy.mri.linop ConvSense / ConvImage); the NUFFT is
 * a plan of this library (tr_data_simulate / tr_data_adjoint), the rest is below.
 * (*) is SigPy's 'valid' convolution (sp.linop.ConvolveInput / ConvolveFilter, mode 'valid',
 * output_multi_channel): img (iz, iy, ix), ker (C, W, W, W), out (C, iz-W+1, iy-W+1, ix-W+1),
 * out[c, n] = sum_k ker[c, k] img[n + W - 1 - k].  W <= 64, W <= each dim <= 1024.
 * work holds tr_js_work_count() complex values (partial sums, added in a fixed order). */
int64_t tr_js_work_count(int iz, int iy, int ix, int C, int W);
int tr_js_conv(const tr_c64* img, int iz, int iy, int ix, const tr_c64* ker, int C, int W, tr_c64* out,
               tr_c64* work, void* stream);
/* Adjoint in the image kernel: out (iz, iy, ix) = sum_c conv^H(z_c; ker_c), z (C, iz-W+1, ..). */
int tr_js_conv_adj_img(const tr_c64* z, int iz, int iy, int ix, const tr_c64* ker, int C, int W, tr_c64* out,
                       tr_c64* work, void* stream);
/* Adjoint in the map kernel: out (C, W, W, W), out[c, k] = sum_n z[c, n] conj(img[n + W - 1 - k]). */
int tr_js_conv_adj_ker(const tr_c64* z, const tr_c64* img, int iz, int iy, int ix, int C, int W, tr_c64* out,
                       tr_c64* work, void* stream);
/* out[o, j, t] = sum_i mat[j, i] x[o, i, t] for x (outer, n_in, inner), mat (n_out, n_in) on the device,
 * out (outer, n_out, inner), not aliasing x: one axis of sp.fft / sp.ifft (centred, orthonormal) on
 * the calibration grid, and of the k-space resizes of JsenseRecon's output and of
 * jsense.py:93-96 (ifft(resize(fft(map)))). */
int tr_js_axis_apply(const tr_c64* x, int64_t outer, int n_in, int64_t inner, const tr_c64* mat, int n_out,
                     tr_c64* out, void* stream);
/* maps (C, N): each voxel divided by the root sum of squares over the coils (no floor), as
 * JsenseRecon's output step. */
int tr_js_rss_normalize(tr_c64* maps, int C, int64_t N, void* stream);
/* y += a x (n complex values): the lamda * I term of JsenseRecon's regularised least squares. */
int tr_js_axpy(tr_c64* y, const tr_c64* x, int64_t n, float a, void* stream);

/* ---- Respiratory self-gating (src/toporecon/preprocessing/respiratory.py:86-92): the band-pass of
 * every (echo, coil) centre trace.  traces (ntraces, T) float32 (magnitude minus its mean, computed
 * by the caller as the reference does), out (ntraces, T) float32 = sosfiltfilt(hi, sosfiltfilt(lo,
 * trace)) with SciPy's semantics: odd extension of pad_* samples at each end (in the input's
 * precision: float32 for the first filter, float64 for the second), a forward and a backward pass of
 * the n_* second-order sections sos_* (n, 6) started from zi_* (n, 2) times the pass's first sample,
 * float64 arithmetic with every product and sum rounded in the order of SciPy's loop.  n_* <= 16,
 * pad_* < T.  work holds tr_resp_work_count() float64 values.  One thread per trace. */
int64_t tr_resp_work_count(int64_t ntraces, int64_t T, int pad_lo, int pad_hi);
int tr_resp_bandpass(const float* traces, int64_t ntraces, int64_t T, const double* sos_lo, const double* zi_lo,
                     int n_lo, int pad_lo, const double* sos_hi, const double* zi_hi, int n_hi, int pad_hi,
                     double* work, float* out, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* TOPORECON_HIP_H_ */
