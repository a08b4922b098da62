#!/usr/bin/env python3
"""TVME time-to-tolerance benchmark of the HIP PDHG hot path on one MI355X.

Workload (default): TVME PDHG to tol 1e-3 (max 300 iterations, sigma = tau =
1/6, lambda_motion = lambda_echo = 1e-5, L2 echo coupling) at 80x192x192,
32 coils, 6 echoes (TE = 1.0 + 1.3 e ms), 6 motion bins, 9000 cone
interleaves x 512 samples per echo (readout fraction 0.98, 2/98 % respiratory
outlier drop), synthetic seeded data simulated on the GPU.  A step is one
full solve; inputs are resident in HBM when the timed region starts.

Two legs run on the same problem: plain PDHG from zeros (as the reference),
and the learned warm start (density-compensated adjoint -> 3-D CNN on MFMA ->
per-(bin, echo) data-fitted scale -> 4 conjugate-gradient data-consistency
steps -> PDHG from there to the same tolerance, the warm-start work inside the
timed region).  ``value`` is the learned leg's time-to-tol when its final
objective is within +-0.1 % of the plain leg's, otherwise the plain leg's (the
fallback); both legs are reported.  Variant starts (CNN without CG, CG from
zeros without the CNN, CNN with 3 or 5 CG steps) are timed too and reported
with their signed objective differences; they never set ``value``.

Prints one JSON line: time-to-tol (value, seconds), iterations, per-kernel
algorithmic GB/s from HIP events recorded on the launch stream (plain leg),
the dominant kernel's fraction of the 8 TB/s HBM peak with its measured HBM
traffic, forward/adjoint NUFFT GB/s, and a single-core CPU baseline (the
float64 oracle timed on a bounded sample and extrapolated to the plain
iteration count).
"""

from __future__ import annotations

import argparse
import json
import math
import os
import subprocess
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(REPO, "multinodemri-miccai26_amd")
for _p in (REPO, PKG):
    if _p not in sys.path:
        sys.path.insert(0, _p)

METRIC = "TVME time-to-tol 1e-3 (learned warm start vs plain) on 1 GPU; NUFFT GB/s"
HBM_PEAK_GBPS = 8000.0
TRAFFIC_FILE = os.path.join(REPO, "profiles", "traffic.json")


def parse_args(argv=None):
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", choices=("full", "small", "batched", "tvm48", "cnn"), default="full",
                    help="full: TVME 80x192x192 (the headline); small: a reduced TVME; batched: 12 uneven "
                         "TVM instances, HIP-graph batch vs sequential solves; tvm48: plain TVM at "
                         "48x64x64, 8 coils, 2 echoes, 4 bins beside one full float64 oracle iteration; cnn: "
                         "the warm-start network's MFMA convolutions (inference at 80x192x192, training "
                         "steps at 64x96x96)")
    ap.add_argument("--max-iter", type=int, default=None, help="override max_iter (profiling runs)")
    ap.add_argument("--tol", type=float, default=None)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-baseline-worker", action="store_true", help=argparse.SUPPRESS)
    ap.add_argument("--verbose", action="store_true")
    ap.add_argument("--traffic-from", nargs=2, metavar=("FETCH_CSV", "WRITE_CSV"),
                    help="summarize two rocprofv3 --pmc counter_collection.csv files (FETCH_SIZE, "
                         "WRITE_SIZE) of a run with --copy-calibration-mib into profiles/traffic.json")
    ap.add_argument("--params", default=None,
                    help="JSON launch parameters (interp_xcd_chunk, interp_kgroups, spread_xcd_chunk, "
                         "lines_xcd_lines, bin_streams) to set instead of autotuning "
                         "(profiling runs)")
    ap.add_argument("--legs", choices=("all", "plain"), default="all",
                    help="plain: only the plain PDHG leg (no warm start, no power iteration)")
    ap.add_argument("--cg-steps", type=int, default=None,
                    help="CG steps of the learned leg (default: the warm-start module's default, 4)")
    ap.add_argument("--variants", default="cnn+0,zero+4,cnn+3,cnn+5",
                    help="comma list of extra warm starts 'cnn+K' / 'zero+K' (K CG steps) timed and "
                         "reported beside the learned leg; '' for none")
    ap.add_argument("--copy-calibration-mib", type=int, default=0,
                    help="before the solve, copy a buffer of this many MiB once (a kernel of known "
                         "traffic for scaling hardware byte counters)")
    return ap.parse_args(argv)


def log(msg, verbose=True):
    if verbose:
        print(f"[bench] {msg}", file=sys.stderr, flush=True)


def make_config(name):
    from toporecon_hip import synth
    if name == "full":
        return synth.config_full_tvme()
    cfg = synth.config_scaled(synth.config_full_tvme(), (40, 64, 64), coils=8, echoes=3, bins=4,
                              interleaves=2000, readouts=256)
    cfg.algorithm = "tvme"
    cfg.l2_coupling = True
    return cfg


def nufft_model_bytes(plan_samples, E, C, img_shape, grid_shape):
    """Algorithmic bytes of one forward and one adjoint pass over all bins, on the unpruned model
    (every grid pass moves the whole oversampled grid), for the NUFFT GB/s figures:

    forward per bin: pad 8EN + 8CN + 8KG, fft3 16KG, interp 8KG + 12M + 24KM;
    adjoint per bin: spread 8KM + 16M + 8KG, ifft3 16KG, crop 8KN + 8CN + 8EN.
    """
    K = E * C
    N = math.prod(img_shape)
    G = math.prod(grid_shape)
    fwd = adj = 0.0
    for M in plan_samples:
        fwd += 8 * E * N + 8 * C * N + 8 * K * G + 16 * K * G + 8 * K * G + 12 * M + 24 * K * M
        adj += 8 * K * M + 16 * M + 8 * K * G + 16 * K * G + 8 * K * N + 8 * C * N + 8 * E * N
    return fwd, adj


def cpu_baseline_worker(cfg_name):
    """Time the float64 oracle on one bin x one channel forward + adjoint and one bin's TV steps."""
    import numpy as np

    from oracle import nufft as on
    from oracle import pdhg as op
    from toporecon_hip import synth
    cfg = make_config(cfg_name)
    t0 = time.perf_counter()
    acq = synth.make_acquisition(cfg)
    idx = acq.bin_index(0)
    coord = acq.coord[idx].reshape(-1, 3)
    M = coord.shape[0]
    img = (acq.truth[0, 0] * acq.mps[0]).astype(np.complex128)
    gen_s = time.perf_counter() - t0
    t0 = time.perf_counter()
    y = on.nufft(img, coord)
    t_fwd = time.perf_counter() - t0
    t0 = time.perf_counter()
    on.nufft_adjoint(y * acq.dcf[idx].reshape(-1), coord, cfg.img_shape)
    t_adj = time.perf_counter() - t0
    # element-wise TVME work of one bin (all echoes), float64
    u = acq.truth[:1].astype(np.complex128)
    t0 = time.perf_counter()
    de = op.echo_diff(u)
    px, py, pz = op.prox_iso3(op.fdiff(de, -1), op.fdiff(de, -2), op.fdiff(de, -3), cfg.lambda_echo)
    pm = op.prox_motion(op.motion_diff(u), cfg.lambda_motion, True)
    divp = op.fdiff_adj(px, -1) + op.fdiff_adj(py, -2) + op.fdiff_adj(pz, -3)
    _ = u - (1 / 6) * op.motion_div_tvme(pm) - (1 / 6) * op.echo_div(divp)
    t_ew = time.perf_counter() - t0
    K = cfg.echoes * cfg.coils
    per_iter = cfg.bins * (K * (t_fwd + t_adj) + t_ew)
    print(json.dumps({"t_fwd": t_fwd, "t_adj": t_adj, "t_ew_bin": t_ew, "per_iter": per_iter,
                      "samples": M, "K": K, "bins": cfg.bins, "gen_s": gen_s}))


def run_cpu_baseline(cfg_name, iterations, verbose):
    env = dict(os.environ)
    for k in ("OMP_NUM_THREADS", "OPENBLAS_NUM_THREADS", "MKL_NUM_THREADS", "NUMEXPR_NUM_THREADS"):
        env[k] = "1"
    cmd = [sys.executable, os.path.abspath(__file__), "--cpu-baseline-worker", "--config", cfg_name]
    log("cpu baseline: timing the float64 oracle on 1 core", verbose)
    r = subprocess.run(cmd, capture_output=True, text=True, env=env, timeout=900)
    if r.returncode != 0:
        log(f"cpu baseline failed: {r.stderr[-2000:]}", True)
        return None
    d = json.loads(r.stdout.strip().splitlines()[-1])
    est = d["per_iter"] * iterations
    sample = (f"float64 oracle, 1 thread: forward+adjoint NUFFT of 1 bin x 1 (coil, echo) channel "
              f"({d['samples']} samples; {d['t_fwd']:.1f} s + {d['t_adj']:.1f} s) plus one bin of TVME "
              f"element-wise steps ({d['t_ew_bin']:.2f} s), scaled x{d['K']} channels x{d['bins']} bins "
              f"= {d['per_iter']:.0f} s/iteration, x{iterations} iterations")
    out = {"value": est, "unit": "s", "cores": 1, "kind": "port", "sample": sample,
           "s_per_iteration": d["per_iter"]}
    # all host cores: the (bin, channel) NUFFTs are independent, so W single-threaded processes
    # each time the same one-channel sample at once (memory bandwidth shared as in a real run)
    W = max(1, min(16, len(os.sched_getaffinity(0))))
    log(f"cpu baseline: {W} concurrent single-threaded oracle processes", verbose)
    procs = [subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True, env=env)
             for _ in range(W)]
    res = []
    for pr in procs:
        so, se = pr.communicate(timeout=1200)
        if pr.returncode == 0:
            res.append(json.loads(so.strip().splitlines()[-1]))
    if len(res) == W:
        t_ch = sum(x["t_fwd"] + x["t_adj"] for x in res) / W
        t_ew = sum(x["t_ew_bin"] for x in res) / W
        per_iter = d["bins"] * (d["K"] * t_ch / W + t_ew / W)
        out["all_cores"] = {"value": per_iter * iterations, "unit": "s", "cores": W, "kind": "port",
                            "s_per_iteration": per_iter,
                            "sample": f"{W} processes x 1 thread, each timing the 1-channel forward+adjoint "
                                      f"sample ({t_ch:.1f} s under load) and one bin's element-wise steps; "
                                      f"{d['bins']} bins x {d['K']} channels spread over {W} cores"}
    return out


# Kernel symbols -> the logical kernel names used in the JSON line.  The line
# passes share one symbol per length; forward and inverse passes are told
# apart by the kernel that ran before them on the stream.
def _logical_kernel(sym, prev):
    if "k_fwd_x_pad" in sym:
        return "fwd_x_pad"
    if "k_adj_x_crop" in sym:
        return "adj_x_crop"
    if "k_interp" in sym:
        return "interp"
    if "k_spread" in sym:
        return "spread"
    if "k_fft_lines" in sym:
        return {"fwd_x_pad": "fft_y_fwd", "fft_y_fwd": "fft_z_fwd", "spread": "fft_z_inv",
                "fft_z_inv": "fft_y_inv"}.get(prev)
    for name in ("tvme_dual", "tvme_primal", "tvm_dual", "tvm_primal", "tvmw_dual", "tvmw_primal"):
        if f"k_{name}(" in sym:
            return name
    return None


def _launch_counters(path, counter):
    """{logical kernel: [bytes per launch as counted]} and the calibration copy's count."""
    import csv
    rows = sorted(csv.DictReader(open(path)), key=lambda r: int(r["Dispatch_Id"]))
    rows = [r for r in rows if r["Counter_Name"] == counter]
    copy_counted = [float(r["Counter_Value"]) * 1024.0 for r in rows
                    if "copyBuffer" in r["Kernel_Name"] and float(r["Counter_Value"]) > 1024 * 1024]
    launches, prev = {}, None
    for r in rows:
        name = _logical_kernel(r["Kernel_Name"], prev)
        if name is None:
            continue
        v = float(r["Counter_Value"]) * 1024.0  # the counters are in KiB
        if name == "spread" and prev == "spread":
            launches[name][-1] += v  # segment and combine kernels of the same spread launch
            continue
        launches.setdefault(name, []).append(v)
        prev = name
    return launches, copy_counted


def write_traffic(fetch_csv, write_csv, copy_mib, out_path):
    """HBM bytes per launch of every logical kernel from FETCH_SIZE / WRITE_SIZE, each
    scaled by copy bytes / what the counter reads for the calibration copy."""
    import statistics
    fetch, fcopy = _launch_counters(fetch_csv, "FETCH_SIZE")
    write, wcopy = _launch_counters(write_csv, "WRITE_SIZE")
    if not fcopy or not wcopy:
        raise SystemExit("no calibration copy found: profile the bench with --copy-calibration-mib")
    copy_bytes = copy_mib * float(1 << 20)
    f_factor = copy_bytes / max(fcopy)
    w_factor = copy_bytes / max(wcopy)
    kernels = {}
    for name in sorted(set(fetch) | set(write)):
        fb = statistics.median(fetch.get(name, [0.0]))
        wb = statistics.median(write.get(name, [0.0]))
        kernels[name] = {"fetch_bytes_counted": fb, "write_bytes_counted": wb,
                         "bytes": fb * f_factor + wb * w_factor,
                         "launches_sampled": len(fetch.get(name, []))}
    out = {"calibration": {"copy_bytes": copy_bytes, "fetch_factor": f_factor, "write_factor": w_factor,
                           "note": "factor = copy_bytes / counter value of one device copy of copy_bytes"},
           "kernels": kernels}
    with open(out_path, "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)
    return out


def bench_batched(args, verbose=True):
    """Twelve uneven TVM instances: sequential eager solves vs one HIP-graph batch."""
    import numpy as np
    import torch

    from toporecon_hip import synth
    from toporecon_hip.algorithms import TvmReconstructor
    from toporecon_hip.batch import GraphBatch

    dev = torch.device("cuda", 0)
    cfgs = synth.batched_instance_configs()
    solvers, rows = [], []
    t0 = time.perf_counter()
    for cfg in cfgs:
        acq = synth.make_acquisition(cfg)
        s = TvmReconstructor(None, acq.coord, acq.dcf, acq.mps, acq.resp, None, cfg.bins, cfg.l2_coupling,
                             lamda=cfg.lambda_motion, sigma=1 / 6, tau=1 / 6, max_iter=cfg.max_iter, tol=cfg.tol,
                             margin=0, device=dev, bin_edges=acq.bin_edges, show_pbar=False,
                             echoes=len(cfg.echo_times_ms))
        s.simulate_kspace(acq.truth, cfg.snr_db, seed=cfg.seed + 104729)
        tuned = s.autotune(reps=1, bin_streams=False)  # graph batches keep one stream per instance
        solvers.append(s)
        rows.append({"seed": cfg.seed, "shape": list(cfg.img_shape), "coils": cfg.coils,
                     "echoes": len(cfg.echo_times_ms), "bins": cfg.bins,
                     "samples_per_echo": int(acq.coord.shape[0] * acq.coord.shape[1]),
                     "launch_params": {k: tuned.get(k) for k in ("interp_xcd_chunk", "spread_xcd_chunk")}})
        log(f"instance seed {cfg.seed}: {cfg.img_shape} C={cfg.coils} E={len(cfg.echo_times_ms)} B={cfg.bins} "
            f"samples/echo={rows[-1]['samples_per_echo']} ({time.perf_counter() - t0:.0f}s)", verbose)
    for s in solvers:  # warm every code path once
        s.max_iter, full = 2, s.max_iter
        s.run()
        s.max_iter = full
    torch.cuda.synchronize(dev)
    # sequential: each instance alone, eager launches, one stop-sum readback per iteration
    results = []
    t0 = time.perf_counter()
    for s, r in zip(solvers, rows):
        ta = time.perf_counter()
        results.append(s.run().clone())
        torch.cuda.synchronize(dev)
        r["sequential"] = {"iterations": s.iterations, "seconds": time.perf_counter() - ta}
    t_seq = time.perf_counter() - t0
    log(f"sequential: {t_seq:.2f} s", verbose)
    # batched: one HIP-graph launch per iteration for every running instance, and the same
    # iteration issued eagerly with one stream per instance
    legs = {}
    for mode in ("graph", "streams"):
        batch = GraphBatch(solvers)
        torch.cuda.synchronize(dev)
        t0 = time.perf_counter()
        out = batch.run(mode=mode)
        torch.cuda.synchronize(dev)
        tb = time.perf_counter() - t0
        same = all(bool(torch.equal(s.ubar, ref)) for s, ref in zip(solvers, results))
        for r, o in zip(rows, out):
            r[mode] = {"iterations": o["iterations"], "seconds": o["seconds"]}
        legs[mode] = {"seconds": tb, "captures": batch.captures, "bit_identical_to_sequential": same,
                      "iterations": sum(o["iterations"] for o in out)}
        log(f"{mode}: {tb:.2f} s ({batch.captures} graph captures), identical results: {same}", verbose)
    its = sum(r["sequential"]["iterations"] for r in rows)
    for r in rows:
        for leg in ("sequential", "graph", "streams"):
            r[leg]["iterations_per_s"] = r[leg]["iterations"] / max(r[leg]["seconds"], 1e-9)
    best = min(legs, key=lambda m: legs[m]["seconds"])
    t_batch = legs[best]["seconds"]
    its_b = legs[best]["iterations"]
    same = all(v["bit_identical_to_sequential"] for v in legs.values())
    line = {
        "metric": "TVM batched uneven instances: aggregate PDHG iterations/s (batched)",
        "value": its_b / t_batch, "unit": "iterations/s", "n_gpus": 1, "steps": 1, "warmup": 1,
        "ms_per_step": t_batch * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic: seeded TVM instances from the small-TVM generator",
        "config": {"workload": "12 uneven TVM instances (32^3 .. 96x128x128, 8-32 coils, 1-6 echoes, "
                               "4-8 bins, 0.2M-3M samples/echo) to tol 1e-3, max 100 iterations",
                   "batched_mode": best, "sequential_iterations_per_s": its / t_seq, "sequential_seconds": t_seq,
                   "legs": legs, "speedup": t_seq / t_batch, "bit_identical_to_sequential": same,
                   "instances": rows},
    }
    print(json.dumps(line))
    return 0


def tvm48_oracle_iteration():
    """One full float64-oracle TVM PDHG iteration at 48x64x64 (all bins, coils, echoes), 1 thread."""
    import numpy as np

    from oracle import nufft as on
    from oracle import pdhg as op
    from toporecon_hip import synth
    cfg = synth.config_small_tvm()
    acq = synth.make_acquisition(cfg)
    ksp = synth.simulate_kspace_numpy(acq, lambda x, c: on.nufft(x.astype(np.complex128), c))
    prob = op.bin_problem(ksp, acq.coord, acq.dcf, acq.mps, acq.resp, acq.bin_edges)
    t0 = time.perf_counter()
    op.run(prob, "tvm", lam_m=cfg.lambda_motion, l2_coupling=cfg.l2_coupling, max_iter=1, tol=0.0)
    print(json.dumps({"per_iter": time.perf_counter() - t0, "bins": prob.B, "K": prob.E * prob.C,
                      "samples": [int(k.shape[-1]) for k in prob.ksp]}))


def bench_tvm48(args, verbose=True):
    """Plain TVM to tol 1e-3 at 48x64x64 (8 coils, 2 echoes, 4 bins, 1536x256 cone samples) on the GPU,
    beside one full float64-oracle iteration on one host core, scaled to the GPU's iteration count."""
    import torch

    from toporecon_hip import synth
    from toporecon_hip.algorithms import TvmReconstructor
    dev = torch.device("cuda", 0)
    cfg = synth.config_small_tvm()
    acq = synth.make_acquisition(cfg)
    s = TvmReconstructor(None, acq.coord, acq.dcf, acq.mps, acq.resp, None, cfg.bins, cfg.l2_coupling,
                         lamda=cfg.lambda_motion, sigma=1 / 6, tau=1 / 6, max_iter=cfg.max_iter, tol=cfg.tol,
                         margin=0, device=dev, bin_edges=acq.bin_edges, show_pbar=False,
                         echoes=len(cfg.echo_times_ms))
    s.simulate_kspace(acq.truth, cfg.snr_db, seed=cfg.seed + 104729)
    tuned = s.autotune(reps=2)
    for _ in range(max(args.warmup, 1)):
        s.max_iter, full = 3, s.max_iter
        s.run()
        s.max_iter = full
    torch.cuda.synchronize(dev)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        s.run()
    torch.cuda.synchronize(dev)
    t = (time.perf_counter() - t0) / args.steps
    n_it = s.iterations
    log(f"tvm48: {n_it} iterations, {t * 1e3:.1f} ms to tol ({t / n_it * 1e3:.2f} ms per iteration)", verbose)
    cpu = None
    if not args.no_cpu_baseline:
        env = dict(os.environ)
        for k in ("OMP_NUM_THREADS", "OPENBLAS_NUM_THREADS", "MKL_NUM_THREADS", "NUMEXPR_NUM_THREADS"):
            env[k] = "1"
        log("cpu baseline: one float64 oracle iteration on 1 core", verbose)
        r = subprocess.run([sys.executable, os.path.abspath(__file__), "--cpu-baseline-worker", "--config", "tvm48"],
                           capture_output=True, text=True, env=env, timeout=900)
        if r.returncode == 0:
            d = json.loads(r.stdout.strip().splitlines()[-1])
            cpu = {"value": d["per_iter"] * n_it, "unit": "s", "cores": 1, "kind": "port",
                   "s_per_iteration": d["per_iter"],
                   "sample": f"one full float64-oracle TVM iteration ({d['bins']} bins x {d['K']} channels, "
                             f"{sum(d['samples'])} samples) = {d['per_iter']:.1f} s on 1 thread, "
                             f"x{n_it} iterations"}
        else:
            log(f"cpu baseline failed: {r.stderr[-2000:]}", True)
    line = {
        "metric": "TVM time-to-tol 1e-3 at 48x64x64, 8 coils, 2 echoes, 4 bins on 1 GPU",
        "value": t, "unit": "s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": t * 1e3, "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic: seeded cone trajectory, ellipsoid phantom, Gaussian-blob coil maps, "
                                "k-space simulated by the HIP forward NUFFT + 30 dB complex noise",
        "config": {"workload": f"plain TVM PDHG to tol {cfg.tol} (max {cfg.max_iter} it), 48x64x64, "
                               f"{cfg.coils} coils, 2 echoes, {cfg.bins} bins, 1536x256 cone samples/echo",
                   "iterations": n_it, "s_per_iteration": t / max(n_it, 1),
                   "final_tol": s.history[-1] if s.history else None,
                   "launch_params": {k: tuned.get(k) for k in ("interp_xcd_chunk", "interp_kgroups",
                                                                "spread_xcd_chunk", "bin_streams")}},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))
    return 0


BF16_MFMA_PEAK_TFLOPS = 2500.0


def bench_cnn(args, verbose=True):
    """The warm-start CNN on MFMA: inference over the 36 (bin, echo) images of an 80x192x192 TVME problem
    (chunks of 6, as the warm start runs it) and full training steps (forward, loss, backward, Adam) on
    64x96x96 images; executed bf16 MFMA flops over wall time between device synchronisations."""
    import torch

    from toporecon_hip.learned import WarmStartNet, default_weights_path
    from toporecon_hip.learned.net import CIN
    dev = torch.device("cuda", 0)
    net = WarmStartNet.load(default_weights_path(), dev)
    flop_voxel = sum(2 * 27 * cin * 32 for cin in CIN)  # five 3x3x3 layers as executed (32 output lanes)
    gen = torch.Generator(device=dev)
    gen.manual_seed(3)

    def rand_input(n, shape):
        x = torch.zeros((n, *shape, 8), dtype=torch.bfloat16, device=dev)
        x[..., :2] = torch.randn((n, *shape, 2), generator=gen, device=dev).to(torch.bfloat16)
        return x

    shape, n, chunk = (80, 192, 192), 36, 6
    x = rand_input(n, shape)
    for _ in range(max(args.warmup, 1)):
        net.forward(x[:chunk], chunk, shape)
    torch.cuda.synchronize(dev)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        for i0 in range(0, n, chunk):
            net.forward(x[i0:i0 + chunk], chunk, shape)
    torch.cuda.synchronize(dev)
    t_inf = (time.perf_counter() - t0) / args.steps
    vox = n * shape[0] * shape[1] * shape[2]
    inf_tflops = flop_voxel * vox / t_inf / 1e12
    log(f"cnn inference: {n} images of {shape} in {t_inf * 1e3:.1f} ms, {inf_tflops:.0f} TFLOP/s", verbose)
    del x
    tshape, steps = (64, 96, 96), 20
    xt = rand_input(1, tshape)
    target = torch.randn((1, *tshape, 2), generator=gen, device=dev)
    # a training step: forward (4 bf16 layers + fp32 output) + input-gradient convolutions of layers
    # 4..1 + weight gradients of all five layers, each the same contraction size as a forward layer
    flop_step_voxel = flop_voxel + sum(2 * 27 * 32 * 32 for _ in range(4)) + flop_voxel
    for _ in range(2):
        out, acts = net.forward(xt, 1, tshape, keep=True)
        _, dy, _ = net.loss_grad(out, target, 1, tshape)
        net.backward(acts, dy, 1, tshape)
    torch.cuda.synchronize(dev)
    t0 = time.perf_counter()
    for _ in range(steps):
        out, acts = net.forward(xt, 1, tshape, keep=True)
        _, dy, _ = net.loss_grad(out, target, 1, tshape)
        gw, gb = net.backward(acts, dy, 1, tshape)
        net.adam_step(gw, gb, 1e-6)
    torch.cuda.synchronize(dev)
    t_step = (time.perf_counter() - t0) / steps
    tv = tshape[0] * tshape[1] * tshape[2]
    step_tflops = flop_step_voxel * tv / t_step / 1e12
    log(f"cnn training step at {tshape}: {t_step * 1e3:.2f} ms, {step_tflops:.0f} TFLOP/s", verbose)
    line = {
        "metric": "warm-start CNN inference time over the 36 (bin, echo) images of an 80x192x192 TVME problem",
        "value": t_inf, "unit": "s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": t_inf * 1e3, "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
        "dtype": "bf16", "data": "synthetic: random bf16 inputs, committed weights",
        "config": {"workload": "5-layer 3x3x3 CNN, 32 channels, bf16 activations, fp32 accumulation on "
                               "v_mfma_f32_32x32x8_bf16; inference 36 x 80x192x192 in chunks of 6; training "
                               "steps (forward, MSE, backward, Adam) on one 64x96x96 image",
                   "executed_flop_per_voxel_forward": flop_voxel,
                   "training_step_ms": t_step * 1e3, "training_step_tflops": step_tflops},
        "roofline": {"bound": "mfma", "achieved": inf_tflops, "peak": BF16_MFMA_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": inf_tflops / BF16_MFMA_PEAK_TFLOPS, "traffic": None},
        "cpu_baseline": None,
    }
    print(json.dumps(line))
    return 0


def main(argv=None):
    args = parse_args(argv)
    if args.config == "cnn":
        return bench_cnn(args)
    if args.cpu_baseline_worker and args.config == "tvm48":
        tvm48_oracle_iteration()
        return 0
    if args.config == "tvm48":
        if args.gpus != 1:
            print("bench.py measures one GPU (run with --gpus 1)", file=sys.stderr)
            return 2
        return bench_tvm48(args)
    if args.config == "batched":
        if args.gpus != 1:
            print("bench.py measures one GPU (run with --gpus 1)", file=sys.stderr)
            return 2
        return bench_batched(args)
    if args.traffic_from:
        out = write_traffic(args.traffic_from[0], args.traffic_from[1], args.copy_calibration_mib or 2048,
                            TRAFFIC_FILE)
        print(json.dumps(out["calibration"]))
        return 0
    if args.cpu_baseline_worker:
        cpu_baseline_worker(args.config)
        return 0
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: the HIP path has no multi-GPU sharding "
              "(run with --gpus 1)", file=sys.stderr)
        return 2
    verbose = True
    import numpy as np
    import torch

    from toporecon_hip import _lib, synth
    from toporecon_hip.algorithms import TvmeReconstructor

    if not torch.cuda.is_available():
        print("bench.py needs a GPU", file=sys.stderr)
        return 2
    dev = torch.device("cuda", 0)
    cfg = make_config(args.config)
    if args.max_iter is not None:
        cfg.max_iter = args.max_iter
    if args.tol is not None:
        cfg.tol = args.tol
    t0 = time.perf_counter()
    acq = synth.make_acquisition(cfg)
    log(f"synthetic acquisition {cfg.img_shape} C={cfg.coils} E={cfg.echoes} B={cfg.bins} "
        f"traj={acq.coord.shape} in {time.perf_counter() - t0:.1f} s", verbose)
    t0 = time.perf_counter()
    solver = TvmeReconstructor(None, acq.coord, acq.dcf, acq.mps, acq.resp, None, cfg.bins, cfg.l2_coupling,
                               lamda_m=cfg.lambda_motion, lamda_e=cfg.lambda_echo, sigma=1 / 6, tau=1 / 6,
                               max_iter=cfg.max_iter, tol=cfg.tol, margin=0, device=dev,
                               bin_edges=acq.bin_edges, show_pbar=False, echoes=cfg.echoes)
    solver.simulate_kspace(acq.truth, cfg.snr_db, seed=cfg.seed + 104729)
    torch.cuda.synchronize(dev)
    samples = [solver.plan.samples(b) for b in range(cfg.bins)]
    log(f"plan + simulated k-space in {time.perf_counter() - t0:.1f} s; samples/bin {samples}", verbose)
    if args.copy_calibration_mib > 0:
        n = args.copy_calibration_mib * (1 << 20) // 4
        src = torch.ones(n, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        torch.cuda.synchronize(dev)
        dst.copy_(src)
        torch.cuda.synchronize(dev)
        log(f"copy calibration: one copy of {4 * n} bytes (read {4 * n}, write {4 * n})", verbose)
        del src, dst
    t0 = time.perf_counter()
    if args.params:
        tuned = json.loads(args.params)
        for k in ("interp_xcd_chunk", "interp_kgroups", "spread_xcd_chunk", "lines_xcd_lines"):
            if k in tuned:
                solver.plan.set_param(k, int(tuned[k]))
        solver.set_bin_streams(int(tuned.get("bin_streams", 1)))
        log(f"launch parameters set: {tuned}", verbose)
    else:
        tuned = solver.autotune()
        log(f"autotune in {time.perf_counter() - t0:.1f} s: {tuned}", verbose)

    from toporecon_hip import objective
    from toporecon_hip.learned import DEFAULT_CG_STEPS, WarmStartNet, default_weights_path, warm_start
    cg_steps = DEFAULT_CG_STEPS if args.cg_steps is None else args.cg_steps
    net = None
    if args.legs == "plain":
        log("plain leg only", verbose)
    elif os.path.exists(default_weights_path()):
        net = WarmStartNet.load(default_weights_path(), dev)
    else:
        log("no trained warm-start weights: the learned leg is skipped (plain PDHG only)", verbose)

    # warmup: short solves through both paths (kernels, plans and caches warm)
    full_iter = solver.max_iter
    solver.max_iter = min(full_iter, 5)
    for w in range(args.warmup):
        solver.run()
        if net is not None:
            solver.run(init=warm_start(solver, net, cg_steps=cg_steps), dual_init="residual")
        torch.cuda.synchronize(dev)
        log(f"warmup {w} done", verbose)
    solver.max_iter = full_iter

    # plain PDHG from zeros
    torch.cuda.synchronize(dev)
    t0 = time.perf_counter()
    iters = []
    for k in range(args.steps):
        u_plain = solver.run(progress_every=25)
        iters.append(solver.iterations)
    torch.cuda.synchronize(dev)
    elapsed = time.perf_counter() - t0
    t_plain = elapsed / args.steps
    n_it = iters[-1]
    u_plain = u_plain.clone()
    phi_plain = objective.objective(solver, u_plain)
    del u_plain
    log(f"plain: {args.steps} solve(s), {n_it} iterations, {t_plain:.2f} s per solve, objective {phi_plain:.6e}",
        verbose)

    # learned warm start: dcf adjoint -> CNN (MFMA, bf16) -> data-fitted scale -> CG steps -> PDHG
    def timed_start(use_net, k_cg, steps):
        torch.cuda.synchronize(dev)
        t0 = time.perf_counter()
        for _ in range(steps):
            ta = time.perf_counter()
            u0 = warm_start(solver, net if use_net else None, cg_steps=k_cg)
            torch.cuda.synchronize(dev)
            t_start = time.perf_counter() - ta
            u_warm = solver.run(init=u0, dual_init="residual", progress_every=25)
        torch.cuda.synchronize(dev)
        t_warm = (time.perf_counter() - t0) / steps
        phi_warm = objective.objective(solver, u_warm)
        rel = (phi_warm - phi_plain) / phi_plain
        del u0, u_warm
        return {"start": ("cnn" if use_net else "zero") + f"+{k_cg}cg", "cg_steps": k_cg,
                "time_to_tol_s": t_warm, "iterations": solver.iterations, "warm_start_s": t_start,
                "objective": phi_warm, "objective_plain": phi_plain, "objective_rel_diff": rel,
                "within_0.1pct": abs(rel) <= 1e-3, "not_worse_than_plain_0.1pct": rel <= 1e-3,
                "speedup_vs_plain": t_plain / t_warm}

    warm = None
    variants = []
    if net is not None:
        warm = timed_start(True, cg_steps, args.steps)
        warm.update(weights=os.path.basename(default_weights_path()), trained_on=net.meta)
        log(f"warm start {warm['start']}: {warm['iterations']} iterations, {warm['time_to_tol_s']:.2f} s "
            f"(start {warm['warm_start_s']:.2f} s), objective {warm['objective']:.6e} "
            f"({warm['objective_rel_diff']:+.2e} vs plain), speedup {warm['speedup_vs_plain']:.2f}x", verbose)
        for v in filter(None, (x.strip() for x in args.variants.split(","))):
            kind, k_cg = v.split("+")
            r = timed_start(kind == "cnn", int(k_cg), 1)
            variants.append(r)
            log(f"variant {r['start']}: {r['iterations']} iterations, {r['time_to_tol_s']:.2f} s, "
                f"objective {r['objective_rel_diff']:+.2e} vs plain, speedup {r['speedup_vs_plain']:.2f}x", verbose)
    # step-size check: float64 power iteration for ||K|| (outside the timed legs)
    knorm = None
    if args.legs == "all":
        t0 = time.perf_counter()
        knorm = solver.operator_norm(iters=20)
        knorm = {k: knorm[k] for k in ("K_norm", "sigma_tau_K_norm_sq", "iterations")}
        log(f"power iteration: ||K|| = {knorm['K_norm']:.4f}, sigma tau ||K||^2 = "
            f"{knorm['sigma_tau_K_norm_sq']:.4f} ({time.perf_counter() - t0:.1f} s)", verbose)
    use_warm = warm is not None and warm["within_0.1pct"]
    t_step = warm["time_to_tol_s"] if use_warm else t_plain

    final_tol = solver.history[-1] if solver.history else None
    # per-kernel times: a short serial pass (bins on one stream), so that no two kernels overlap
    # and each HIP-event interval is one kernel's own duration
    streams_used = len(getattr(solver, "_bstreams", None) or [None])
    solver.set_bin_streams(1)
    solver.max_iter, prof_iters = min(3, full_iter), min(3, full_iter)
    solver.tol = 0.0
    _lib.profile_enable(True)
    solver.run()
    torch.cuda.synchronize(dev)
    prof = _lib.profile_read()
    _lib.profile_enable(False)
    solver.max_iter, solver.tol = full_iter, cfg.tol
    solver.set_bin_streams(streams_used)

    kernels = {}
    for name, (ms, n, by) in prof.items():
        kernels[name] = {"launches": n, "ms_avg": ms / n, "ms_total": ms,
                         "gbps": by / (ms * 1e-3) / 1e9 if ms > 0 else None,
                         "bytes_per_launch": by / n}
    dom = max(kernels, key=lambda k: kernels[k]["ms_total"])
    dk = kernels[dom]
    traffic = None
    if os.path.exists(TRAFFIC_FILE):
        with open(TRAFFIC_FILE) as f:
            traffic = json.load(f).get("kernels", {}).get(dom, {}).get("bytes")
    roofline = {"bound": "hbm", "kernel": dom, "achieved": dk["gbps"], "peak": HBM_PEAK_GBPS,
                "unit": "GB/s", "frac": dk["gbps"] / HBM_PEAK_GBPS, "traffic": traffic,
                "algorithmic_bytes_per_launch": dk["bytes_per_launch"], "avg_launch_ms": dk["ms_avg"]}
    # the two gathers take about the same time at this size, so the dominant one can flip between
    # runs: both are reported
    gathers = {}
    all_traffic = {}
    if os.path.exists(TRAFFIC_FILE):
        with open(TRAFFIC_FILE) as f:
            all_traffic = json.load(f).get("kernels", {})
    for k in ("interp", "spread"):
        if k in kernels and kernels[k]["gbps"]:
            gathers[k] = {"achieved": kernels[k]["gbps"], "frac": kernels[k]["gbps"] / HBM_PEAK_GBPS,
                          "avg_launch_ms": kernels[k]["ms_avg"],
                          "algorithmic_bytes_per_launch": kernels[k]["bytes_per_launch"],
                          "traffic": all_traffic.get(k, {}).get("bytes")}
    roofline["gathers"] = gathers
    fwd_b, adj_b = nufft_model_bytes(samples, cfg.echoes, cfg.coils, cfg.img_shape,
                                      solver.plan.grid_shape)
    fwd_ms = sum(kernels[k]["ms_total"] for k in ("fwd_x_pad", "fft_y_fwd", "fft_z_fwd", "interp")
                 if k in kernels)
    adj_ms = sum(kernels[k]["ms_total"] for k in ("spread", "fft_z_inv", "fft_y_inv", "adj_x_crop")
                 if k in kernels)
    total_iters = prof_iters
    nufft = {"forward_gbps": fwd_b * total_iters / (fwd_ms * 1e-3) / 1e9 if fwd_ms else None,
             "adjoint_gbps": adj_b * total_iters / (adj_ms * 1e-3) / 1e9 if adj_ms else None,
             "forward_ms_per_iter": fwd_ms / total_iters, "adjoint_ms_per_iter": adj_ms / total_iters}

    cpu = None
    if not args.no_cpu_baseline:
        cpu = run_cpu_baseline(args.config, n_it, verbose)

    shape = "x".join(str(v) for v in cfg.img_shape)
    out = {
        "metric": METRIC,
        "value": t_step,
        "unit": "s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": t_step * 1e3,
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: seeded cone trajectory, ellipsoid phantom, Gaussian-blob coil maps, "
                "k-space simulated by the HIP forward NUFFT + 30 dB complex noise",
        "config": {"workload": f"TVME PDHG to tol {cfg.tol} (max {cfg.max_iter} it) "
                               f"{f'from the learned CNN + {cg_steps} CG warm start' if use_warm else 'from zeros (plain)'}, "
                               f"{shape}, {cfg.coils} coils, {cfg.echoes} echoes, {cfg.bins} bins, "
                               f"{cfg.interleaves}x{cfg.readouts} cone samples/echo, "
                               f"l2 coupling, lambda {cfg.lambda_motion}",
                   "start": "learned" if use_warm else "plain",
                   "plain": {"time_to_tol_s": t_plain, "iterations": n_it, "objective": phi_plain,
                             "s_per_iteration": t_plain / max(n_it, 1)},
                   "learned_warm_start": warm,
                   "warm_start_variants": variants,
                   "operator_norm": knorm,
                   "iterations": warm["iterations"] if use_warm else n_it,
                   "s_per_iteration": t_plain / max(n_it, 1),
                   "samples_per_bin": samples, "grid": list(solver.plan.grid_shape),
                   "launch_params": {k: tuned.get(k) for k in ("interp_xcd_chunk", "interp_kgroups",
                                                                "spread_xcd_chunk", "bin_streams")},
                   "final_tol": final_tol, "kernel_profile": f"serial pass, {prof_iters} iterations, "
                                                             "bins on one stream"},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "nufft": nufft,
        "kernels": kernels,
    }
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
