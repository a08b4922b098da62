#!/usr/bin/env python3
"""TVME time-to-tolerance benchmark of the HIP PDHG hot path on one MI355X.

Workload (default): TVME PDHG to tol 1e-3 (max 300 iterations, sigma = tau =
1/6, lambda_motion = lambda_echo = 1e-5, L2 echo coupling) at 80x192x192,
32 coils, 6 echoes (TE = 1.0 + 1.3 e ms), 6 motion bins, 9000 cone
interleaves x 512 samples per echo (readout fraction 0.98, 2/98 % respiratory
outlier drop), synthetic seeded data simulated on the GPU.  Inputs are
resident in HBM when a timed region starts.

A step is one time-to-tolerance solve of the headline leg: the learned warm
start (density-compensated adjoint -> 3-D CNN on MFMA -> per-(bin, echo)
data-fitted scale -> 4 conjugate-gradient data-consistency steps) followed by
PDHG to the tolerance, the warm-start work inside the timed region.  The
plain leg (PDHG from zeros, as the reference runs it) is timed once as the
denominator.  ``value`` is the learned leg's mean time-to-tol over the timed
steps when its final objective is within +-0.1 % of the plain leg's (plain
PDHG is the fallback otherwise, and then ``value`` is the plain time).
Warmup steps are learned solves capped at a few PDHG iterations.  Variant
starts (``--variants``) are off by default.  A wall-clock guard stops the
timed steps early (and records how many ran) if the run would pass
``--budget-s`` seconds since the process started.

Per-kernel times come from HIP events that the library records on each
kernel's own launch stream during the timed steps (bins on one stream, so no
two kernels overlap); ``roofline`` reports the dominant kernel against the
8 TB/s HBM peak on the algorithmic-byte model (per-unit bytes x units per
launch) and, labelled separately, on the bytes the kernel must move after
k-space support pruning.  ``traffic`` is replayed from the counter profile
named in ``traffic_source`` (two rocprofv3 --pmc runs of this bench).

The CPU baseline is the single-precision NumPy restatement of the reference's
``--device -1`` path (oracle/sigpy32.py: SigPy's nufft / nufft_adjoint with
numpy.fft, one thread), timed on one (coil, echo) channel of every motion bin
plus a z-slab of the TV element-wise work, extrapolated to a full iteration
and to the plain solve's iteration count; and the same channels spread over
all host cores.
"""

from __future__ import annotations

import argparse
import json
import math
import os
import subprocess
import sys
import tempfile
import time

T_PROCESS0 = time.perf_counter()

REPO = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(REPO, "multinodemri-miccai26_amd")
for _p in (REPO, PKG):
    if _p not in sys.path:
        sys.path.insert(0, _p)

METRIC = "TVME time-to-tol 1e-3 (learned warm start vs plain) on 1 GPU; NUFFT GB/s"
HBM_PEAK_GBPS = 8000.0
TRAFFIC_FILE = os.path.join(REPO, "profiles", "traffic.json")
WARMUP_ITERS = 3          # PDHG iterations of a warmup solve
DEFAULT_BUDGET_S = 450.0  # wall-clock guard from process start (the driver kills at 600 s)
CPU_BUDGET_S = 45.0       # single-core CPU-baseline sample
EW_SLAB_Z = 8             # z planes of the element-wise CPU sample


def parse_args(argv=None):
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", choices=("full", "small", "batched", "tvm48", "cnn", "jsense", "tvmw"),
                    default="full",
                    help="full: TVME 80x192x192 (the headline); small: a reduced TVME; batched: 12 uneven "
                         "TVM instances, HIP-graph batch vs sequential solves; tvm48: plain TVM at "
                         "48x64x64, 8 coils, 2 echoes, 4 bins beside one full single-precision CPU "
                         "iteration; cnn: the warm-start network's MFMA convolutions; jsense: coil-map "
                         "estimation on echo 0 of the headline acquisition; tvmw: TVMW to tolerance at the "
                         "headline acquisition")
    ap.add_argument("--max-iter", type=int, default=None, help="override max_iter (profiling runs)")
    ap.add_argument("--tol", type=float, default=None)
    ap.add_argument("--budget-s", type=float, default=DEFAULT_BUDGET_S,
                    help="stop the timed steps early when the run would pass this many seconds")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-baseline-worker", default=None, help=argparse.SUPPRESS)
    ap.add_argument("--no-timed-events", action="store_true",
                    help="no per-kernel HIP events in the timed steps (no roofline)")
    ap.add_argument("--verbose", action="store_true")
    ap.add_argument("--traffic-from", nargs=2, metavar=("FETCH_CSV", "WRITE_CSV"),
                    help="summarize two rocprofv3 --pmc counter_collection.csv files (FETCH_SIZE, "
                         "WRITE_SIZE) of a run with --copy-calibration-mib into profiles/traffic.json")
    ap.add_argument("--traffic-command", default=None,
                    help="with --traffic-from: the bench command the counters were collected on")
    ap.add_argument("--params", default=None,
                    help="JSON launch parameters (interp_xcd_chunk, interp_kgroups, spread_xcd_chunk, "
                         "lines_xcd_lines) to set instead of autotuning (profiling runs)")
    ap.add_argument("--legs", choices=("all", "plain"), default="all",
                    help="plain: only the plain PDHG leg, timed --steps times (no warm start)")
    ap.add_argument("--cg-steps", type=int, default=None,
                    help="CG steps of the learned leg (default: the warm-start module's default, 4)")
    ap.add_argument("--variants", default="",
                    help="comma list of extra warm starts 'cnn+K' / 'zero+K' (K CG steps) timed once each "
                         "and reported beside the learned leg, e.g. cnn+0,zero+4,cnn+3,cnn+5")
    ap.add_argument("--geometry", default="default",
                    help="sample geometry of the plan: 'default' (sort tile 8, segments of 1024 pairs), "
                         "'tune' (the autotuner picks both) or 'TILExSEGMENT', e.g. 4x256")
    ap.add_argument("--copy-calibration-mib", type=int, default=0,
                    help="before the solve, copy a buffer of this many MiB once (a kernel of known "
                         "traffic for scaling hardware byte counters)")
    return ap.parse_args(argv)


def log(msg, verbose=True):
    if verbose:
        print(f"[bench {time.perf_counter() - T_PROCESS0:6.1f}s] {msg}", file=sys.stderr, flush=True)


def work_plan(args, variants=None):
    """What a run of the full configuration does, in solves (used by main and checked by a test).

    warmup_solves: learned solves capped at WARMUP_ITERS iterations; plain_solves: 1 (all with
    --legs plain: --steps); learned_solves: --steps (0 with --legs plain); variant_solves: one per
    --variants entry; power_iterations: the ||K|| estimate; cpu: single-core sample bounded by
    CPU_BUDGET_S plus one concurrent round on all host cores."""
    variants = [v for v in (x.strip() for x in (args.variants or "").split(",")) if v] if variants is None \
        else variants
    plain_only = args.legs == "plain"
    return {"warmup_solves": args.warmup, "warmup_iterations": WARMUP_ITERS,
            "plain_solves": args.steps if plain_only else 1,
            "learned_solves": 0 if plain_only else args.steps,
            "variant_solves": 0 if plain_only else len(variants),
            "power_iterations": 0 if plain_only else 20,
            "cpu_single_core_budget_s": 0.0 if args.no_cpu_baseline else CPU_BUDGET_S,
            "cpu_all_cores_rounds": 0 if args.no_cpu_baseline else 1,
            "budget_s": args.budget_s}


def estimate_seconds(plan, *, t_setup, t_iter, t_start, n_plain, n_learned, t_cpu_round):
    """Wall-clock estimate of a run: t_setup (imports, data, plan, autotune), t_iter (one PDHG
    iteration), t_start (the warm-start work), iteration counts of the two legs, t_cpu_round (one
    concurrent all-cores CPU round); a power iteration costs about two PDHG iterations."""
    learned = t_start + n_learned * t_iter
    return (t_setup + plan["warmup_solves"] * (t_start + plan["warmup_iterations"] * t_iter)
            + plan["plain_solves"] * n_plain * t_iter
            + (plan["learned_solves"] + plan["variant_solves"]) * learned
            + plan["power_iterations"] * 2 * t_iter
            + plan["cpu_single_core_budget_s"] + plan["cpu_all_cores_rounds"] * t_cpu_round)


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
    """Algorithmic bytes of one forward and one adjoint pass over all bins (every grid pass moves the
    whole oversampled grid), for the NUFFT GB/s figures:

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


def model_bytes_per_launch(name, M, E, C, N, G, B):
    """Algorithmic bytes of one launch of a logical kernel on the per-unit model (each input read once,
    each output written once, every grid pass over the whole oversampled grid).  M: samples of the
    launch's bin; the line passes are one 1-D FFT pass over the grid (read + write)."""
    K = E * C
    return {
        "fwd_x_pad": 8 * E * N + 8 * C * N + 8 * K * G,
        "fft_y_fwd": 16 * K * G, "fft_z_fwd": 16 * K * G, "fft_z_inv": 16 * K * G, "fft_y_inv": 16 * K * G,
        "interp": 8 * K * G + 12 * M + 24 * K * M,
        "spread": 8 * K * M + 16 * M + 8 * K * G,
        "adj_x_crop": 8 * K * N + 8 * C * N + 8 * E * N,
        "tvme_dual": 9 * 8 * B * E * N, "tvme_primal": 8 * 8 * B * E * N,
    }.get(name)


# ---------------------------------------------------------------------------------------------
# CPU baseline (single-precision NumPy restatement of the reference's --device -1 path)
# ---------------------------------------------------------------------------------------------

def _cpu_worker(job_path):
    """Time the jobs of one CPU-baseline worker; prints one JSON line."""
    import numpy as np

    from oracle import sigpy32 as s32
    with open(job_path) as f:
        job = json.load(f)
    if job.get("core") is not None:
        os.sched_setaffinity(0, {int(job["core"])})
    d = job["dir"]
    shape = tuple(job["shape"])
    rng = np.random.default_rng(job.get("seed", 0))
    while time.time() < job.get("start_at", 0.0):
        time.sleep(0.01)
    t_begin = time.perf_counter()
    out = {"channels": [], "ew_slab_s": None}
    for ch in job["channels"]:
        if time.perf_counter() - t_begin > job.get("budget_s", 1e9):
            break
        b, c, e = ch["bin"], ch["coil"], ch["echo"]
        coord = np.load(os.path.join(d, f"coord_{b}.npy"), mmap_mode="r")
        dcf = np.load(os.path.join(d, f"dcf_{b}.npy"))
        ub = np.load(os.path.join(d, f"img_{b}_{e}.npy"))
        mp = np.load(os.path.join(d, f"map_{c}.npy"))
        M = coord.shape[0]
        y = (rng.standard_normal(M, dtype=np.float32) + 1j * rng.standard_normal(M, dtype=np.float32)) \
            .astype(np.complex64) * np.float32(1e-3)
        q = np.zeros(M, dtype=np.complex64)
        g = np.zeros(shape, dtype=np.complex64)
        coord = np.ascontiguousarray(coord)
        t0 = time.perf_counter()
        s32.data_channel(ub, mp, y, q, dcf, coord, g)
        out["channels"].append({"bin": b, "coil": c, "echo": e, "samples": int(M),
                                "seconds": time.perf_counter() - t0})
    if job.get("ew_slab"):
        B, E = job["bins"], job["echoes"]
        sh = (B, E, min(EW_SLAB_Z, shape[0]), shape[1], shape[2])

        def rnd():
            return (rng.standard_normal(sh, dtype=np.float32)
                    + 1j * rng.standard_normal(sh, dtype=np.float32)).astype(np.complex64)
        u, ub, g = rnd(), rnd(), rnd()
        z = np.zeros(sh, dtype=np.complex64)
        t0 = time.perf_counter()
        s32.tvme_elementwise(u, ub, g, z, z, z, z, job["lam_m"], job["lam_e"], True)
        out["ew_slab_s"] = time.perf_counter() - t0
        out["ew_slab_z"] = sh[2]
    print(json.dumps(out))


def _cpu_env():
    env = dict(os.environ)
    for k in ("OMP_NUM_THREADS", "OPENBLAS_NUM_THREADS", "MKL_NUM_THREADS", "NUMEXPR_NUM_THREADS"):
        env[k] = "1"
    return env


def _spawn_worker(job, tmpdir, name):
    path = os.path.join(tmpdir, f"{name}.json")
    with open(path, "w") as f:
        json.dump(job, f)
    return subprocess.Popen([sys.executable, os.path.abspath(__file__), "--cpu-baseline-worker", path],
                            stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True, env=_cpu_env())


def _cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def run_cpu_baseline(acq, cfg, n_plain, verbose, deadline):
    """The single-precision CPU path on one channel of every bin (1 pinned core, bounded by
    CPU_BUDGET_S) and on W concurrent channels (all host cores), extrapolated to the plain solve."""
    import numpy as np
    B, E, C = cfg.bins, cfg.echoes, cfg.coils
    K = E * C
    Z = cfg.img_shape[0]
    cores = sorted(os.sched_getaffinity(0))
    W = max(1, min(16, len(cores)))
    tmp = tempfile.mkdtemp(prefix="tr_cpu_")
    chans = [{"bin": b, "coil": (7 * b + 3) % C, "echo": b % E} for b in range(B)]
    par = [{"bin": w % B, "coil": (5 * w + 1) % C, "echo": (w // B + w) % E} for w in range(W)]
    try:
        for b in range(B):
            np.save(os.path.join(tmp, f"coord_{b}.npy"), acq.coord[acq.bin_index(b)].reshape(-1, 3))
            np.save(os.path.join(tmp, f"dcf_{b}.npy"), acq.dcf[acq.bin_index(b)].reshape(-1))
        for ch in chans + par:
            p = os.path.join(tmp, f"img_{ch['bin']}_{ch['echo']}.npy")
            if not os.path.exists(p):
                np.save(p, acq.truth[ch["bin"], ch["echo"]])
            p = os.path.join(tmp, f"map_{ch['coil']}.npy")
            if not os.path.exists(p):
                np.save(p, acq.mps[ch["coil"]])
        base = {"dir": tmp, "shape": list(cfg.img_shape), "bins": B, "echoes": E,
                "lam_m": cfg.lambda_motion, "lam_e": cfg.lambda_echo}
        log(f"cpu baseline: single-precision CPU path, 1 pinned core, one channel of each of {B} bins", verbose)
        pr = _spawn_worker(dict(base, channels=chans, core=cores[0], ew_slab=True, budget_s=CPU_BUDGET_S),
                           tmp, "single")
        so, se = pr.communicate(timeout=CPU_BUDGET_S + 240)
        if pr.returncode != 0:
            log(f"cpu baseline failed: {se[-2000:]}", True)
            return None
        d1 = json.loads(so.strip().splitlines()[-1])
        done = d1["channels"]
        t_ch = sum(x["seconds"] for x in done) / len(done)
        ew_iter = d1["ew_slab_s"] * Z / d1["ew_slab_z"]
        per_iter = B * K * t_ch + ew_iter
        bins_done = sorted({x["bin"] for x in done})
        out = {"value": per_iter * n_plain, "unit": "s", "cores": 1, "kind": "port",
               "s_per_iteration": per_iter, "iterations": n_plain, "cpu_model": _cpu_model(),
               "sample": (f"single-precision NumPy restatement of the reference's --device -1 path "
                          f"(oracle/sigpy32.py, numpy.fft, 1 thread pinned to one core): the data term of one "
                          f"(coil, echo) channel in each of bins {bins_done} "
                          f"({', '.join('%.2f' % x['seconds'] for x in done)} s; forward NUFFT, dual prox, "
                          f"weighted adjoint, coil accumulation over {done[0]['samples']} samples) and the TVME "
                          f"element-wise steps on a {d1['ew_slab_z']}-plane z-slab ({d1['ew_slab_s']:.2f} s); "
                          f"per iteration = {B} bins x {K} channels x {t_ch:.2f} s + {ew_iter:.1f} s = "
                          f"{per_iter:.0f} s, x{n_plain} iterations of the plain solve "
                          f"(extrapolation factor {B * K * n_plain / len(done):.0f} on the sampled channel time)")}
        if time.perf_counter() - T_PROCESS0 + 2.5 * t_ch + 10 > deadline:
            log("cpu baseline: no time left for the all-cores round", verbose)
            return out
        log(f"cpu baseline: {W} concurrent single-threaded workers, one channel each", verbose)
        start_at = time.time() + 6.0
        procs = [_spawn_worker(dict(base, channels=[par[w]], core=None, ew_slab=False, start_at=start_at,
                                    seed=w), tmp, f"par{w}") for w in range(W)]
        res = []
        for p in procs:
            so, se = p.communicate(timeout=600)
            if p.returncode == 0:
                res.append(json.loads(so.strip().splitlines()[-1]))
        if len(res) == W and all(r["channels"] for r in res):
            t_round = max(r["channels"][0]["seconds"] for r in res)
            per_iter_all = B * K / W * t_round + ew_iter / W
            out["all_cores"] = {"value": per_iter_all * n_plain, "unit": "s", "cores": W, "kind": "port",
                                "s_per_iteration": per_iter_all,
                                "sample": (f"{W} single-threaded processes started together, each running the "
                                           f"data term of a different (bin, coil, echo) channel; the round took "
                                           f"{t_round:.2f} s (slowest worker); per iteration = {B * K} channels "
                                           f"/ {W} x {t_round:.2f} s + the element-wise steps / {W}")}
        return out
    finally:
        import shutil
        shutil.rmtree(tmp, ignore_errors=True)


# ---------------------------------------------------------------------------------------------
# hardware counters -> profiles/traffic.json
# ---------------------------------------------------------------------------------------------

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


def write_traffic(fetch_csv, write_csv, copy_mib, out_path, command=None):
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
           "command": command, "kernels": kernels}
    with open(out_path, "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)
    return out


# ---------------------------------------------------------------------------------------------
# other configurations
# ---------------------------------------------------------------------------------------------

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
        tuned = s.autotune(reps=1, bin_streams=False, geometry=True)  # batches keep one stream per instance
        solvers.append(s)
        rows.append({"seed": cfg.seed, "shape": list(cfg.img_shape), "coils": cfg.coils,
                     "echoes": len(cfg.echo_times_ms), "bins": cfg.bins,
                     "samples_per_echo": int(acq.coord.shape[0] * acq.coord.shape[1]),
                     "launch_params": {k: tuned.get(k) for k in ("interp_xcd_chunk", "interp_kgroups",
                                                                   "spread_xcd_chunk", "sort_tile",
                                                                   "spread_segment")}})
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
    modes = ("persistent", "graph", "streams")
    for mode in modes:
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
        if mode == "persistent":
            legs[mode]["replays"] = batch.replays
            legs[mode]["timing"] = ("wall clock around run() (prepare, one capture, replays); per-instance "
                                    "seconds from device clock stamps when each instance's gate closed")
        log(f"{mode}: {tb:.2f} s ({batch.captures} graph captures), identical results: {same}", verbose)
    its = sum(r["sequential"]["iterations"] for r in rows)
    for r in rows:
        for leg in ("sequential",) + modes:
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


def _tvm48_cpu_iteration():
    """One full single-precision CPU TVM iteration at 48x64x64 (every bin, coil and echo), 1 thread."""
    import numpy as np

    from oracle import pdhg as op
    from oracle import sigpy32 as s32
    from toporecon_hip import synth
    cfg = synth.config_small_tvm()
    acq = synth.make_acquisition(cfg)
    ksp = synth.simulate_kspace_numpy(acq, lambda x, c: np.stack(
        [np.stack([s32.nufft(x[e, k], c) for k in range(x.shape[1])]) for e in range(x.shape[0])]))
    prob = op.bin_problem(ksp, acq.coord, acq.dcf, acq.mps.astype(np.complex64), acq.resp, acq.bin_edges)
    shape = (prob.B, prob.E) + prob.img_shape
    st = {"u": np.zeros(shape, np.complex64), "ubar": np.zeros(shape, np.complex64),
          "p_m": np.zeros(shape, np.complex64),
          "q": [np.zeros(k.shape, np.complex64) for k in prob.ksp]}
    prob.ksp = [k.astype(np.complex64) for k in prob.ksp]
    prob.dcf = [d.astype(np.float32) for d in prob.dcf]
    t0 = time.perf_counter()
    s32.tvm_iteration(prob, st, cfg.lambda_motion, cfg.l2_coupling)
    print(json.dumps({"per_iter": time.perf_counter() - t0, "bins": prob.B, "K": prob.E * prob.C,
                      "samples": [int(k.shape[-1]) for k in prob.ksp]}))


FP32_PEAK_TFLOPS = 157.3  # MI355X published dense fp32 peak


def _jsense_cpu_sample(job_path):
    """Time the float64 JSENSE oracle's pieces on a calibration problem of the benchmark's shape
    (random values; the cost depends on shapes only); prints one JSON line."""
    import numpy as np

    from oracle import jsense as oj
    from oracle import nufft as on
    with open(job_path) as f:
        job = json.load(f)
    if job.get("core") is not None:
        os.sched_setaffinity(0, {int(job["core"])})
    rng = np.random.default_rng(0)
    C, W, grd, M, frac = job["C"], job["W"], tuple(job["grd"]), job["M"], job["nufft_fraction"]
    I = tuple(g + W - 1 for g in grd)

    def cr(*shape):
        return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    img, ker, z = cr(*I), cr(C, W, W, W), cr(C, *grd)
    t = {}
    t0 = time.perf_counter()
    oj.conv_valid(img[None], ker)
    t["conv"] = time.perf_counter() - t0
    t0 = time.perf_counter()
    oj.conv_valid_adj_img(z, ker, I)
    t["conv_adj_img"] = time.perf_counter() - t0
    t0 = time.perf_counter()
    oj.conv_valid_adj_ker(z, img, ker.shape)
    t["conv_adj_ker"] = time.perf_counter() - t0
    t0 = time.perf_counter()
    oj.ifft_c(oj.fft_c(z))
    t["fft_pair"] = time.perf_counter() - t0
    ms = max(1, int(M * frac))
    coord = rng.uniform(-11.9, 11.9, (ms, 3)).astype(np.float32)
    t0 = time.perf_counter()
    k = on.nufft(z, coord)
    t["nufft"] = (time.perf_counter() - t0) * M / ms
    t0 = time.perf_counter()
    on.nufft_adjoint(k, coord, grd)
    t["nufft_adjoint"] = (time.perf_counter() - t0) * M / ms
    full = cr(*job["img_shape"])
    t0 = time.perf_counter()
    oj.ifft_c(full)
    t["map_fft"] = time.perf_counter() - t0
    print(json.dumps({"t": t, "nufft_samples_timed": ms}))


def bench_jsense(args, verbose=True):
    """JSENSE coil-map estimation at the benchmark's acquisition (80x192x192, 32 coils, 9000 x 512 cone
    samples of echo 0, readout fraction 0.95, legacy_max coordinates) with SigPy's JsenseRecon defaults
    (map kernels 16^3, calibration width 24, 10 x 10 CG steps).  A step is one estimation from
    calibration data resident in HBM: the alternating minimisation and the output maps at the image
    shape (JsenseHip.run + maps); the host part (calibration selection, upload, plan) and the full
    host-to-host call are reported beside it."""
    import numpy as np
    import torch

    from toporecon_hip import _lib, synth
    from toporecon_hip.data.trajectory import prepare_trajectory
    from toporecon_hip.jsense import JsenseHip
    from toporecon_hip.nufft_backend import HipNufftBackend
    dev = torch.device("cuda", 0)
    cfg = synth.config_full_tvme()
    shape = tuple(cfg.img_shape)
    coord = synth.cone_trajectory(cfg.interleaves, cfg.readouts, shape, cfg.turns, cfg.cone_angle)
    raw_xyz = np.ascontiguousarray(coord[..., ::-1] / np.asarray(shape[::-1], np.float32))
    traj = prepare_trajectory(raw_xyz, synth.density(coord), image_shape_zyx=shape, voxel_size_zyx=(1, 1, 1),
                              readout_fraction=0.95, scaling="legacy_max")
    c = traj.coordinates_zyx
    maps = synth.coil_maps(cfg.coils, shape, cfg.seed)
    truth = synth.phantom(shape, cfg.echo_times_ms[:1], 1, cfg.seed, cfg.ellipsoids)[0]
    be = HipNufftBackend(shape, [c], dev)
    ksp = be.forward_batch(torch.from_numpy(truth).to(dev), torch.from_numpy(maps).to(dev), 0)
    ksp = ksp.reshape(1, cfg.coils, *c.shape[:2]).cpu().numpy()
    be.plan.close()
    del be
    log(f"jsense: k-space {ksp.shape}, trajectory {c.shape}", verbose)

    def one(profile=False):
        t0 = time.perf_counter()
        s = JsenseHip(ksp[0], c, traj.density, dev)
        torch.cuda.synchronize(dev)
        t1 = time.perf_counter()
        if profile:
            _lib.profile_enable(True)
        s.run()
        m = s.maps(shape)
        torch.cuda.synchronize(dev)
        t2 = time.perf_counter()
        prof = _lib.profile_read() if profile else None
        if profile:
            _lib.profile_enable(False)
        host = m.cpu().numpy()
        t3 = time.perf_counter()
        info = {"M": s.M, "grd": list(s.grd_shape), "img_shape": list(s.img_shape), "C": s.C, "W": s.W,
                "objective": s.objective(), "max_iter": s.max_iter, "max_inner_iter": s.max_inner_iter}
        s.close()
        return t1 - t0, t2 - t1, t3 - t0, prof, info, host

    for _ in range(max(args.warmup, 1)):
        one()
    dev_t, host_t, e2e_t = [], [], []
    host_maps = None
    for _ in range(args.steps):
        th, td, te, _, info, host_maps = one()
        dev_t.append(td)
        host_t.append(th)
        e2e_t.append(te)
    _, _, _, prof, info, m_prof = one(profile=True)
    assert np.array_equal(m_prof, host_maps), "JSENSE maps differ between runs"
    t_dev = sum(dev_t) / len(dev_t)
    log(f"jsense: device {t_dev:.3f} s, host prep {sum(host_t) / len(host_t):.3f} s, "
        f"host-to-host {sum(e2e_t) / len(e2e_t):.3f} s; calibration grid {info['grd']}, {info['M']} samples",
        verbose)
    C, W = info["C"], info["W"]
    Gn = math.prod(info["grd"])
    conv_flops = 8.0 * C * Gn * W ** 3
    kernels = {}
    for name, (ms, n, by) in prof.items():
        e = {"launches": n, "ms_avg": ms / n, "ms_total": ms}
        if name.startswith("js_conv"):
            e["tflops"] = conv_flops / (ms / n * 1e-3) / 1e12
        kernels[name] = e
    dom = max(kernels, key=lambda k: kernels[k]["ms_total"])
    dk = kernels[dom]
    if dom.startswith("js_conv"):
        roofline = {"bound": "fp32", "kernel": dom, "achieved": dk["tflops"], "peak": FP32_PEAK_TFLOPS,
                    "unit": "TFLOP/s", "frac": dk["tflops"] / FP32_PEAK_TFLOPS, "traffic": None,
                    "flops_model": "8 C G W^3 per launch (a complex multiply-add per coil, calibration-grid "
                                   "point and tap; C coils, G calibration-grid points, W^3 taps)",
                    "algorithmic_flops_per_launch": conv_flops, "avg_launch_ms": dk["ms_avg"],
                    "launches": dk["launches"],
                    "note": "FP32 FMA on the vector ALUs without packed FP32, whose ceiling is half the "
                            "published fp32 peak used here"}
    else:
        # the calibration NUFFT: C coil images (no maps), one bin of M samples on the small grid
        g = [math.ceil(1.25 * v) for v in info["grd"]]
        model = model_bytes_per_launch(dom, info["M"], C, 1, Gn, math.prod(g), 1)
        achieved = model / (dk["ms_avg"] * 1e-3) / 1e9 if model else None
        roofline = {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                    "frac": achieved / HBM_PEAK_GBPS if achieved else None, "traffic": None,
                    "algorithmic_bytes_per_launch": model, "avg_launch_ms": dk["ms_avg"],
                    "launches": dk["launches"],
                    "note": f"calibration NUFFT: K = {C} channels, M = {info['M']} samples on a "
                            f"{'x'.join(map(str, g))} grid; every sample's 64 taps land in a grid that fits in "
                            f"cache, so the kernel is bound by the per-tap work, not by HBM"}
    cpu = None
    if not args.no_cpu_baseline:
        core = sorted(os.sched_getaffinity(0))[0]
        job = {"core": core, "C": C, "W": W, "grd": info["grd"], "M": info["M"], "nufft_fraction": 0.125,
               "img_shape": info["img_shape"]}
        with tempfile.TemporaryDirectory() as td:
            path = os.path.join(td, "jsense_job.json")
            with open(path, "w") as f:
                json.dump(job, f)
            log("cpu baseline: float64 JSENSE oracle pieces on 1 pinned core", verbose)
            r = subprocess.run([sys.executable, os.path.abspath(__file__), "--cpu-baseline-worker",
                                "jsense:" + path], capture_output=True, text=True, env=_cpu_env(), timeout=900)
        if r.returncode == 0:
            d = json.loads(r.stdout.strip().splitlines()[-1])["t"]
            data = d["nufft"] + d["nufft_adjoint"] + d["fft_pair"]
            n_mps = info["max_inner_iter"] + 1  # SigPy's CG applies the normal operator once before its steps
            per_round = (data + d["conv_adj_ker"] + n_mps * (d["conv"] + data + d["conv_adj_ker"])
                         + data + d["conv_adj_img"] + n_mps * (d["conv"] + data + d["conv_adj_img"]))
            total = info["max_iter"] * per_round + 3 * C * d["map_fft"]
            cpu = {"value": total, "unit": "s", "cores": 1, "kind": "port", "cpu_model": _cpu_model(),
                   "pieces_s": d,
                   "sample": ("the float64 oracle (oracle/jsense.py, oracle/nufft.py) on 1 pinned thread: one "
                              "valid convolution and each adjoint at the full shape, the centred FFT pair, the "
                              "NUFFT pair on 1/8 of the calibration samples (x8), one map-size inverse FFT; "
                              "projected to SigPy's schedule: per round two least-squares solves, each A^H y "
                              "plus (inner + 1) normal-operator applications, then 3 map-size FFTs per coil")}
        else:
            log(f"cpu baseline failed: {r.stderr[-2000:]}", True)
    line = {
        "metric": "JSENSE coil-map estimation time (SigPy JsenseRecon defaults) on 1 GPU",
        "value": t_dev, "unit": "s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": t_dev * 1e3, "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: seeded cone trajectory, ellipsoid phantom, Gaussian-blob coil maps, echo-0 k-space "
                "simulated by the HIP forward NUFFT (no noise)",
        "config": {"workload": (f"JSENSE on echo 0 of {'x'.join(map(str, shape))}, {cfg.coils} coils, "
                                f"{cfg.interleaves}x{cfg.readouts} cone samples (readout fraction 0.95, legacy_max "
                                f"coordinates): map kernels {W}^3, calibration width 24, "
                                f"{info['max_iter']} x {info['max_inner_iter']} CG steps"),
                   "step": "JsenseHip.run + maps at the image shape, calibration data resident in HBM",
                   "calibration_grid": info["grd"], "calibration_samples": info["M"],
                   "host_prep_s": sum(host_t) / len(host_t),
                   "host_to_host_s": sum(e2e_t) / len(e2e_t),
                   "final_objective": info["objective"],
                   "run_wall_s": time.perf_counter() - T_PROCESS0},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "kernels": kernels,
    }
    print(json.dumps(line), flush=True)
    return 0


def bench_tvm48(args, verbose=True):
    """Plain TVM to tol 1e-3 at 48x64x64 (8 coils, 2 echoes, 4 bins, 1536x256 cone samples) on the GPU,
    beside one full single-precision CPU iteration on one host core, scaled to the GPU's
    iteration count."""
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
        log("cpu baseline: one single-precision CPU iteration on 1 pinned core", verbose)
        core = sorted(os.sched_getaffinity(0))[0]
        r = subprocess.run([sys.executable, "-c", "import os, sys; os.sched_setaffinity(0, {%d}); "
                            "sys.argv = ['bench.py', '--cpu-baseline-worker', 'tvm48']; "
                            "sys.path.insert(0, %r); import bench; bench._tvm48_cpu_iteration()" % (core, REPO)],
                           capture_output=True, text=True, env=env, timeout=900)
        if r.returncode == 0:
            d = json.loads(r.stdout.strip().splitlines()[-1])
            cpu = {"value": d["per_iter"] * n_it, "unit": "s", "cores": 1, "kind": "port",
                   "s_per_iteration": d["per_iter"],
                   "cpu_model": _cpu_model(),
                   "sample": f"one full TVM iteration of the single-precision NumPy restatement of the "
                             f"reference's --device -1 path (oracle/sigpy32.py; {d['bins']} bins x {d['K']} "
                             f"channels, {sum(d['samples'])} samples) = {d['per_iter']:.1f} s on 1 pinned "
                             f"thread, x{n_it} iterations of the GPU solve"}
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


def bench_tvmw(args, verbose=True):
    """TVMW (db1 echo wavelet + db6 zero-mode spatial wavelet at the automatic level + motion TV) at the
    headline acquisition, 80x192x192, 32 coils, 6 echoes, 6 bins, 9000x512 cone samples per echo:
    plain PDHG from zeros to tol 1e-3 (max 300 iterations, sigma = tau = 1/6, lambda1 = lambda2 =
    lambda3 = 1e-5, L2 coupling), timed --steps times; per-kernel HIP events in one more solve.  The
    reference's TVMW is GPU-only (it moves arrays through CuPy and torch.cuda), so there is no CPU
    path to time beside it."""
    import torch

    from toporecon_hip import _lib, synth
    from toporecon_hip.algorithms import TvmwReconstructor
    dev = torch.device("cuda", 0)
    cfg = synth.config_full_tvme()
    acq = synth.make_acquisition(cfg)
    s = TvmwReconstructor(None, acq.coord, acq.dcf, acq.mps, acq.resp, None, cfg.bins, True, lambda1=1e-5,
                          lambda2=1e-5, lambda3=1e-5, sigma=1 / 6, tau=1 / 6, max_iter=args.max_iter or 300,
                          tol=args.tol or 1e-3, margin=0, device=dev, bin_edges=acq.bin_edges, show_pbar=False,
                          echoes=cfg.echoes)
    s.simulate_kspace(acq.truth, cfg.snr_db, seed=cfg.seed + 104729)
    tuned = s.autotune(bin_streams=False)
    log(f"tvmw: level {s.level_w2}, {s.ncoef} db6 coefficients per image, autotune {tuned.get('interp_xcd_chunk')}"
        f"/{tuned.get('interp_kgroups')}/{tuned.get('spread_xcd_chunk')}", verbose)
    for _ in range(max(args.warmup, 1)):
        s.max_iter, full = 3, s.max_iter
        s.run()
        s.max_iter = full
    torch.cuda.synchronize(dev)
    times = []
    for _ in range(args.steps):
        t0 = time.perf_counter()
        s.run()
        torch.cuda.synchronize(dev)
        times.append(time.perf_counter() - t0)
    t = sum(times) / len(times)
    n_it = s.iterations
    log(f"tvmw: {n_it} iterations, {t:.2f} s to tol ({t / n_it * 1e3:.1f} ms per iteration)", verbose)
    _lib.profile_enable(True)
    s.max_iter, full = min(20, full), s.max_iter
    s.run()
    s.max_iter = full
    prof = _lib.profile_read()
    _lib.profile_enable(False)
    kernels = {}
    for name, (ms, n, by) in prof.items():
        kernels[name] = {"launches": n, "ms_avg": ms / n, "ms_total": ms, "bytes_per_launch": by / n,
                         "gbps": by / (ms * 1e-3) / 1e9 if ms > 0 else None}
    ew = {k: v for k, v in kernels.items() if k.startswith(("w2_", "tvmw_"))}
    dom = max(ew, key=lambda k: ew[k]["ms_total"]) if ew else None
    roofline = None
    if dom:
        d = ew[dom]
        roofline = {"bound": "hbm", "kernel": dom, "achieved": d["gbps"], "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                    "frac": (d["gbps"] or 0.0) / HBM_PEAK_GBPS, "traffic": None,
                    "algorithmic_bytes_per_launch": d["bytes_per_launch"], "avg_launch_ms": d["ms_avg"],
                    "launches": d["launches"],
                    "note": "the dominant wavelet / element-wise kernel (the NUFFT kernels are those of the "
                            "headline line); bytes = each array read and written once per pass"}
    line = {
        "metric": "TVMW time-to-tol 1e-3 at 80x192x192, 32 coils, 6 echoes, 6 bins on 1 GPU",
        "value": t, "unit": "s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": t * 1e3, "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic: seeded cone trajectory, ellipsoid phantom, Gaussian-blob coil maps, "
                                "k-space simulated by the HIP forward NUFFT + 30 dB complex noise",
        "config": {"workload": (f"plain TVMW PDHG to tol {s.tol} (max {full} it), 80x192x192, {cfg.coils} coils, "
                                f"{cfg.echoes} echoes, {cfg.bins} bins, 9000x512 cone samples/echo, db6 level "
                                f"{s.level_w2}, lambda1 = lambda2 = lambda3 = 1e-5, l2 coupling"),
                   "iterations": n_it, "s_per_iteration": t / max(n_it, 1),
                   "final_tol": s.history[-1] if s.history else None,
                   "launch_params": {k: tuned.get(k) for k in ("interp_xcd_chunk", "interp_kgroups",
                                                                "spread_xcd_chunk")},
                   "run_wall_s": time.perf_counter() - T_PROCESS0},
        "roofline": roofline,
        "cpu_baseline": None,
        "kernels": kernels,
    }
    print(json.dumps(line), flush=True)
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


# ---------------------------------------------------------------------------------------------
# the headline: TVME at 80x192x192, learned warm start vs plain
# ---------------------------------------------------------------------------------------------

def _set_params(solver, params, verbose):
    tuned = json.loads(params)
    for k in ("interp_xcd_chunk", "interp_kgroups", "spread_xcd_chunk", "lines_xcd_lines"):
        if k in tuned:
            solver.plan.set_param(k, int(tuned[k]))
    log(f"launch parameters set: {tuned}", verbose)
    return tuned


def _kernel_table(prof, cfg, samples, grid_shape):
    """Per logical kernel: launches, average HIP-event duration, GB/s on the must-move bytes the
    library credits, and GB/s on the per-unit byte model (launches spread evenly over the bins)."""
    E, C, B = cfg.echoes, cfg.coils, cfg.bins
    N = math.prod(cfg.img_shape)
    G = math.prod(grid_shape)
    M_avg = sum(samples) / max(len(samples), 1)
    out = {}
    for name, (ms, n, by) in prof.items():
        model = model_bytes_per_launch(name, M_avg, E, C, N, G, B)
        out[name] = {"launches": n, "ms_avg": ms / n, "ms_total": ms,
                     "must_move_bytes_per_launch": by / n,
                     "gbps_must_move": by / (ms * 1e-3) / 1e9 if ms > 0 else None,
                     "model_bytes_per_launch": model,
                     "gbps_model": model * n / (ms * 1e-3) / 1e9 if (model and ms > 0) else None}
    return out


def bench_full(args, verbose=True):
    import torch

    from toporecon_hip import _lib, objective, synth
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
    plan_w = work_plan(args)
    log(f"work plan: {plan_w}", verbose)
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
        tuned = _set_params(solver, args.params, verbose)
    else:
        # bins stay on one stream: the per-kernel HIP events then time one kernel each
        if args.geometry not in ("default", "tune"):
            tile, seg = (int(v) for v in args.geometry.lower().split("x"))
            solver.set_geometry(tile, seg)
        tuned = solver.autotune(bin_streams=False, geometry=args.geometry == "tune")
        log(f"autotune in {time.perf_counter() - t0:.1f} s: "
            f"{ {k: tuned.get(k) for k in ('interp_xcd_chunk', 'interp_kgroups', 'spread_xcd_chunk')} } "
            f"geometry {solver.plan.get_param('sort_tile')}x{solver.plan.get_param('spread_segment')} "
            f"{tuned.get('timings_ms', {}).get('geometry', '')}", verbose)

    from toporecon_hip.learned import DEFAULT_CG_STEPS, WarmStartNet, default_weights_path, warm_start
    cg_steps = DEFAULT_CG_STEPS if args.cg_steps is None else args.cg_steps
    net = None
    if args.legs == "plain":
        log("plain leg only", verbose)
    elif os.path.exists(default_weights_path()):
        net = WarmStartNet.load(default_weights_path(), dev)
    else:
        log("no trained warm-start weights: the learned leg is skipped (plain PDHG only)", verbose)

    def learned_solve(progress=0):
        u0, q0 = warm_start(solver, net, cg_steps=cg_steps, with_dual=True)
        u = solver.run(init=u0, dual_init=q0, progress_every=progress)
        del u0, q0
        return u

    # warmup: learned solves capped at WARMUP_ITERS PDHG iterations (every kernel and code path runs)
    full_iter = solver.max_iter
    solver.max_iter = min(full_iter, WARMUP_ITERS)
    for w in range(args.warmup):
        if net is not None:
            learned_solve()
        else:
            solver.run()
        torch.cuda.synchronize(dev)
        log(f"warmup step {w + 1}/{args.warmup} done", verbose)
    solver.max_iter = full_iter

    # plain PDHG from zeros: the denominator, timed once (--legs plain: the timed steps)
    n_plain_solves = plan_w["plain_solves"]
    timed_events = not args.no_timed_events
    if args.legs == "plain" and timed_events:
        _lib.profile_enable(True)
    torch.cuda.synchronize(dev)
    t0 = time.perf_counter()
    done_plain = 0
    for k in range(n_plain_solves):
        u_plain = solver.run(progress_every=25)
        done_plain += 1
        if args.legs == "plain" and timed_events:
            _lib.profile_read()  # drain this solve's events
        el = time.perf_counter() - t0
        if time.perf_counter() - T_PROCESS0 + el / done_plain > args.budget_s:
            break
    torch.cuda.synchronize(dev)
    t_plain = (time.perf_counter() - t0) / done_plain
    n_plain = solver.iterations
    u_plain = u_plain.clone()
    phi_plain = objective.objective(solver, u_plain)
    del u_plain
    log(f"plain: {done_plain} solve(s), {n_plain} iterations, {t_plain:.2f} s per solve, "
        f"objective {phi_plain:.6e}", verbose)

    # learned leg: the timed steps
    warm = None
    steps_done = 0
    t_learned = None
    if net is not None:
        if timed_events:
            _lib.profile_enable(True)
        torch.cuda.synchronize(dev)
        t0 = time.perf_counter()
        t_last = 0.0
        for k in range(args.steps):
            if k > 0 and time.perf_counter() - T_PROCESS0 + 1.2 * t_last > args.budget_s:
                log(f"wall-clock guard: stopping after {k} of {args.steps} timed steps", True)
                break
            ta = time.perf_counter()
            u_warm = learned_solve(progress=25 if k == 0 else 0)
            torch.cuda.synchronize(dev)
            t_last = time.perf_counter() - ta
            steps_done += 1
            if timed_events:
                _lib.profile_read()  # drain this step's events (keeps the event pool small)
            log(f"timed step {k + 1}/{args.steps}: {t_last:.3f} s, {solver.iterations} iterations", verbose)
        torch.cuda.synchronize(dev)
        t_learned = (time.perf_counter() - t0) / steps_done
        n_learned = solver.iterations
        phi_warm = objective.objective(solver, u_warm)
        del u_warm
        rel = (phi_warm - phi_plain) / phi_plain
        warm = {"start": f"cnn+{cg_steps}cg", "cg_steps": cg_steps, "time_to_tol_s": t_learned,
                "iterations": n_learned, "objective": phi_warm, "objective_rel_diff": rel,
                "within_0.1pct": abs(rel) <= 1e-3, "speedup_vs_plain": t_plain / t_learned,
                "weights": os.path.basename(default_weights_path()), "trained_on": net.meta}
        log(f"learned: {steps_done} step(s), {n_learned} iterations, {t_learned:.3f} s per solve, objective "
            f"{phi_warm:.6e} ({rel:+.2e} vs plain), speedup {t_plain / t_learned:.2f}x", verbose)
    prof = _lib.profile_read() if timed_events else {}
    _lib.profile_enable(False)

    variants = []
    for v in filter(None, (x.strip() for x in args.variants.split(","))):
        if net is None:
            break
        kind, k_cg = v.split("+")
        torch.cuda.synchronize(dev)
        ta = time.perf_counter()
        u0, q0 = warm_start(solver, net if kind == "cnn" else None, cg_steps=int(k_cg), with_dual=True)
        u_v = solver.run(init=u0, dual_init=q0)
        torch.cuda.synchronize(dev)
        tv = time.perf_counter() - ta
        phi = objective.objective(solver, u_v)
        del u0, q0, u_v
        variants.append({"start": f"{kind}+{k_cg}cg", "time_to_tol_s": tv, "iterations": solver.iterations,
                         "objective_rel_diff": (phi - phi_plain) / phi_plain, "speedup_vs_plain": t_plain / tv})
        log(f"variant {v}: {solver.iterations} iterations, {tv:.2f} s, objective "
            f"{variants[-1]['objective_rel_diff']:+.2e} vs plain", verbose)

    knorm = None
    if plan_w["power_iterations"]:
        t0 = time.perf_counter()
        knorm = solver.operator_norm(iters=plan_w["power_iterations"])
        knorm = {k: knorm[k] for k in ("K_norm", "sigma_tau_K_norm_sq", "iterations")}
        log(f"power iteration: ||K|| = {knorm['K_norm']:.4f}, sigma tau ||K||^2 = "
            f"{knorm['sigma_tau_K_norm_sq']:.4f} ({time.perf_counter() - t0:.1f} s)", verbose)

    use_warm = warm is not None and warm["within_0.1pct"]
    if args.legs == "plain":
        t_step, steps_done = t_plain, done_plain
    else:
        t_step = warm["time_to_tol_s"] if use_warm else t_plain

    kernels = _kernel_table(prof, cfg, samples, solver.plan.grid_shape)
    roofline = None
    if kernels:
        dom = max(kernels, key=lambda k: kernels[k]["ms_total"])
        dk = kernels[dom]
        traffic, traffic_src = None, None
        if os.path.exists(TRAFFIC_FILE):
            with open(TRAFFIC_FILE) as f:
                tj = json.load(f)
            traffic = tj.get("kernels", {}).get(dom, {}).get("bytes")
            traffic_src = (f"replayed from profiles/traffic.json (rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE of "
                           f"`{tj.get('command') or 'bench.py'}`, scaled by a calibration copy)")
        achieved = dk["gbps_model"] if dk["gbps_model"] is not None else dk["gbps_must_move"]
        roofline = {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                    "frac": achieved / HBM_PEAK_GBPS, "traffic": traffic, "traffic_source": traffic_src,
                    "bytes_model": ("per-unit algorithmic bytes x units per launch (interp 8KG + 12M + 24KM, "
                                    "spread 8KM + 16M + 8KG, pad 8EN + 8CN + 8KG, crop 8KN + 8CN + 8EN, a line "
                                    "pass 16KG; K = channels, G = oversampled grid points, M = samples of the bin)"),
                    "algorithmic_bytes_per_launch": dk["model_bytes_per_launch"],
                    "avg_launch_ms": dk["ms_avg"], "launches": dk["launches"],
                    "timing": (f"HIP events on the launch stream around every launch in the {steps_done} timed "
                               f"step(s)"),
                    "must_move": {"bytes_per_launch": dk["must_move_bytes_per_launch"],
                                  "achieved": dk["gbps_must_move"],
                                  "frac": (dk["gbps_must_move"] or 0.0) / HBM_PEAK_GBPS,
                                  "note": "bytes the kernel must move after k-space support pruning "
                                          "(touched grid points only)"}}
        gathers = {}
        for k in ("interp", "spread"):
            if k in kernels:
                gathers[k] = {"avg_launch_ms": kernels[k]["ms_avg"], "achieved": kernels[k]["gbps_model"],
                              "frac": (kernels[k]["gbps_model"] or 0.0) / HBM_PEAK_GBPS,
                              "achieved_must_move": kernels[k]["gbps_must_move"]}
        roofline["gathers"] = gathers
    fwd_b, adj_b = nufft_model_bytes(samples, cfg.echoes, cfg.coils, cfg.img_shape, solver.plan.grid_shape)
    nufft = None
    if kernels:
        n_fwd = kernels.get("interp", {}).get("launches", 0) / max(cfg.bins, 1)
        n_adj = kernels.get("spread", {}).get("launches", 0) / max(cfg.bins, 1)
        fwd_ms = sum(kernels[k]["ms_total"] for k in ("fwd_x_pad", "fft_y_fwd", "fft_z_fwd", "interp")
                     if k in kernels)
        adj_ms = sum(kernels[k]["ms_total"] for k in ("spread", "fft_z_inv", "fft_y_inv", "adj_x_crop")
                     if k in kernels)
        nufft = {"forward_gbps": fwd_b * n_fwd / (fwd_ms * 1e-3) / 1e9 if fwd_ms else None,
                 "adjoint_gbps": adj_b * n_adj / (adj_ms * 1e-3) / 1e9 if adj_ms else None,
                 "forward_ms_all_bins": fwd_ms / max(n_fwd, 1), "adjoint_ms_all_bins": adj_ms / max(n_adj, 1),
                 "bytes_model": "pad 8EN+8CN+8KG, fft3 16KG, interp 8KG+12M+24KM per bin (forward); "
                                "spread 8KM+16M+8KG, ifft3 16KG, crop 8KN+8CN+8EN per bin (adjoint)"}

    cpu = None
    if not args.no_cpu_baseline:
        cpu = run_cpu_baseline(acq, cfg, n_plain, verbose, deadline=args.budget_s + 100)

    shape = "x".join(str(v) for v in cfg.img_shape)
    out = {
        "metric": METRIC,
        "value": t_step,
        "unit": "s",
        "n_gpus": 1,
        "steps": steps_done,
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
                   "step": ("one learned-warm-start solve to tolerance (warm start + PDHG)" if args.legs == "all"
                            else "one plain PDHG solve to tolerance"),
                   "steps_requested": args.steps,
                   "start": "learned" if use_warm else "plain",
                   "plain": {"time_to_tol_s": t_plain, "iterations": n_plain, "objective": phi_plain,
                             "s_per_iteration": t_plain / max(n_plain, 1), "solves_timed": done_plain},
                   "learned_warm_start": warm,
                   "warm_start_variants": variants,
                   "operator_norm": knorm,
                   "iterations": warm["iterations"] if use_warm else n_plain,
                   "s_per_iteration": t_plain / max(n_plain, 1),
                   "samples_per_bin": samples, "grid": list(solver.plan.grid_shape),
                   "launch_params": {k: tuned.get(k) for k in ("interp_xcd_chunk", "interp_kgroups",
                                                                "spread_xcd_chunk")},
                   "geometry": {"sort_tile": solver.plan.get_param("sort_tile"),
                                "spread_segment": solver.plan.get_param("spread_segment"),
                                "mode": args.geometry},
                   "bin_streams": 1,
                   "final_tol": solver.history[-1] if solver.history else None,
                   "run_wall_s": time.perf_counter() - T_PROCESS0},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "nufft": nufft,
        "kernels": kernels,
    }
    print(json.dumps(out), flush=True)
    return 0


def main(argv=None):
    args = parse_args(argv)
    if args.cpu_baseline_worker:
        if args.cpu_baseline_worker.startswith("jsense:"):
            _jsense_cpu_sample(args.cpu_baseline_worker[len("jsense:"):])
        else:
            _cpu_worker(args.cpu_baseline_worker)
        return 0
    if args.traffic_from:
        out = write_traffic(args.traffic_from[0], args.traffic_from[1], args.copy_calibration_mib or 2048,
                            TRAFFIC_FILE, command=args.traffic_command)
        print(json.dumps(out["calibration"]))
        return 0
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: this build has no multi-GPU sharding, so --gpus must be 1 "
              "(how the path would shard is described in the design notes)", file=sys.stderr)
        return 2
    if args.config == "cnn":
        return bench_cnn(args)
    if args.config == "jsense":
        return bench_jsense(args)
    if args.config == "tvmw":
        return bench_tvmw(args)
    if args.config == "tvm48":
        return bench_tvm48(args)
    if args.config == "batched":
        return bench_batched(args)
    return bench_full(args)


if __name__ == "__main__":
    sys.exit(main())
