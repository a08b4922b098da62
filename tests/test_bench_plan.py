"""The benchmark's planned work and its CPU-baseline leg (CPU only, no GPU calls).

The driver runs ``bench.py --gpus 1 --steps 20 --warmup 5`` under a 600 s limit.  The plan
must time the plain leg once, the learned leg ``--steps`` times, no variant solves unless asked,
and its wall-clock estimate from measured full-size figures must stay well inside the limit; the
guard budget leaves room for the work after the timed steps.
"""

import math

import numpy as np

import bench


def test_driver_command_plan_fits_the_limit():
    args = bench.parse_args(["--gpus", "1", "--steps", "20", "--warmup", "5"])
    plan = bench.work_plan(args)
    assert plan["plain_solves"] == 1
    assert plan["learned_solves"] == 20
    assert plan["variant_solves"] == 0
    assert plan["warmup_solves"] == 5 and plan["warmup_iterations"] <= 5
    # full-size figures measured on the MI355X: 0.111 s per PDHG iteration, 0.73 s of warm-start
    # work, 226 plain / 46 learned iterations; setup includes a cold first `import torch` (120 s)
    est = bench.estimate_seconds(plan, t_setup=180.0, t_iter=0.111, t_start=0.73, n_plain=226, n_learned=46,
                                 t_cpu_round=15.0)
    assert est < 420.0, est
    # the guard stops timed steps at budget_s; what follows (objective, power iteration, CPU
    # baseline, one last step's overrun) must still fit under 600 s
    tail = plan["power_iterations"] * 2 * 0.111 + plan["cpu_single_core_budget_s"] + 30.0 + 6.0
    assert plan["budget_s"] + tail < 560.0


def test_default_run_is_short():
    plan = bench.work_plan(bench.parse_args([]))
    assert plan["learned_solves"] <= 3 and plan["warmup_solves"] <= 1
    est = bench.estimate_seconds(plan, t_setup=180.0, t_iter=0.111, t_start=0.73, n_plain=226, n_learned=46,
                                 t_cpu_round=15.0)
    assert est < 300.0


def test_variants_are_opt_in():
    plan = bench.work_plan(bench.parse_args(["--variants", "cnn+0,zero+4"]))
    assert plan["variant_solves"] == 2
    plain = bench.work_plan(bench.parse_args(["--legs", "plain", "--steps", "4"]))
    assert plain["plain_solves"] == 4 and plain["learned_solves"] == 0 and plain["power_iterations"] == 0


def test_multi_gpu_request_exits_with_a_message(capsys):
    assert bench.main(["--gpus", "2"]) == 2
    assert "one GPU" in capsys.readouterr().err


def test_model_bytes_match_the_per_unit_figures():
    E, C, B = 6, 32, 6
    K = E * C
    N = 80 * 192 * 192
    G = 100 * 240 * 240
    M = 721_000
    assert bench.model_bytes_per_launch("spread", M, E, C, N, G, B) == 8 * K * M + 16 * M + 8 * K * G
    assert bench.model_bytes_per_launch("interp", M, E, C, N, G, B) == 8 * K * G + 12 * M + 24 * K * M
    fwd, adj = bench.nufft_model_bytes([M] * B, E, C, (80, 192, 192), (100, 240, 240))
    # about 436 GB per PDHG iteration at this size (forward + adjoint over all bins)
    assert math.isclose((fwd + adj) / 1e9, 436.0, rel_tol=0.01)


def test_cpu_baseline_leg_runs_on_a_small_problem():
    from toporecon_hip import synth
    cfg = synth.config_scaled(synth.config_small_tvm(), (8, 12, 12), coils=3, echoes=2, bins=2,
                              interleaves=60, readouts=32, seed=4)
    acq = synth.make_acquisition(cfg)
    out = bench.run_cpu_baseline(acq, cfg, n_plain=10, verbose=False, deadline=1e9)
    assert out is not None and out["cores"] == 1 and out["kind"] == "port"
    assert out["value"] > 0 and np.isclose(out["value"], 10 * out["s_per_iteration"])
    assert "bins [0, 1]" in out["sample"]
    assert out["all_cores"]["cores"] >= 1
