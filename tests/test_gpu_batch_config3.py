"""The batched workload: twelve independent TVM instances of uneven size (32^3 ... 96x128x128,
8-32 coils, 1-6 echoes, 4-8 motion bins, 0.2-3 M samples per echo; tol 1e-3, at most 100
iterations), each with its gridding geometry and launch parameters picked by the autotuner, solved
alone with eager launches, then advanced together by the persistent graph (captured once, every
instance gated by its own device-side stopping rule), by one HIP-graph launch per iteration and by
one stream per instance.  Every instance's iteration count, stopping history and final image are
bit-identical across the four; each instance runs its bins on one stream (graph batches refuse
several bin streams per instance)."""

import time

import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_mod():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


@pytest.mark.timeout(900)
def test_twelve_uneven_instances_batched_modes_match_sequential(torch_mod):
    torch = torch_mod
    from toporecon_hip import synth
    from toporecon_hip.algorithms import TvmReconstructor
    from toporecon_hip.batch import GraphBatch
    dev = torch.device("cuda", 0)
    solvers = []
    t0 = time.perf_counter()
    for cfg in synth.batched_instance_configs():
        acq = synth.make_acquisition(cfg)
        s = TvmReconstructor(None, acq.coord, acq.dcf, acq.mps, acq.resp, None, cfg.bins, cfg.l2_coupling,
                             lamda=cfg.lambda_motion, sigma=1 / 6, tau=1 / 6, max_iter=cfg.max_iter, tol=cfg.tol,
                             margin=0, device=dev, bin_edges=acq.bin_edges, show_pbar=False,
                             echoes=len(cfg.echo_times_ms))
        s.simulate_kspace(acq.truth, cfg.snr_db, seed=cfg.seed + 104729)
        tuned = s.autotune(reps=1, bin_streams=False, geometry=True)
        solvers.append(s)
        print(f"instance seed {cfg.seed}: {cfg.img_shape} C={cfg.coils} E={len(cfg.echo_times_ms)} B={cfg.bins} "
              f"tile {tuned['sort_tile']} segment {tuned['spread_segment']} ({time.perf_counter() - t0:.0f} s)",
              flush=True)
    ref = []
    t0 = time.perf_counter()
    for s in solvers:
        u = s.run().clone()
        ref.append((u, s.iterations, list(s.history)))
    torch.cuda.synchronize(dev)
    t_seq = time.perf_counter() - t0
    its = [n for _, n, _ in ref]
    print(f"sequential {t_seq:.2f} s, iterations {its}", flush=True)
    assert len(set(its)) > 1  # instances leave the batch at different replays
    for mode in ("persistent", "graph", "streams"):
        print(f"batch mode {mode}", flush=True)
        batch = GraphBatch(solvers)
        torch.cuda.synchronize(dev)
        t0 = time.perf_counter()
        out = batch.run(mode=mode)
        torch.cuda.synchronize(dev)
        tb = time.perf_counter() - t0
        print(f"{mode}: {tb:.2f} s, {batch.captures} captures", flush=True)
        for i, (s, (u, n, hist), o) in enumerate(zip(solvers, ref, out)):
            assert o["iterations"] == n, (mode, i, o["iterations"], n)
            assert o["history"] == hist, (mode, i)
            assert torch.equal(s.ubar, u), (mode, i)
        if mode == "graph":
            assert batch.captures == len(set(its))
        if mode == "persistent":
            assert batch.captures == 1
    for s in solvers:
        s.plan.close()
