"""Independent PDHG instances advanced by one HIP-graph launch per iteration.

Every replay is checked against the same instance run eagerly on its own, with eager work
and synchronisations of the program between the replays; instances stop at different
iteration counts, so the graph is captured again as the set shrinks."""

import numpy as np
import pytest

from problems import make_problem

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_mod():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def _solver(alg, acq, ksp, max_iter, tol):
    from toporecon_hip.algorithms import TvmeReconstructor, TvmReconstructor, TvmwReconstructor
    common = dict(sigma=1 / 6, tau=1 / 6, max_iter=max_iter, tol=tol, margin=0, device=0,
                  bin_edges=acq.bin_edges, show_pbar=False)
    B = len(acq.bin_edges) - 1
    dq = np.zeros_like(ksp)
    if alg == "tvm":
        return TvmReconstructor(ksp, acq.coord, acq.dcf, acq.mps, acq.resp, dq, B, True, lamda=2e-3, **common)
    if alg == "tvme":
        return TvmeReconstructor(ksp, acq.coord, acq.dcf, acq.mps, acq.resp, dq, B, True, lamda_m=2e-3,
                                 lamda_e=2e-3, **common)
    return TvmwReconstructor(ksp, acq.coord, acq.dcf, acq.mps, acq.resp, dq, B, True, lambda1=2e-3,
                             lambda2=2e-3, lambda3=2e-3, **common)


@pytest.mark.parametrize("mode", ["graph", "streams"])
def test_graph_batch_matches_eager_on_every_replay(torch_mod, mode):
    torch = torch_mod
    from toporecon_hip.batch import GraphBatch
    specs = [("tvm", dict(shape=(8, 10, 12), seed=31), 6), ("tvme", dict(shape=(10, 12, 12), coils=2, echoes=3,
             bins=2, seed=32), 4), ("tvmw", dict(shape=(12, 12, 16), seed=33, algorithm="tvmw"), 5),
             ("tvm", dict(shape=(16, 16, 16), coils=4, seed=34), 3)]
    solvers = []
    for alg, kw, n in specs:
        acq, ksp, _ = make_problem(**kw)
        solvers.append(_solver(alg, acq, ksp, n, 0.0))
    eager = []
    for s in solvers:
        s.pdhg_init(None)
        s.reset_dual()
        seq = []
        for k in range(s.max_iter):
            s.pdhg(k)
            seq.append(s.ubar.clone())
        eager.append(seq)
    checked = []

    def on_replay(k, active):
        for i in active:
            assert torch.equal(solvers[i].ubar, eager[i][k - 1]), (i, k)
            checked.append((i, k))
        junk = torch.randn(4096, device="cuda")  # the program's own eager work between replays
        float(junk.sum())
        torch.cuda.synchronize()

    batch = GraphBatch(solvers)
    out = batch.run(on_replay=on_replay, mode=mode)
    assert [o["iterations"] for o in out] == [n for _, _, n in specs]
    assert len(checked) == sum(n for _, _, n in specs)
    assert batch.captures == (len({n for _, _, n in specs}) if mode == "graph" else 0)
    for s, seq in zip(solvers, eager):
        assert torch.equal(s.ubar, seq[-1])


def test_graph_batch_stops_each_instance_on_its_rule(torch_mod):
    torch = torch_mod
    from toporecon_hip.batch import GraphBatch
    solvers, ref = [], []
    for seed, shape in ((41, (8, 12, 12)), (42, (12, 16, 16))):
        acq, ksp, _ = make_problem(shape=shape, seed=seed)
        s = _solver("tvm", acq, ksp, 100, 1e-3)
        ref.append((s.run().clone(), s.iterations))
        solvers.append(s)
    out = GraphBatch(solvers).run()
    for s, (u, n), o in zip(solvers, ref, out):
        assert o["iterations"] == n
        assert torch.equal(s.ubar, u)


def test_graph_mode_refuses_per_bin_streams(torch_mod):
    """Graph batches run each instance's bins on one stream: an instance set to several bin
    streams is refused before anything is captured."""
    from toporecon_hip.batch import GraphBatch
    acq, ksp, _ = make_problem(shape=(8, 10, 12), bins=3, seed=35)
    s = _solver("tvm", acq, ksp, 2, 0.0)
    s.set_bin_streams(2)
    with pytest.raises(ValueError):
        GraphBatch([s]).run(mode="graph")


@pytest.mark.parametrize("chunk", [4, None])
def test_persistent_batch_matches_solves_alone(torch_mod, chunk):
    """One graph captured once: every instance's iteration gated by its own device flag and
    followed by its stopping rule on the device.  Instances with different caps and
    tolerances stop mid-replay (chunk 4) or inside a single launch (chunk None); each final
    image, iteration count and stopping history equals the instance solved alone."""
    torch = torch_mod
    from toporecon_hip.batch import GraphBatch
    specs = [("tvm", dict(shape=(8, 10, 12), seed=31), 6, 0.0),
             ("tvme", dict(shape=(10, 12, 12), coils=2, echoes=3, bins=2, seed=32), 9, 0.0),
             ("tvmw", dict(shape=(12, 12, 16), seed=33, algorithm="tvmw"), 5, 0.0),
             ("tvm", dict(shape=(8, 12, 12), seed=41), 100, 1e-2),
             ("tvm", dict(shape=(12, 16, 16), seed=42), 100, 3e-3)]
    solvers, ref = [], []
    for alg, kw, n, tol in specs:
        acq, ksp, _ = make_problem(**kw)
        s = _solver(alg, acq, ksp, n, tol)
        u = s.run().clone()
        ref.append((u, s.iterations, list(s.history)))
        solvers.append(s)
    assert ref[3][1] < 100 and ref[4][1] < 100, [r[1] for r in ref]  # the tolerance stops them
    batch = GraphBatch(solvers)
    out = batch.run(mode="persistent", chunk=chunk)
    assert batch.captures == 1
    assert batch.replays == (1 if chunk is None else -(-max(r[1] for r in ref) // chunk))
    for s, (u, n, hist), o in zip(solvers, ref, out):
        assert o["iterations"] == n
        assert o["history"] == hist  # float64 values of the device rule equal the host's
        assert torch.equal(s.ubar, u)
        assert o["seconds"] > 0


def test_persistent_batch_refuses_per_bin_streams(torch_mod):
    from toporecon_hip.batch import GraphBatch
    acq, ksp, _ = make_problem(shape=(8, 10, 12), bins=3, seed=35)
    s = _solver("tvm", acq, ksp, 2, 0.0)
    s.set_bin_streams(2)
    with pytest.raises(ValueError):
        GraphBatch([s]).run(mode="persistent")
