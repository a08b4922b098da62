"""HIP TVM / TVME solver parity against the float64 oracle iterations."""

import numpy as np
import pytest

from oracle import pdhg as op
from problems import make_problem, rel

pytestmark = pytest.mark.gpu

SIGMA = TAU = 1 / 6


@pytest.fixture(scope="module")
def torch_mod():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def _hip_solver(alg, acq, ksp, max_iter, tol, l2, lam=2e-3):
    from toporecon_hip.algorithms import TvmeReconstructor, TvmReconstructor
    common = dict(sigma=SIGMA, tau=TAU, max_iter=max_iter, tol=tol, margin=0, device=0,
                  bin_edges=acq.bin_edges, show_pbar=False)
    dual_q = np.zeros_like(ksp)
    B = len(acq.bin_edges) - 1
    if alg == "tvm":
        return TvmReconstructor(ksp, acq.coord, acq.dcf, acq.mps, acq.resp, dual_q, B, l2, lamda=lam,
                                **common)
    return TvmeReconstructor(ksp, acq.coord, acq.dcf, acq.mps, acq.resp, dual_q, B, l2, lamda_m=lam,
                             lamda_e=lam, **common)


def _oracle_run(prob, alg, iters, l2, lam=2e-3, tol=0.0):
    hist = []
    ubar, n, st = op.run(prob, alg, lam_m=lam, lam_e=lam, l2_coupling=l2, max_iter=iters, tol=tol,
                         history=hist)
    return ubar, n, st, hist


@pytest.mark.parametrize("alg", ["tvm", "tvme"])
@pytest.mark.parametrize("l2", [False, True])
@pytest.mark.parametrize("iters,tol", [(1, 1e-5), (5, 2e-5)])
def test_iterates_match_oracle(torch_mod, alg, l2, iters, tol):
    print("iterates", alg, l2, iters, flush=True)
    acq, ksp, prob = make_problem(seed=3, algorithm=alg, l2=l2)
    s = _hip_solver(alg, acq, ksp, iters, 0.0, l2)
    ubar = s.run().cpu().numpy()
    assert s.iterations == iters
    o_ubar, _, st, hist = _oracle_run(prob, alg, iters, l2)
    assert rel(ubar, o_ubar) < tol
    assert rel(s.u.cpu().numpy(), st.u) < tol
    assert rel(s.p_m.cpu().numpy(), st.p_m) < 10 * tol or np.linalg.norm(st.p_m) < 1e-12
    assert s.history[-1] == pytest.approx(hist[-1], rel=1e-4)


@pytest.mark.parametrize("alg", ["tvm", "tvme"])
def test_solve_to_tolerance_matches_oracle(torch_mod, alg):
    print("solve", alg, flush=True)
    acq, ksp, prob = make_problem(seed=5, algorithm=alg, l2=True)
    s = _hip_solver(alg, acq, ksp, 60, 1e-3, True)
    ubar = s.run().cpu().numpy()
    o_ubar, n, st, hist = _oracle_run(prob, alg, 60, True, tol=1e-3)
    assert abs(s.iterations - n) <= 1
    m = min(s.iterations, n)
    assert rel(ubar, o_ubar) < 1e-4 or s.iterations != n
    assert np.allclose(s.history[:m], hist[:m], rtol=1e-3)


def test_edge_shapes_single_bin_single_echo(torch_mod):
    acq, ksp, prob = make_problem(shape=(8, 10, 6), coils=1, echoes=1, bins=1, seed=7, algorithm="tvme")
    s = _hip_solver("tvme", acq, ksp, 4, 0.0, False)
    ubar = s.run().cpu().numpy()
    o_ubar, _, _, _ = _oracle_run(prob, "tvme", 4, False)
    assert rel(ubar, o_ubar) < 2e-5


def test_rerun_is_bitwise_deterministic(torch_mod):
    acq, ksp, prob = make_problem(seed=9)
    a = _hip_solver("tvme", acq, ksp, 3, 0.0, True).run().cpu().numpy()
    b = _hip_solver("tvme", acq, ksp, 3, 0.0, True).run().cpu().numpy()
    assert np.array_equal(a.view(np.uint64), b.view(np.uint64))


@pytest.mark.gpu
def test_launch_params_and_autotune_keep_results_bitwise(torch_mod):
    """The XCD block order of the gather kernels is a tuning knob only: every
    choice, and whatever the autotuner picks, gives the same bits."""
    acq, ksp, prob = make_problem(seed=11)
    base = _hip_solver("tvme", acq, ksp, 3, 0.0, True)
    ref = base.run().cpu().numpy()
    for ci, cs in [(0, 0), (1, 1), (16, 64), (1024, 4096)]:
        base.plan.set_param("interp_xcd_chunk", ci)
        base.plan.set_param("spread_xcd_chunk", cs)
        assert base.plan.get_param("interp_xcd_chunk") == ci
        out = base.run().cpu().numpy()
        assert np.array_equal(out.view(np.uint64), ref.view(np.uint64)), (ci, cs)
    tuned = base.autotune(reps=1, force=True)
    assert tuned["interp_xcd_chunk"] in (0, 16, 64, 256, 1024)
    out = base.run().cpu().numpy()
    assert np.array_equal(out.view(np.uint64), ref.view(np.uint64))
    with pytest.raises(Exception):
        base.plan.set_param("interp_xcd_chunk", 5000)
