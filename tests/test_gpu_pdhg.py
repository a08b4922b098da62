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
    for ci, cs, kg, xl in [(0, 0, 0, 0), (1, 1, 1, 1), (16, 64, 0, 4), (1024, 4096, 1, 64)]:
        base.plan.set_param("interp_xcd_chunk", ci)
        base.plan.set_param("spread_xcd_chunk", cs)
        base.plan.set_param("interp_kgroups", kg)
        base.plan.set_param("lines_xcd_lines", xl)
        assert base.plan.get_param("interp_xcd_chunk") == ci
        assert base.plan.get_param("lines_xcd_lines") == xl
        out = base.run().cpu().numpy()
        assert np.array_equal(out.view(np.uint64), ref.view(np.uint64)), (ci, cs)
    for n in (3, 2, 1):  # bins on several streams, one grid workspace per stream
        base.set_bin_streams(n)
        out = base.run().cpu().numpy()
        assert np.array_equal(out.view(np.uint64), ref.view(np.uint64)), n
    tuned = base.autotune(reps=1, force=True)
    assert tuned["interp_xcd_chunk"] in (0, 16, 64, 256, 1024)
    out = base.run().cpu().numpy()
    assert np.array_equal(out.view(np.uint64), ref.view(np.uint64))
    with pytest.raises(Exception):
        base.plan.set_param("interp_xcd_chunk", 5000)
    with pytest.raises(Exception):
        base.plan.set_param("lines_xcd_lines", 65)


@pytest.mark.gpu
def test_interp_channel_group_modes_bitwise(torch_mod):
    """96 channels (16 coils x 6 echoes): the interpolation's channel-group modes (all channels
    per block, 64-channel groups, 16-byte loads over 96-channel groups, and the same with one tap
    gather per group of samples with identical taps) give the same bits."""
    acq, ksp, prob = make_problem(shape=(10, 12, 14), coils=16, echoes=6, bins=2, interleaves=400, seed=13)
    s = _hip_solver("tvme", acq, ksp, 3, 0.0, True)
    outs = []
    for kg in (0, 1, 2, 3):
        s.plan.set_param("interp_kgroups", kg)
        outs.append(s.run().cpu().numpy())
    for o in outs[1:]:
        assert np.array_equal(o.view(np.uint64), outs[0].view(np.uint64))
    o_ubar, _, _, _ = _oracle_run(prob, "tvme", 3, True)
    assert rel(outs[0], o_ubar) < 2e-5


@pytest.mark.gpu
def test_learned_warm_start_gains_over_plain(torch_mod):
    """The committed network, on a held-out problem of the size it was trained at
    (64x96x96, 16 coils, 4 echoes, 6 bins; the training seeds are 1000-1083): the
    warm-started solve reaches the same tolerance in fewer iterations and the final
    objective agrees with the plain solve within 0.1 %."""
    import os
    from toporecon_hip import objective
    from toporecon_hip.learned import WarmStartNet, cnn_warm_start, default_weights_path
    from toporecon_hip.learned.train import make_solver, training_config
    assert os.path.exists(default_weights_path())
    solver = make_solver(training_config(9003), 0)
    u_plain = solver.run().clone()
    it_plain = solver.iterations
    phi_plain = objective.objective(solver, u_plain)
    net = WarmStartNet.load(default_weights_path(), 0)
    u0 = cnn_warm_start(solver, net)
    # the warm start is much closer to the converged image than the scaled adjoint alone
    err = float(torch_mod.linalg.vector_norm(u0 - u_plain) / torch_mod.linalg.vector_norm(u_plain))
    assert err < 0.35
    u_warm = solver.run(init=u0, dual_init="residual")
    it_warm = solver.iterations
    phi_warm = objective.objective(solver, u_warm)
    assert it_warm <= 0.85 * it_plain, (it_warm, it_plain)
    assert abs(phi_warm - phi_plain) / phi_plain < 1e-3
    # CNN followed by the default conjugate-gradient data-consistency steps: about a third of the
    # iterations (0.26-0.34 of plain on seeds 9000-9002), the same objective within 0.1 %
    from toporecon_hip.learned import warm_start
    u_cg = solver.run(init=warm_start(solver, net), dual_init="residual")
    it_cg = solver.iterations
    phi_cg = objective.objective(solver, u_cg)
    print("plain", it_plain, "cnn", it_warm, "cnn+cg", it_cg, "objective", phi_plain, phi_warm, phi_cg, flush=True)
    assert it_cg <= 0.45 * it_plain, (it_cg, it_plain)
    assert abs(phi_cg - phi_plain) / phi_plain < 1e-3


@pytest.mark.gpu
def test_residual_dual_init_matches_definition(torch_mod):
    """q = A x - y from tr_data_residual equals tr_data_simulate minus y."""
    torch = torch_mod
    from toporecon_hip import _lib
    acq, ksp, prob = make_problem(seed=12)
    solver = _hip_solver("tvme", acq, ksp, 1, 0.0, True)
    x = torch.from_numpy((np.random.default_rng(1).standard_normal(solver.vol_dims())
                          + 0j).astype(np.complex64)).cuda()
    solver.init_dual_residual(x)
    for b in range(solver.B):
        ax = torch.empty_like(solver.y[b])
        _lib.call("tr_data_simulate", solver.plan.handle, b, x[b].data_ptr(), solver.maps.data_ptr(),
                  solver.E, solver.C, ax.data_ptr(), solver.stream)
        assert torch.equal(solver.q[b], ax - solver.y[b])


@pytest.mark.gpu
def test_support_pruning_changes_no_bit(torch_mod):
    """The line passes and the spreading skip grid points outside the samples' k-space
    support; the cone trajectory leaves the grid's corners empty, and skipping them
    changes no result bit of a solve (forward, adjoint and the PDHG iterates)."""
    acq, ksp, prob = make_problem(seed=13)
    solver = _hip_solver("tvme", acq, ksp, 4, 0.0, True)
    plan = solver.plan
    G = int(np.prod(plan.grid_shape))
    pts, cols = plan.get_param("support_points"), plan.get_param("live_columns")
    assert 0 < pts < G and 0 < cols <= plan.grid_shape[1] * plan.grid_shape[2]
    assert plan.get_param("prune") == 1
    pruned = solver.run().cpu().numpy()
    plan.set_param("prune", 0)
    full = solver.run().cpu().numpy()
    plan.set_param("prune", 1)
    assert np.array_equal(pruned.view(np.uint64), full.view(np.uint64))
    with pytest.raises(Exception):
        plan.set_param("support_points", 5)


def _oracle_normal(prob, x, alg):
    """K^H K x composed from the oracle's own step pieces (sigma = 1, no prox, tau = -1)."""
    g = op.adjoint_op(prob, op.forward_op(prob, x))
    dm = op.motion_diff(x)
    if alg == "tvm":
        return g - op.motion_div_tvm(dm)
    de = op.echo_diff(x)
    divp = op.fdiff_adj(op.fdiff(de, -1), -1) + op.fdiff_adj(op.fdiff(de, -2), -2) \
        + op.fdiff_adj(op.fdiff(de, -3), -3)
    return g + op.motion_div_tvme(dm) + op.echo_div(divp)


@pytest.mark.gpu
@pytest.mark.parametrize("alg", ["tvm", "tvme"])
def test_operator_norm_power_iteration(torch_mod, alg):
    """The GPU normal operator K^H K equals the oracle's, and the float64 power iteration's
    ||K||^2 agrees with a float64 power iteration on the oracle."""
    torch = torch_mod
    from toporecon_hip import _lib
    acq, ksp, prob = make_problem(seed=14)
    solver = _hip_solver(alg, acq, ksp, 1, 0.0, True)
    solver._alloc_reduction()
    rng = np.random.default_rng(3)
    dims = solver.vol_dims()
    xn = (rng.standard_normal(dims) + 1j * rng.standard_normal(dims)).astype(np.complex64)
    x = torch.from_numpy(xn).cuda()
    g = torch.empty_like(x)
    u = torch.zeros_like(x)
    scr = torch.empty_like(x)
    q = [torch.empty_like(solver.y[b]) for b in range(solver.B)]
    for b in range(solver.B):
        _lib.call("tr_data_simulate", solver.plan.handle, b, x[b].data_ptr(), solver.maps.data_ptr(), solver.E,
                  solver.C, q[b].data_ptr(), solver.stream)
        _lib.call("tr_data_adjoint", solver.plan.handle, b, q[b].data_ptr(), solver.dcf[b].data_ptr(),
                  solver.maps.data_ptr(), solver.E, solver.C, g[b].data_ptr(), 0, solver.stream)
    num = solver._tv_normal(x, u, scr, g, {})
    ref = _oracle_normal(prob, xn.astype(np.complex128), alg)
    assert rel(u.cpu().numpy(), ref) < 2e-5
    assert abs(num - float(np.sum(np.abs(ref) ** 2))) / float(np.sum(np.abs(ref) ** 2)) < 5e-5
    est = solver.operator_norm(iters=60, seed=1)
    v = rng.standard_normal(dims) + 1j * rng.standard_normal(dims)
    v /= np.linalg.norm(v)
    lam = 0.0
    for _ in range(60):
        w = _oracle_normal(prob, v, alg)
        lam = np.linalg.norm(w)
        v = w / lam
    print(alg, "K^2 gpu", est["K_norm_sq"], "oracle", lam, "sigma tau K^2", est["sigma_tau_K_norm_sq"], flush=True)
    assert abs(est["K_norm_sq"] - lam) / lam < 1e-2
    assert est["sigma_tau_K_norm_sq"] == pytest.approx(lam / 36.0, rel=1e-2)


def test_steps_beyond_the_bound_run_to_the_cap_like_the_oracle(torch_mod):
    """sigma = tau = 1 on the TVME test problem puts sigma * tau * ||K||^2 far above PDHG's bound of 1
    (the reference fixes sigma = tau = 1/6 and, at full size, lands at 1.28): the relative change stops
    falling and rises again, the solve runs to max_iter without an error and with one history entry
    per iteration (the reference's behaviour: its overflow back-off never triggers), and the iterates
    and stop-rule values follow the float64 oracle's."""
    from toporecon_hip.algorithms import TvmeReconstructor
    print("steps beyond the bound", flush=True)
    acq, ksp, prob = make_problem(seed=3, algorithm="tvme", l2=True)
    B = len(acq.bin_edges) - 1
    s = TvmeReconstructor(ksp, acq.coord, acq.dcf, acq.mps, acq.resp, np.zeros_like(ksp), B, True, lamda_m=2e-3,
                          lamda_e=2e-3, sigma=1.0, tau=1.0, max_iter=12, tol=1e-9, margin=0, device=0,
                          bin_edges=acq.bin_edges, show_pbar=False)
    kn = s.operator_norm(iters=20)
    assert kn["sigma_tau_K_norm_sq"] > 4.0, kn
    ubar = s.run().cpu().numpy()
    assert s.iterations == 12 and len(s.history) == 12
    hist = []
    o_ubar, n, _ = op.run(prob, "tvme", lam_m=2e-3, lam_e=2e-3, sigma=1.0, tau=1.0, l2_coupling=True, max_iter=12,
                          tol=1e-9, history=hist)
    assert n == 12
    assert min(hist) < hist[-1] and min(s.history) < s.history[-1]  # not settling
    assert np.all(np.isfinite(ubar))
    np.testing.assert_allclose(s.history, hist, rtol=1e-3)
    assert rel(ubar, o_ubar) < 1e-4


def test_sample_geometry_rebuild_keeps_the_solution(torch_mod):
    """A solver whose plan is rebuilt with another sample geometry (set_geometry moves k-space,
    data duals and DCF to the new sample order) follows the oracle iterates like the default
    plan; the geometry autotuner picks one of its candidates and keeps the solver on the oracle."""
    from toporecon_hip import autotune
    print("geometry rebuild", flush=True)
    acq, ksp, prob = make_problem(seed=3, algorithm="tvme", l2=True)
    o_ubar, _, st, _ = _oracle_run(prob, "tvme", 5, True)
    s = _hip_solver("tvme", acq, ksp, 5, 0.0, True)
    s.set_geometry(4, 64)
    assert (s.plan.get_param("sort_tile"), s.plan.get_param("spread_segment")) == (4, 64)
    ubar = s.run().cpu().numpy()
    assert rel(ubar, o_ubar) < 2e-5
    assert rel(s.u.cpu().numpy(), st.u) < 2e-5
    s2 = _hip_solver("tvme", acq, ksp, 5, 0.0, True)
    tuned = s2.autotune(reps=1, bin_streams=False, geometry=True)
    assert tuned["sort_tile"] in autotune.TILE_CANDIDATES
    assert tuned["spread_segment"] in autotune.SEGMENT_CANDIDATES
    assert s2.plan.get_param("sort_tile") == tuned["sort_tile"]
    assert rel(s2.run().cpu().numpy(), o_ubar) < 2e-5
