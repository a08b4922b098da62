"""TVMW precision study on a 32x64x64 sub-problem: the FP32 HIP path against the float64 oracle.

Reports (and bounds) the relative L2 error of every operator the solver applies (NUFFT
forward and density-weighted adjoint over all coils / echoes / bins, db6 zero-mode
analysis and synthesis, db1 echo wavelet, the whole normal operator K^H K), the
power-iteration estimate of ||K|| (float64 reductions on the GPU vs float64 on the CPU),
and the iterates after 1, 10 and 40 PDHG iterations of the study's problem (4 coils, 3 echoes,
3 bins; the float64 oracle runs the 40 iterations once and is compared after 1 and 10 on the
way).  TR_PRECISION_QUICK=1 runs a smaller problem (2 coils, 2 echoes, 2 bins; 1, 5 and 20
iterations).  With TR_PRECISION_REPORT set, the numbers are also written there as JSON."""

import json
import os

import numpy as np
import pytest

from oracle import pdhg as op
from oracle import wavelet as ow
from problems import make_problem, rel

pytestmark = pytest.mark.gpu

SHAPE = (32, 64, 64)
LAM = dict(lambda1=2e-3, lambda2=2e-3, lambda3=2e-3)
FULL = os.environ.get("TR_PRECISION_QUICK") != "1"
PROBLEM = (dict(coils=4, echoes=3, bins=3, interleaves=700) if FULL
           else dict(coils=2, echoes=2, bins=2, interleaves=400))
ITERS = (1, 10, 40) if FULL else (1, 5, 20)


@pytest.fixture(scope="module")
def torch_mod():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def _flat(coeffs):
    return np.concatenate([a.reshape(-1) for a in ow.tree_leaves(coeffs)])


def test_tvmw_precision_study(torch_mod):
    torch = torch_mod
    from toporecon_hip import _lib
    from toporecon_hip.algorithms import TvmwReconstructor
    print("precision study", SHAPE, flush=True)
    acq, ksp, prob = make_problem(shape=SHAPE, readouts=64, seed=21, algorithm="tvmw", l2=True, **PROBLEM)
    B = len(acq.bin_edges) - 1
    solver = TvmwReconstructor(ksp, acq.coord, acq.dcf, acq.mps, acq.resp, np.zeros_like(ksp), B, True,
                               sigma=1 / 6, tau=1 / 6, max_iter=1, tol=0.0, margin=0, device=0,
                               bin_edges=acq.bin_edges, show_pbar=False, **LAM)
    solver.pdhg_init()
    s = solver.stream
    E, C = solver.E, solver.C
    K = E * C
    dims = solver.vol_dims()
    rng = np.random.default_rng(4)
    report = {"shape": list(SHAPE), "coils": C, "echoes": E, "bins": B,
              "samples_per_bin": [int(solver.plan.samples(b)) for b in range(B)], "db6_level": solver.level_w2}
    xn = (rng.standard_normal(dims) + 1j * rng.standard_normal(dims)).astype(np.complex64)
    x = torch.from_numpy(xn).cuda()
    x64 = xn.astype(np.complex128)

    # NUFFT forward A (all bins, coils, echoes), compared in the trajectory's own order
    ref_a = op.forward_op(prob, x64)
    errs = []
    for b in range(B):
        M = solver.y[b].shape[0]
        a_sorted = torch.empty_like(solver.y[b])
        _lib.call("tr_data_simulate", solver.plan.handle, b, x[b].data_ptr(), solver.maps.data_ptr(), E, C,
                  a_sorted.data_ptr(), s)
        a_orig = torch.empty((K, M), dtype=torch.complex64, device="cuda")
        _lib.call("tr_plan_scatter_samples", solver.plan.handle, b, a_sorted.data_ptr(), K, a_orig.data_ptr(), s)
        errs.append((a_orig.cpu().numpy().reshape(E, C, M), ref_a[b]))
    report["A"] = rel(np.concatenate([g.ravel() for g, _ in errs]), np.concatenate([r.ravel() for _, r in errs]))

    # density-weighted adjoint A^H W on random k-space
    qs = [(rng.standard_normal((E, C, solver.y[b].shape[0]))
           + 1j * rng.standard_normal((E, C, solver.y[b].shape[0]))).astype(np.complex64) for b in range(B)]
    ref_g = op.adjoint_op(prob, [q.astype(np.complex128) for q in qs])
    g = torch.empty_like(x)
    held = []
    for b in range(B):
        q_orig = torch.from_numpy(qs[b].reshape(K, -1)).cuda()
        q_sorted = torch.empty_like(solver.y[b])
        held.append((q_orig, q_sorted))
        _lib.call("tr_plan_gather_samples", solver.plan.handle, b, q_orig.data_ptr(), K, q_sorted.data_ptr(), s)
        _lib.call("tr_data_adjoint", solver.plan.handle, b, q_sorted.data_ptr(), solver.dcf[b].data_ptr(),
                  solver.maps.data_ptr(), E, C, g[b].data_ptr(), 0, s)
    report["AHW"] = rel(g.cpu().numpy(), ref_g)

    # db6 zero-mode analysis / synthesis and db1 echo wavelet
    coef = torch.empty((B * E, solver.ncoef), dtype=torch.complex64, device="cuda")
    solver.w2_forward(x, coef)
    ref_w2 = np.stack([_flat(ow.wavedec3(x64[b, e])) for b in range(B) for e in range(E)])
    report["W2"] = rel(coef.cpu().numpy(), ref_w2)
    cn = (rng.standard_normal(ref_w2.shape) + 1j * rng.standard_normal(ref_w2.shape)).astype(np.complex64)
    ct = torch.from_numpy(cn).cuda()
    back = torch.empty_like(x)
    solver.w2_adjoint(ct, back)
    lv = solver.level_w2
    shapes = ow.level_shapes(SHAPE, lv)

    def unflat(f):
        out, off = [], 0
        n = int(np.prod(shapes[-1]))
        out.append(f[off:off + n].reshape(shapes[-1]))
        off += n
        for level in range(lv, 0, -1):
            n = int(np.prod(shapes[level]))
            d = {}
            for k in ow._KEYS:
                d[k] = f[off:off + n].reshape(shapes[level])
                off += n
            out.append(d)
        return out

    ref_back = np.stack([ow.waverec3(unflat(cn[i].astype(np.complex128)), SHAPE) for i in range(B * E)])
    report["W2H"] = rel(back.cpu().numpy().reshape(B * E, *SHAPE), ref_back)
    p_m = torch.zeros_like(x)
    p_w1 = torch.zeros_like(x)
    _lib.call("tr_tvmw_dual", *dims, x.data_ptr(), p_m.data_ptr(), p_w1.data_ptr(), 1.0, 1e30, 1e30, 1, s)
    report["W1"] = rel(p_w1.cpu().numpy(), op.w1_fwd(x64))
    report["Dm"] = rel(p_m.cpu().numpy(), op.motion_diff(x64))

    # the whole normal operator K^H K and the power iteration for ||K||
    def oracle_normal(v):
        gg = op.adjoint_op(prob, op.forward_op(prob, v))
        w2 = np.stack([np.stack([ow.waverec3(ow.wavedec3(v[b, e]), SHAPE) for e in range(E)]) for b in range(B)])
        return gg - op.motion_div_tvmw(op.motion_diff(v)) + op.w1_adj(op.w1_fwd(v)) + w2

    u = torch.zeros_like(x)
    scr = torch.empty_like(x)
    for b in range(B):
        a_sorted = torch.empty_like(solver.y[b])
        _lib.call("tr_data_simulate", solver.plan.handle, b, x[b].data_ptr(), solver.maps.data_ptr(), E, C,
                  a_sorted.data_ptr(), s)
        _lib.call("tr_data_adjoint", solver.plan.handle, b, a_sorted.data_ptr(), solver.dcf[b].data_ptr(),
                  solver.maps.data_ptr(), E, C, g[b].data_ptr(), 0, s)
    solver._tv_normal(x, u, scr, g, {})
    report["KHK"] = rel(u.cpu().numpy(), oracle_normal(x64))
    est = solver.operator_norm(iters=20, seed=2)
    v = rng.standard_normal(dims) + 1j * rng.standard_normal(dims)
    v /= np.linalg.norm(v)
    lam = 0.0
    for _ in range(20):
        w = oracle_normal(v)
        lam = np.linalg.norm(w)
        v = w / lam
    report["K_norm_sq_gpu"] = est["K_norm_sq"]
    report["K_norm_sq_oracle"] = float(lam)
    report["sigma_tau_K_norm_sq"] = est["sigma_tau_K_norm_sq"]

    # iterates: the FP32 solver against the FP64 oracle from the same start
    kw = dict(lam1=LAM["lambda1"], lam2=LAM["lambda2"], lam3=LAM["lambda3"], l2_coupling=True, tol=0.0)
    o_st, done = None, 0
    for n_it in ITERS:
        solver.max_iter = n_it
        ub = solver.run().cpu().numpy()
        o_ub, _, o_st = op.run(prob, "tvmw", max_iter=n_it - done, init=o_st, **kw)  # continues the oracle run
        done = n_it
        report[f"ubar_{n_it}"] = rel(ub, o_ub)

    print("TVMW precision study (relative L2, FP32 GPU vs FP64 oracle):", flush=True)
    for k, v in report.items():
        print(f"  {k:20s} {v}", flush=True)
    path = os.environ.get("TR_PRECISION_REPORT")
    if path:
        with open(path, "w") as f:
            json.dump(report, f, indent=1)
    assert report["A"] < 1e-5 and report["AHW"] < 1e-5
    assert report["W2"] < 1e-6 and report["W2H"] < 1e-6 and report["W1"] < 1e-6 and report["Dm"] < 1e-6
    assert report["KHK"] < 2e-5
    assert abs(report["K_norm_sq_gpu"] - report["K_norm_sq_oracle"]) / report["K_norm_sq_oracle"] < 2e-2
    assert report["sigma_tau_K_norm_sq"] < 1.0
    for n_it, tol in zip(ITERS, (1e-5, 3e-5, 1e-4)):
        assert report[f"ubar_{n_it}"] < tol
