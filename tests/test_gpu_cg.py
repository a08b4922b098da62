"""Conjugate-gradient refinement of a warm start: the vector kernels against float64
definitions and the whole refinement against float64 CG over the oracle's operators."""

import numpy as np
import pytest

from oracle import pdhg as op
from problems import make_problem, rel

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_mod():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def _cvec(rng, n):
    return (rng.standard_normal(n) + 1j * rng.standard_normal(n)).astype(np.complex64)


@pytest.mark.parametrize("n,nblocks", [(1, 1), (1000, 7), (300001, 1024)])
def test_cg_vector_kernels(torch_mod, n, nblocks):
    torch = torch_mod
    from toporecon_hip import _lib
    print("cg kernels", n, nblocks, flush=True)
    rng = np.random.default_rng(n)
    xn, rn, pn, an = (_cvec(rng, n) for _ in range(4))
    x, r, p, ap = (torch.from_numpy(v).cuda() for v in (xn, rn, pn, an))
    part = torch.empty(nblocks, dtype=torch.float64, device="cuda")
    s = torch.cuda.current_stream().cuda_stream
    alpha = 0.37
    _lib.call("tr_cg_update", x.data_ptr(), r.data_ptr(), p.data_ptr(), ap.data_ptr(), n, alpha,
              part.data_ptr(), nblocks, s)
    x_ref = xn.astype(np.complex128) + np.float32(alpha) * pn
    r_ref = rn.astype(np.complex128) - np.float32(alpha) * an
    assert rel(x.cpu().numpy(), x_ref) < 1e-6 and rel(r.cpu().numpy(), r_ref) < 1e-6
    rr = np.sum(np.abs(r.cpu().numpy().astype(np.complex128)) ** 2)  # of the stored float32 residual
    assert part.cpu().numpy().sum() == pytest.approx(rr, rel=1e-12)
    _lib.call("tr_cdot_re", p.data_ptr(), ap.data_ptr(), n, part.data_ptr(), nblocks, s)
    dot = np.real(np.vdot(pn.astype(np.complex128), an.astype(np.complex128)))
    assert part.cpu().numpy().sum() == pytest.approx(dot, rel=1e-10, abs=1e-9 * n)
    beta = -1.25
    _lib.call("tr_cg_direction", p.data_ptr(), r.data_ptr(), n, beta, s)
    p_ref = r.cpu().numpy().astype(np.complex128) + beta * pn
    assert rel(p.cpu().numpy(), p_ref) < 1e-6


def _oracle_cg(prob, x0, steps):
    """Textbook CG (Hestenes-Stiefel) on A^H W A x = A^H W y in float64 with the oracle's operators."""
    normal = lambda v: op.adjoint_op(prob, op.forward_op(prob, v))
    x = x0.astype(np.complex128)
    r = op.adjoint_op(prob, prob.ksp) - normal(x)
    p = r.copy()
    rr = np.vdot(r, r).real
    data = [0.5 * sum(np.sum(prob.dcf[b][None, None] * np.abs(a - y) ** 2)
                      for b, (a, y) in enumerate(zip(op.forward_op(prob, x), prob.ksp)))]
    for _ in range(steps):
        ap = normal(p)
        alpha = rr / np.vdot(p, ap).real
        x = x + alpha * p
        r = r - alpha * ap
        rr_new = np.vdot(r, r).real
        p = r + (rr_new / rr) * p
        rr = rr_new
        data.append(0.5 * sum(np.sum(prob.dcf[b][None, None] * np.abs(a - y) ** 2)
                              for b, (a, y) in enumerate(zip(op.forward_op(prob, x), prob.ksp))))
    return x, data


@pytest.mark.parametrize("start,steps", [("zeros", 4), ("random", 4), ("random", 1)])
def test_cg_refine_matches_float64_cg(torch_mod, start, steps):
    """cg_refine against Hestenes-Stiefel CG in float64 over the oracle's operators; its last step
    takes the step length from <A p, W A p> in k-space and runs no adjoint, so steps = 1 is that
    path alone."""
    torch = torch_mod
    from toporecon_hip.algorithms import TvmeReconstructor
    from toporecon_hip.learned import cg_refine
    print("cg refine", start, steps, flush=True)
    acq, ksp, prob = make_problem(shape=(16, 24, 24), coils=3, echoes=2, bins=2, interleaves=120,
                                  readouts=40, seed=8, algorithm="tvme")
    B = len(acq.bin_edges) - 1
    s = TvmeReconstructor(ksp, acq.coord, acq.dcf, acq.mps, acq.resp, np.zeros_like(ksp), B, True,
                          lamda_m=2e-3, lamda_e=2e-3, sigma=1 / 6, tau=1 / 6, max_iter=1, tol=0.0, margin=0,
                          device=0, bin_edges=acq.bin_edges, show_pbar=False)
    s.pdhg_init()
    dims = s.vol_dims()
    rng = np.random.default_rng(5)
    x0n = np.zeros(dims, np.complex64) if start == "zeros" else 0.1 * _cvec(rng, int(np.prod(dims))).reshape(dims)
    x0 = torch.from_numpy(x0n).cuda()
    x_gpu = cg_refine(s, x0, steps).cpu().numpy()
    assert np.array_equal(x0.cpu().numpy(), x0n)  # the start is not modified
    x_ref, data = _oracle_cg(prob, x0n, steps)
    # CG minimises the data term over a growing Krylov space: strictly decreasing
    assert all(b < a for a, b in zip(data, data[1:])), data
    assert rel(x_gpu, x_ref) < 1e-4
    assert cg_refine(s, x0, 0) is x0


def test_warm_start_reuses_rhs_and_zero_start(torch_mod):
    """warm_start's shortcuts (the CNN's input reused as the CG rhs; A u0 formed as c * A v from the
    scale fit; no normal-operator product for a zero start) give the plain composition's result up
    to float32 rounding."""
    torch = torch_mod
    from toporecon_hip.algorithms import TvmeReconstructor
    from toporecon_hip.learned import WarmStartNet, cg_refine, cnn_warm_start, default_weights_path, warm_start
    acq, ksp, prob = make_problem(shape=(12, 16, 16), coils=4, echoes=2, bins=2, interleaves=200,
                                  readouts=40, seed=9, algorithm="tvme")
    B = len(acq.bin_edges) - 1
    s = TvmeReconstructor(ksp, acq.coord, acq.dcf, acq.mps, acq.resp, np.zeros_like(ksp), B, True,
                          lamda_m=2e-3, lamda_e=2e-3, sigma=1 / 6, tau=1 / 6, max_iter=1, tol=0.0, margin=0,
                          device=0, bin_edges=acq.bin_edges, show_pbar=False)
    s.pdhg_init()
    net = WarmStartNet.load(default_weights_path(), 0)
    a = warm_start(s, net, cg_steps=3).cpu().numpy()
    b = cg_refine(s, cnn_warm_start(s, net), 3).cpu().numpy()
    assert rel(a, b) < 1e-5
    z = torch.zeros(s.vol_dims(), dtype=torch.complex64, device="cuda")
    c = warm_start(s, None, cg_steps=3).cpu().numpy()
    d = cg_refine(s, z, 3).cpu().numpy()
    assert rel(c, d) < 1e-6


@pytest.mark.parametrize("with_net", [True, False])
def test_warm_start_dual_from_tracked_kspace(torch_mod, with_net):
    """warm_start(..., with_dual=True) returns the same image as without, and a data dual A u0 - y
    built from A v (scale fit) and the CG steps' A p that matches the direct residual
    (tr_data_residual, one forward NUFFT per bin) up to float32 rounding; run() accepts it."""
    torch = torch_mod
    from toporecon_hip.algorithms import TvmeReconstructor
    from toporecon_hip.learned import WarmStartNet, default_weights_path, warm_start
    acq, ksp, prob = make_problem(shape=(12, 16, 16), coils=4, echoes=2, bins=2, interleaves=200,
                                  readouts=40, seed=9, algorithm="tvme")
    B = len(acq.bin_edges) - 1
    s = TvmeReconstructor(ksp, acq.coord, acq.dcf, acq.mps, acq.resp, np.zeros_like(ksp), B, True,
                          lamda_m=2e-3, lamda_e=2e-3, sigma=1 / 6, tau=1 / 6, max_iter=3, tol=0.0, margin=0,
                          device=0, bin_edges=acq.bin_edges, show_pbar=False)
    s.pdhg_init()
    net = WarmStartNet.load(default_weights_path(), 0) if with_net else None
    u_plain = warm_start(s, net, cg_steps=3)
    u, q0 = warm_start(s, net, cg_steps=3, with_dual=True)
    assert torch.equal(torch.view_as_real(u), torch.view_as_real(u_plain))
    s.init_dual_residual(u)
    for b in range(B):
        assert rel(q0[b].cpu().numpy(), s.q[b].cpu().numpy()) < 1e-5
    a = s.run(init=u, dual_init=q0).cpu().numpy()
    r = s.run(init=u, dual_init="residual").cpu().numpy()
    assert rel(a, r) < 1e-5
    with pytest.raises(ValueError):
        s.run(init=u, dual_init=q0[:1])

