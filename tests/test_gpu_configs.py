"""Parity at the two benchmark configurations.

TVM 48x64x64, 8 coils, 2 echoes, 4 bins, 1536x256 cone samples (the reference's CPU-runnable case):
one full PDHG iteration and the data-term operators against the float64 oracle.

TVME 80x192x192, 32 coils, 6 echoes, 6 bins, 9000x512 cone samples (the benchmark workload, too large
for the oracle): size-independent properties of the batched NUFFT on a full bin -- the normal
equation <x, A^H W A x> = ||A x||_W^2, linearity, bitwise reruns, and bitwise equality with the
k-space support pruning switched off.
"""

import numpy as np
import pytest

from oracle import pdhg as op
from problems import oracle_forward, rel

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_mod():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def test_tvm_48x64x64_iteration_matches_oracle(torch_mod):
    torch = torch_mod
    from toporecon_hip import _lib, synth
    from toporecon_hip.algorithms import TvmReconstructor
    print("tvm 48x64x64 c=8 e=2 b=4", flush=True)
    cfg = synth.config_small_tvm()
    acq = synth.make_acquisition(cfg)
    ksp = synth.simulate_kspace_numpy(acq, oracle_forward)
    prob = op.bin_problem(ksp, acq.coord, acq.dcf, acq.mps, acq.resp, acq.bin_edges)
    B = len(acq.bin_edges) - 1
    lam = 2e-3
    s = TvmReconstructor(ksp, acq.coord, acq.dcf, acq.mps, acq.resp, np.zeros_like(ksp), B, cfg.l2_coupling,
                         lamda=lam, sigma=1 / 6, tau=1 / 6, max_iter=1, tol=0.0, margin=0, device=0,
                         bin_edges=acq.bin_edges, show_pbar=False)
    s.pdhg_init()
    E, C, K = s.E, s.C, s.E * s.C
    rng = np.random.default_rng(3)
    dims = s.vol_dims()
    xn = (rng.standard_normal(dims) + 1j * rng.standard_normal(dims)).astype(np.complex64)
    x = torch.from_numpy(xn).cuda()
    ref = op.forward_op(prob, xn.astype(np.complex128))
    got = []
    for b in range(B):
        M = s.y[b].shape[0]
        a_sorted = torch.empty_like(s.y[b])
        _lib.call("tr_data_simulate", s.plan.handle, b, x[b].data_ptr(), s.maps.data_ptr(), E, C,
                  a_sorted.data_ptr(), s.stream)
        a_orig = torch.empty((K, M), dtype=torch.complex64, device="cuda")
        _lib.call("tr_plan_scatter_samples", s.plan.handle, b, a_sorted.data_ptr(), K, a_orig.data_ptr(), s.stream)
        got.append(a_orig.cpu().numpy().reshape(E, C, M))
    e_fwd = rel(np.concatenate([g.ravel() for g in got]), np.concatenate([r.ravel() for r in ref]))
    g = torch.empty_like(x)
    for b in range(B):
        _lib.call("tr_data_adjoint", s.plan.handle, b, s.y[b].data_ptr(), s.dcf[b].data_ptr(), s.maps.data_ptr(),
                  E, C, g[b].data_ptr(), 0, s.stream)
    e_adj = rel(g.cpu().numpy(), op.adjoint_op(prob, prob.ksp))
    ub = s.run().cpu().numpy()
    o_ub, _, _ = op.run(prob, "tvm", lam_m=lam, l2_coupling=cfg.l2_coupling, max_iter=1, tol=0.0)
    e_it = rel(ub, o_ub)
    print(f"A {e_fwd:.2e}  A^H W {e_adj:.2e}  ubar after 1 iteration {e_it:.2e}", flush=True)
    assert e_fwd < 1e-5 and e_adj < 1e-5
    assert e_it < 1e-5


def test_tvme_80x192x192_full_bin_properties(torch_mod):
    torch = torch_mod
    from toporecon_hip import _lib, synth
    from toporecon_hip.algorithms import TvmeReconstructor
    print("tvme 80x192x192 c=32 e=6 b=6", flush=True)
    cfg = synth.config_full_tvme()
    acq = synth.make_acquisition(cfg)
    s = TvmeReconstructor(None, acq.coord, acq.dcf, acq.mps, acq.resp, None, cfg.bins, cfg.l2_coupling,
                          lamda_m=cfg.lambda_motion, lamda_e=cfg.lambda_echo, sigma=1 / 6, tau=1 / 6,
                          max_iter=1, tol=0.0, margin=0, device=0, bin_edges=acq.bin_edges, show_pbar=False,
                          echoes=len(cfg.echo_times_ms))
    E, C = s.E, s.C
    st = s.stream
    b = 0
    gen = torch.Generator(device="cuda")
    gen.manual_seed(5)
    shape = (E,) + tuple(s.img_shape)

    def rand_img():
        return torch.complex(torch.randn(shape, generator=gen, device="cuda"),
                             torch.randn(shape, generator=gen, device="cuda"))

    def fwd(x):
        out = torch.empty_like(s.y[b])
        _lib.call("tr_data_simulate", s.plan.handle, b, x.data_ptr(), s.maps.data_ptr(), E, C, out.data_ptr(), st)
        return out

    def adj(q):
        out = torch.empty(shape, dtype=torch.complex64, device="cuda")
        _lib.call("tr_data_adjoint", s.plan.handle, b, q.data_ptr(), s.dcf[b].data_ptr(), s.maps.data_ptr(),
                  E, C, out.data_ptr(), 0, st)
        return out

    x1, x2 = rand_img(), rand_img()
    a1 = fwd(x1)
    # <x, A^H W A x> = ||A x||_W^2 (float64 sums of the float32 results)
    g1 = adj(a1)
    lhs = torch.sum(x1.conj().to(torch.complex128) * g1.to(torch.complex128))
    w = s.dcf[b].to(torch.float64)[:, None]
    rhs = torch.sum(w * torch.abs(a1.to(torch.complex128)) ** 2)
    print("normal equation", float(lhs.real), float(rhs), float(lhs.imag), flush=True)
    assert abs(float(lhs.real) - float(rhs)) / float(rhs) < 1e-5
    assert abs(float(lhs.imag)) / float(rhs) < 1e-5
    # linearity of the forward operator
    a2 = fwd(x2)
    a12 = fwd(x1 + 2 * x2)
    assert float(torch.linalg.vector_norm(a12 - (a1 + 2 * a2)) / torch.linalg.vector_norm(a12)) < 1e-6
    # bitwise reruns, and bitwise equality with support pruning off
    g1b = adj(a1)
    assert torch.equal(torch.view_as_real(g1), torch.view_as_real(g1b))
    s.plan.set_param("prune", 0)
    try:
        assert torch.equal(torch.view_as_real(adj(a1)), torch.view_as_real(g1))
        assert torch.equal(torch.view_as_real(fwd(x1)), torch.view_as_real(a1))
    finally:
        s.plan.set_param("prune", 1)


def test_tvm_48x64x64_full_solve_matches_oracle(torch_mod):
    """The reference's CPU-runnable configuration solved to its stopping rule (lambda 1e-5, tol 1e-3,
    at most 100 iterations) against the float64 oracle's solve frozen in tests/golden/tvm48_solve.npz
    (tests/golden/make_tvm48_golden.py): both run the 100-iteration cap, the stopping values agree
    iteration by iteration, and the final image agrees to 1e-4 on 20,000 seeded voxels, in norm and
    in 4 seeded projections (stopping values to 1e-2: they are ratios of small differences)."""
    import os

    torch = torch_mod
    from toporecon_hip import synth
    from toporecon_hip.algorithms import TvmReconstructor
    here = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
    gold = np.load(os.path.join(here, "tvm48_solve.npz"))
    import importlib.util
    spec = importlib.util.spec_from_file_location("make_tvm48_golden", os.path.join(here, "make_tvm48_golden.py"))
    mk = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mk)
    print("tvm 48x64x64 full solve", flush=True)
    cfg = synth.config_small_tvm()
    acq = synth.make_acquisition(cfg)
    ksp = synth.simulate_kspace_numpy(acq, oracle_forward)
    B = len(acq.bin_edges) - 1
    s = TvmReconstructor(ksp, acq.coord, acq.dcf, acq.mps, acq.resp, None, B, cfg.l2_coupling,
                         lamda=cfg.lambda_motion, sigma=1 / 6, tau=1 / 6, max_iter=cfg.max_iter, tol=cfg.tol,
                         margin=0, device=0, bin_edges=acq.bin_edges, show_pbar=False)
    ub = s.run().cpu().numpy().astype(np.complex128)
    assert s.iterations == int(gold["iterations"]) == 100
    hist = np.asarray(s.history)
    assert np.max(np.abs(hist - gold["history"]) / gold["history"]) < 1e-2
    assert tuple(ub.shape) == tuple(gold["shape"])
    assert rel(ub.reshape(-1)[gold["voxel_index"]], gold["voxels"]) < 1e-4
    assert abs(np.linalg.norm(ub) - float(gold["norm"])) / float(gold["norm"]) < 1e-5
    proj = np.array([np.vdot(mk.projection_volume(sd, ub.shape), ub) for sd in mk.PROJ_SEEDS])
    assert np.max(np.abs(proj - gold["projections"]) / np.abs(gold["projections"])) < 1e-3
