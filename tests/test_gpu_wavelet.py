"""HIP db6 zero-mode wavelet and TVMW parity against the float64 oracle."""

import numpy as np
import pytest

from oracle import pdhg as op
from oracle import wavelet as ow
from problems import make_problem, rel

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_mod():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def _flat(coeffs):
    return np.concatenate([a.reshape(-1) for a in ow.tree_leaves(coeffs)])


def _unflat(flat, shape, level):
    shapes = ow.level_shapes(shape, level)
    out, off = [], 0
    s = shapes[-1]
    n = int(np.prod(s))
    out.append(flat[off:off + n].reshape(s))
    off += n
    for lv in range(level, 0, -1):
        s = shapes[lv]
        n = int(np.prod(s))
        d = {}
        for k in ow._KEYS:
            d[k] = flat[off:off + n].reshape(s)
            off += n
        out.append(d)
    assert off == flat.size
    return out


@pytest.mark.parametrize("shape", [(22, 24, 24), (44, 46, 45), (23, 30, 25), (12, 16, 16)])
def test_w2_forward_adjoint_match_oracle(torch_mod, shape):
    torch = torch_mod
    from toporecon_hip import _lib
    from toporecon_hip.algorithms.tvmw import db6_max_level
    lib = _lib.load()
    level = db6_max_level(shape)
    assert level == ow.max_level(shape)
    nimg = 3
    rng = np.random.default_rng(sum(shape))
    x = (rng.standard_normal((nimg,) + shape) + 1j * rng.standard_normal((nimg,) + shape)).astype(np.complex64)
    ncoef = lib.tr_w2_coeff_count(*shape, level)
    nwork = lib.tr_w2_work_count(*shape, level)
    xt = torch.from_numpy(x).cuda()
    ct = torch.empty((nimg, ncoef), dtype=torch.complex64, device="cuda")
    wt = torch.empty((max(nimg * nwork, 1),), dtype=torch.complex64, device="cuda")
    _lib.call("tr_w2_forward", xt.data_ptr(), nimg, *shape, level, ct.data_ptr(), wt.data_ptr(),
              _lib.stream_handle())
    c = ct.cpu().numpy()
    for i in range(nimg):
        ref = _flat(ow.wavedec3(x[i].astype(np.complex128), level))
        assert ref.size == ncoef
        assert rel(c[i], ref) < 1e-6
    # adjoint of independent coefficients
    d = (rng.standard_normal((nimg, ncoef)) + 1j * rng.standard_normal((nimg, ncoef))).astype(np.complex64)
    dt = torch.from_numpy(d).cuda()
    out = torch.empty((nimg,) + shape, dtype=torch.complex64, device="cuda")
    _lib.call("tr_w2_adjoint", dt.data_ptr(), nimg, *shape, level, out.data_ptr(), wt.data_ptr(),
              _lib.stream_handle())
    o = out.cpu().numpy()
    for i in range(nimg):
        ref = ow.waverec3(_unflat(d[i].astype(np.complex128), shape, level), shape)
        assert rel(o[i], ref) < 1e-6


def test_w2_dual_update(torch_mod):
    torch = torch_mod
    from toporecon_hip import _lib
    rng = np.random.default_rng(3)
    p = (rng.standard_normal(5000) + 1j * rng.standard_normal(5000)).astype(np.complex64)
    w = (rng.standard_normal(5000) + 1j * rng.standard_normal(5000)).astype(np.complex64)
    for lam in (0.3, 0.0):
        pt = torch.from_numpy(p).cuda()
        wt = torch.from_numpy(w).cuda()
        _lib.call("tr_w2_dual_update", pt.data_ptr(), wt.data_ptr(), 5000, 1 / 6, lam, _lib.stream_handle())
        ref = ow.tree_project_linf([p.astype(np.complex128) + (1 / 6) * w], lam)[0]
        assert np.abs(pt.cpu().numpy() - ref).max() < 1e-6


def _tvmw_solver(acq, ksp, iters, tol, l2, lam=(2e-3, 2e-3, 2e-3)):
    from toporecon_hip.algorithms import TvmwReconstructor
    B = len(acq.bin_edges) - 1
    return TvmwReconstructor(ksp, acq.coord, acq.dcf, acq.mps, acq.resp, np.zeros_like(ksp), B, l2,
                             lambda1=lam[0], lambda2=lam[1], lambda3=lam[2], sigma=1 / 6, tau=1 / 6,
                             max_iter=iters, tol=tol, margin=0, device=0, bin_edges=acq.bin_edges,
                             show_pbar=False)


@pytest.mark.parametrize("iters,tol", [(1, 1e-5), (4, 3e-5)])
@pytest.mark.parametrize("l2", [False, True])
def test_tvmw_iterates_match_oracle(torch_mod, iters, tol, l2):
    print("tvmw iterates", iters, l2, flush=True)
    acq, ksp, prob = make_problem(shape=(22, 24, 24), coils=2, echoes=3, bins=2, interleaves=60,
                                  readouts=48, seed=11, algorithm="tvmw", l2=l2)
    s = _tvmw_solver(acq, ksp, iters, 0.0, l2)
    assert s.level_w2 == 1
    ubar = s.run().cpu().numpy()
    hist = []
    o_ubar, _, st = op.run(prob, "tvmw", lam1=2e-3, lam2=2e-3, lam3=2e-3, l2_coupling=l2, max_iter=iters,
                           tol=0.0, history=hist)
    assert rel(ubar, o_ubar) < tol
    assert rel(s.u.cpu().numpy(), st.u) < tol
    assert s.history[-1] == pytest.approx(hist[-1], rel=1e-3)


def test_tvmw_level0_and_zero_lambdas(torch_mod):
    acq, ksp, prob = make_problem(shape=(10, 12, 12), coils=2, echoes=2, bins=2, seed=12, algorithm="tvmw")
    s = _tvmw_solver(acq, ksp, 3, 0.0, False, lam=(2e-3, 0.0, 0.0))
    assert s.level_w2 == 0
    ubar = s.run().cpu().numpy()
    o_ubar, _, _ = op.run(prob, "tvmw", lam1=2e-3, lam2=0.0, lam3=0.0, max_iter=3, tol=0.0)
    assert rel(ubar, o_ubar) < 2e-5
