"""Pins the float64 PDHG restatement (oracle/pdhg.py) by identities the mathematics fixes."""

import numpy as np
import pytest

from oracle import nufft as on
from oracle import pdhg as op


def _c(rng, shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


@pytest.mark.parametrize("B", [1, 2, 5])
def test_motion_difference_adjoints(B):
    rng = np.random.default_rng(B)
    u = _c(rng, (B, 3, 4, 5, 6))
    p = _c(rng, (B, 3, 4, 5, 6))
    lhs = np.vdot(op.motion_diff(u), p)
    # TVME divergence is +D^T p; TVM's is -D^T p (tvm.py:342-357, tvme.py:421-439).
    assert np.vdot(u, op.motion_div_tvme(p)) == pytest.approx(lhs, rel=1e-12, abs=1e-12)
    assert np.vdot(u, -op.motion_div_tvm(p)) == pytest.approx(lhs, rel=1e-12, abs=1e-12)
    # TVMW's divergence has no last-bin rule; it equals -D^T p when p[B-1] = 0, which the
    # iteration preserves because the last motion difference is 0.
    p[-1] = 0
    lhs = np.vdot(op.motion_diff(u), p)
    assert np.vdot(u, -op.motion_div_tvmw(p)) == pytest.approx(lhs, rel=1e-12, abs=1e-12)


@pytest.mark.parametrize("E", [1, 2, 6])
def test_echo_difference_adjoint(E):
    rng = np.random.default_rng(10 + E)
    u = _c(rng, (2, E, 3, 4, 5))
    p = _c(rng, (2, E, 3, 4, 5))
    if E == 1:
        # One echo: the difference is 0, so the echo duals stay 0 and the reference's
        # divp_e[:, 0] = -divp[:, 0] rule (tvme.py:445-446) only ever sees 0.
        assert not np.any(op.echo_diff(u))
        return
    assert np.vdot(op.echo_diff(u), p) == pytest.approx(np.vdot(u, op.echo_div(p)), rel=1e-12, abs=1e-12)


@pytest.mark.parametrize("axis", [-1, -2, -3])
def test_spatial_difference_adjoint(axis):
    rng = np.random.default_rng(20 - axis)
    u = _c(rng, (2, 2, 5, 6, 7))
    p = _c(rng, (2, 2, 5, 6, 7))
    assert np.vdot(op.fdiff(u, axis), p) == pytest.approx(np.vdot(u, op.fdiff_adj(p, axis)), rel=1e-12)


def test_spatial_difference_boundaries():
    x = np.arange(5.0) ** 2 + 0j
    d = op.fdiff(x, -1)
    assert list(d.real) == [1, 3, 5, 7, 0]
    a = op.fdiff_adj(x, -1)
    assert list(a.real) == [-0.0, 0 - 1, 1 - 4, 4 - 9, 9]


@pytest.mark.parametrize("E", [1, 2, 5, 6])
def test_db1_echo_wavelet_adjoint_and_isometry(E):
    rng = np.random.default_rng(30 + E)
    u = _c(rng, (2, E, 3, 4, 5))
    p = _c(rng, (2, E, 3, 4, 5))
    assert np.vdot(op.w1_fwd(u), p) == pytest.approx(np.vdot(u, op.w1_adj(p)), rel=1e-12)
    if E % 2 == 0:
        assert np.allclose(op.w1_adj(op.w1_fwd(u)), u)


def test_prox_maps_are_projections():
    rng = np.random.default_rng(40)
    p = _c(rng, (2, 3, 4, 5, 6))
    for l2 in (False, True):
        q = op.prox_motion(p, 0.7, l2)
        assert np.allclose(op.prox_motion(q, 0.7, l2), q)
        mag = np.sqrt(np.sum(np.abs(q) ** 2, axis=1)) if l2 else np.abs(q)
        assert mag.max() <= 0.7 * (1 + 1e-12)
    x, y, z = op.prox_iso3(p, 2 * p, -p, 0.5)
    assert np.sqrt(np.abs(x) ** 2 + np.abs(y) ** 2 + np.abs(z) ** 2).max() <= 0.5 * (1 + 1e-12)


def _tiny_problem(seed=0, B=2, E=2, C=2, shape=(4, 6, 6), M=260):
    rng = np.random.default_rng(seed)
    half = np.array(shape) / 2.0
    mps = _c(rng, (C,) + shape)
    mps /= np.sqrt(np.sum(np.abs(mps) ** 2, axis=0))
    coords = [(rng.uniform(-1, 1, (M, 3)) * half).astype(np.float32) for _ in range(B)]
    dcf = [rng.uniform(0.2, 1.0, M) for _ in range(B)]
    truth = _c(rng, (B, E) + shape)
    ksp = []
    for b in range(B):
        img = truth[b][:, None] * mps[None]
        y = on.nufft(img, coords[b])
        ksp.append(y + 0.01 * _c(rng, y.shape))
    return op.BinnedProblem(ksp, coords, dcf, mps)


def _weighted_ls(prob):
    """Dense weighted least squares per (b, e): argmin sum_c ||W^1/2 (A_c u - y_c)||^2."""
    Z, Y, X = prob.img_shape
    N = Z * Y * X
    eye = np.eye(N).reshape(N, Z, Y, X).astype(complex)
    out = np.zeros((prob.B, prob.E, Z, Y, X), complex)
    for b in range(prob.B):
        cols = on.nufft(eye[:, None] * prob.mps[None], prob.coord[b])  # (N, C, M)
        A = np.transpose(cols, (1, 2, 0)).reshape(-1, N)               # (C*M, N)
        w = np.tile(prob.dcf[b], prob.C)
        H = A.conj().T @ (w[:, None] * A)
        for e in range(prob.E):
            rhs = A.conj().T @ (w * prob.ksp[b][e].reshape(-1))
            out[b, e] = np.linalg.solve(H, rhs).reshape(Z, Y, X)
    return out


@pytest.mark.parametrize("alg", ["tvm", "tvme", "tvmw"])
def test_without_regularisation_converges_to_weighted_least_squares(alg):
    prob = _tiny_problem(1)
    ref = _weighted_ls(prob)
    kw = dict(lam_m=1e-14, lam_e=1e-14, lam2=0.0, lam3=0.0, max_iter=400, tol=0.0)
    ubar, iters, st = op.run(prob, alg, **kw)
    rel = np.linalg.norm(st.u - ref) / np.linalg.norm(ref)
    assert rel < 1e-5, rel


def test_regularised_objective_below_least_squares_objective():
    prob = _tiny_problem(2)
    ref = _weighted_ls(prob)
    kw = dict(lam_m=2e-3, lam_e=2e-3, l2_coupling=True)
    ubar, iters, st = op.run(prob, "tvme", max_iter=300, tol=0.0, **kw)
    phi = op.objective(prob, st.u, "tvme", **kw)
    phi_ls = op.objective(prob, ref, "tvme", **kw)
    assert phi <= phi_ls * (1 + 1e-9)


def test_stopping_rules():
    u0 = np.ones((1, 1, 2, 2, 2), complex)
    assert op.rel_change_tvm(u0, u0) == 0.0
    assert op.rel_change_tvm(np.zeros_like(u0), u0) == float("inf")
    assert op.rel_change_tvmw(np.zeros_like(u0), u0) == pytest.approx(np.sqrt(8) / 1e-12)


def test_run_resumes_from_a_state():
    """run(init=state) continues a solve: 1 + 2 iterations equal 3 in one go, bit for bit."""
    from problems import make_problem
    _, _, prob = make_problem(shape=(6, 8, 8), coils=2, echoes=2, bins=2, interleaves=30, readouts=16, seed=5,
                              algorithm="tvmw")
    kw = dict(lam1=2e-3, lam2=2e-3, lam3=2e-3, l2_coupling=True, tol=0.0)
    u3, n3, _ = op.run(prob, "tvmw", max_iter=3, **kw)
    _, _, st = op.run(prob, "tvmw", max_iter=1, **kw)
    u12, _, _ = op.run(prob, "tvmw", max_iter=2, init=st, **kw)
    assert n3 == 3 and np.array_equal(u3, u12)
