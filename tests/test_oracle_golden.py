"""The oracle reproduces its committed vectors (tests/golden/oracle_vectors.npz), and the
NUFFT stays within the Kaiser-Bessel error of the brute-force NUDFT on the fixture."""

import os

import numpy as np
import pytest

from oracle import nufft as on
from oracle import pdhg as op
from problems import make_problem

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "oracle_vectors.npz")


@pytest.fixture(scope="module")
def gold():
    with np.load(GOLD, allow_pickle=False) as f:
        return {k: f[k] for k in f.files}


def test_nufft_vectors(gold):
    c, img, k, dcf = gold["nufft_coord"], gold["nufft_img"], gold["nufft_ksp"], gold["nufft_dcf"]
    fwd = np.stack([on.nufft(img[i].astype(np.complex128), c) for i in range(2)])
    adj = np.stack([on.nufft_adjoint(k[i].astype(np.complex128) * dcf, c, img.shape[1:]) for i in range(2)])
    np.testing.assert_allclose(fwd, gold["nufft_fwd"], rtol=1e-10, atol=1e-12)
    np.testing.assert_allclose(adj, gold["nufft_adj"], rtol=1e-10, atol=1e-12)


def test_nudft_vector_and_kb_error_bound(gold):
    c, img = gold["nudft_coord"], gold["nudft_img"].astype(np.complex128)
    exact = on.nudft(img, c)
    np.testing.assert_allclose(exact, gold["nudft_out"], rtol=1e-12, atol=1e-12)
    approx = on.nufft(img, c)
    assert np.linalg.norm(approx - exact) / np.linalg.norm(exact) < 2e-2


@pytest.mark.parametrize("alg", ["tvm", "tvme", "tvmw"])
def test_pdhg_states(gold, alg):
    acq, ksp, prob = make_problem(shape=(8, 8, 8), coils=2, echoes=3, bins=3, interleaves=48, readouts=24,
                                  seed=20263, algorithm=alg)
    kw = dict(lam_m=2e-3, lam_e=2e-3, lam1=2e-3, lam2=2e-3, lam3=2e-3, l2_coupling=True, tol=0.0)
    for n in (1, 5):
        ubar, _, _ = op.run(prob, alg, max_iter=n, **kw)
        np.testing.assert_allclose(ubar, gold[f"{alg}_ubar_{n}"], rtol=1e-9, atol=1e-12)
