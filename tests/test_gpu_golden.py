"""The HIP path against the committed oracle vectors (tests/golden/oracle_vectors.npz)."""

import os

import numpy as np
import pytest

from problems import make_problem, rel

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "oracle_vectors.npz")


@pytest.fixture(scope="module")
def gold():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    with np.load(GOLD, allow_pickle=False) as f:
        return {k: f[k] for k in f.files}


def test_nufft_matches_golden(gold):
    from toporecon_hip.nufft_backend import create_nufft_backend
    c = gold["nufft_coord"]
    be = create_nufft_backend("hip", img_shape=gold["nufft_img"].shape[1:], coordinates=[c], device=0)
    for i in range(2):
        assert rel(be.forward(gold["nufft_img"][i], 0).cpu().numpy(), gold["nufft_fwd"][i]) < 1e-5
        k = (gold["nufft_ksp"][i] * gold["nufft_dcf"]).astype(np.complex64)
        assert rel(be.adjoint(k, 0).cpu().numpy(), gold["nufft_adj"][i]) < 1e-5


@pytest.mark.parametrize("alg", ["tvm", "tvme", "tvmw"])
def test_pdhg_states_match_golden(gold, alg):
    from toporecon_hip.algorithms import TvmeReconstructor, TvmReconstructor, TvmwReconstructor
    acq, ksp, _ = make_problem(shape=(8, 8, 8), coils=2, echoes=3, bins=3, interleaves=48, readouts=24,
                               seed=20263, algorithm=alg)
    B = len(acq.bin_edges) - 1
    common = dict(sigma=1 / 6, tau=1 / 6, tol=0.0, margin=0, device=0, bin_edges=acq.bin_edges, show_pbar=False)
    dq = np.zeros_like(ksp)
    for n in (1, 5):
        if alg == "tvm":
            s = TvmReconstructor(ksp, acq.coord, acq.dcf, acq.mps, acq.resp, dq, B, True, lamda=2e-3,
                                 max_iter=n, **common)
        elif alg == "tvme":
            s = TvmeReconstructor(ksp, acq.coord, acq.dcf, acq.mps, acq.resp, dq, B, True, lamda_m=2e-3,
                                  lamda_e=2e-3, max_iter=n, **common)
        else:
            s = TvmwReconstructor(ksp, acq.coord, acq.dcf, acq.mps, acq.resp, dq, B, True, lambda1=2e-3,
                                  lambda2=2e-3, lambda3=2e-3, max_iter=n, **common)
        assert rel(s.run().cpu().numpy(), gold[f"{alg}_ubar_{n}"]) < (1e-5 if n == 1 else 2e-5), (alg, n)
