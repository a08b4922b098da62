"""Write tests/golden/oracle_vectors.npz: small input/output vectors of the float64 oracle.

The reference's own tests hold no vectors for the NUFFT or the PDHG steps (their SigPy /
CuPy dependencies are not installed here), so these fixtures freeze the oracle's outputs on
seeded inputs: they pin the oracle against later edits and give the GPU tests fixed targets.
Contents (all seeded, sizes chosen so the oracle runs in seconds):

* NUFFT at 12x16x16, one bin of 400 samples including k = 0, an on-grid point and the
  k-space corner, for 2 channels: forward values and the density-weighted adjoint image;
* a brute-force NUDFT of the same image at 8x8x8 on 64 samples (the NUFFT's approximation
  target, for the error bound the oracle tests assert);
* TVM / TVME / TVMW states after 1 and 5 iterations at 8x8x8, 3 bins, 3 echoes, 2 coils.

Regenerate with ``python tests/golden/make_oracle_golden.py`` (deterministic).
"""

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, REPO)
sys.path.insert(0, os.path.join(REPO, "tests"))
sys.path.insert(0, os.path.join(REPO, "multinodemri-miccai26_amd"))

from oracle import nufft as on  # noqa: E402
from oracle import pdhg as op  # noqa: E402
from problems import make_problem  # noqa: E402


def nufft_case():
    rng = np.random.default_rng(20261)
    shape = (12, 16, 16)
    half = np.array(shape) / 2.0
    c = (rng.uniform(-1, 1, (400, 3)) * half).astype(np.float32)
    c[0] = 0.0
    c[1] = np.round(c[1])
    c[2] = -half
    # inputs are stored as complex64; the outputs are computed from exactly those values
    img = (rng.standard_normal((2,) + shape) + 1j * rng.standard_normal((2,) + shape)).astype(np.complex64)
    k = (rng.standard_normal((2, 400)) + 1j * rng.standard_normal((2, 400))).astype(np.complex64)
    dcf = rng.uniform(0.1, 1.0, 400).astype(np.float32)
    fwd = np.stack([on.nufft(img[i].astype(np.complex128), c) for i in range(2)])
    adj = np.stack([on.nufft_adjoint(k[i].astype(np.complex128) * dcf, c, shape) for i in range(2)])
    return {"nufft_coord": c, "nufft_img": img, "nufft_ksp": k,
            "nufft_dcf": dcf, "nufft_fwd": fwd, "nufft_adj": adj}


def nudft_case():
    rng = np.random.default_rng(20262)
    shape = (8, 8, 8)
    c = (rng.uniform(-1, 1, (64, 3)) * 4.0).astype(np.float32)
    img = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    return {"nudft_coord": c, "nudft_img": img.astype(np.complex64),
            "nudft_out": on.nudft(img.astype(np.complex64).astype(np.complex128), c)}


def pdhg_cases():
    out = {}
    for alg in ("tvm", "tvme", "tvmw"):
        acq, ksp, prob = make_problem(shape=(8, 8, 8), coils=2, echoes=3, bins=3, interleaves=48, readouts=24,
                                      seed=20263, algorithm=alg)
        kw = dict(lam_m=2e-3, lam_e=2e-3, lam1=2e-3, lam2=2e-3, lam3=2e-3, l2_coupling=True, tol=0.0)
        for n in (1, 5):
            ubar, _, _ = op.run(prob, alg, max_iter=n, **kw)
            out[f"{alg}_ubar_{n}"] = ubar
    return out


def main():
    d = {}
    d.update(nufft_case())
    d.update(nudft_case())
    d.update(pdhg_cases())
    path = os.path.join(HERE, "oracle_vectors.npz")
    np.savez_compressed(path, **d)
    print(path, {k: v.shape for k, v in d.items()})


if __name__ == "__main__":
    main()
