"""Write tests/golden/tvm48_solve.npz: the float64 oracle's plain TVM solve at 48x64x64.

The reference's CPU-runnable configuration: TVM at 48x64x64, 8 coils, 2 echoes (TE 1.2, 2.6 ms),
4 respiratory bins, 1536 cone interleaves x 256 samples (readout fraction 0.98, 2/98 % outlier
drop), lambda_motion 1e-5, tol 1e-3, at most 100 iterations, sigma = tau = 1/6, k-space simulated
by the float64 oracle NUFFT plus 30 dB noise (tests/problems.py generators, seed 0).  The whole
solve takes about a quarter of an hour in float64, too long for a test, so its result is frozen
here: the iteration count, the stopping values, the norm of the final image, the final image at
20,000 seeded voxels, and its projections on 4 seeded random volumes.

Regenerate with ``python tests/golden/make_tvm48_golden.py`` (deterministic).
"""

import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, REPO)
sys.path.insert(0, os.path.join(REPO, "tests"))
sys.path.insert(0, os.path.join(REPO, "multinodemri-miccai26_amd"))

from oracle import pdhg as op  # noqa: E402
from problems import oracle_forward  # noqa: E402
from toporecon_hip import synth  # noqa: E402

VOXEL_SEED = 48064064
PROJ_SEEDS = (1, 2, 3, 4)
NVOX = 20000


def sample_indices(n_total):
    return np.sort(np.random.default_rng(VOXEL_SEED).choice(n_total, NVOX, replace=False))


def projection_volume(seed, shape):
    rng = np.random.default_rng(seed)
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


def main():
    cfg = synth.config_small_tvm()
    acq = synth.make_acquisition(cfg)
    t0 = time.time()
    ksp = synth.simulate_kspace_numpy(acq, oracle_forward)
    prob = op.bin_problem(ksp, acq.coord, acq.dcf, acq.mps, acq.resp, acq.bin_edges)
    hist = []
    ubar, iters, _ = op.run(prob, "tvm", lam_m=cfg.lambda_motion, l2_coupling=cfg.l2_coupling,
                            max_iter=cfg.max_iter, tol=cfg.tol, history=hist)
    print(f"{iters} iterations in {time.time() - t0:.0f} s, last stopping value {hist[-1]:.4e}")
    idx = sample_indices(ubar.size)
    proj = np.array([np.vdot(projection_volume(s, ubar.shape), ubar) for s in PROJ_SEEDS])
    np.savez_compressed(os.path.join(HERE, "tvm48_solve.npz"), iterations=iters, history=np.array(hist),
                        norm=np.linalg.norm(ubar), voxel_index=idx, voxels=ubar.reshape(-1)[idx],
                        projections=proj, shape=np.array(ubar.shape))


if __name__ == "__main__":
    main()
