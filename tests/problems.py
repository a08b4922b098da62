"""Small seeded reconstruction problems shared by the parity tests (k-space from the oracle NUFFT)."""

import numpy as np

from oracle import nufft as on
from oracle import pdhg as op
from toporecon_hip import synth


def oracle_forward(img, coord):
    return on.nufft(img.astype(np.complex128), coord)


def make_problem(shape=(10, 12, 12), coils=3, echoes=2, bins=3, interleaves=60, readouts=40,
                 seed=0, algorithm="tvme", l2=True):
    base = synth.config_small_tvm()
    cfg = synth.config_scaled(base, shape, coils=coils, echoes=echoes, bins=bins,
                              interleaves=interleaves, readouts=readouts, seed=seed)
    cfg.algorithm = algorithm
    cfg.l2_coupling = l2
    acq = synth.make_acquisition(cfg)
    ksp = synth.simulate_kspace_numpy(acq, oracle_forward)
    prob = op.bin_problem(ksp, acq.coord, acq.dcf, acq.mps, acq.resp, acq.bin_edges)
    return acq, ksp, prob


def rel(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))
