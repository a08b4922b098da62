"""Respiratory self-gating with the band-pass on the GPU (``tr_resp_bandpass``): the reference's
recorded outputs (tests/golden/respiratory_vectors.npz) bit for bit, SciPy's ``sosfiltfilt`` chain
bit for bit on every trace, at the benchmark acquisition's size too (32 coils x 6 echoes x 9000
trajectories), and the dataset entry point / CLI with --device."""

import os
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "respiratory_vectors.npz")


@pytest.fixture(scope="module")
def torch_mod():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


@pytest.fixture(scope="module")
def gold():
    return np.load(GOLDEN)


@pytest.mark.parametrize("case", ["3echo", "1echo", "custom"])
def test_matches_reference_vectors(torch_mod, gold, case):
    from toporecon_hip.preprocessing import RespiratoryConfig, estimate_respiratory_signal
    print("resp golden", case, flush=True)
    x, tr = gold["centres"], float(gold["tr_seconds"])
    if case == "3echo":
        got = estimate_respiratory_signal(x, tr, device=0)
    elif case == "1echo":
        got = estimate_respiratory_signal(x[:1], tr, device=0)
    else:
        lo, hi, olp, ohp = gold["custom"]
        got = estimate_respiratory_signal(x, tr, config=RespiratoryConfig(float(lo), float(hi), int(olp), int(ohp)),
                                          device=0)
    want = gold[f"resp_{case}"]
    assert got.dtype == np.float32 and np.array_equal(got, want), float(np.max(np.abs(got - want)))


@pytest.mark.parametrize("shape", [(3, 4, 700), (6, 32, 9000)])
def test_bandpass_matches_scipy_on_every_trace(torch_mod, shape):
    from scipy import signal

    from toporecon_hip.preprocessing import bandpass_hip
    print("resp band-pass", shape, flush=True)
    rng = np.random.default_rng(shape[-1])
    t = np.arange(shape[-1]) * 0.005
    x = (4.0 + np.sin(2 * np.pi * 0.25 * t) + 0.2 * rng.standard_normal(shape)).astype(np.float32)
    x -= x.mean(axis=-1, keepdims=True)
    lo = signal.butter(10, 2.0, "low", fs=200.0, output="sos")
    hi = signal.butter(5, 0.1, "high", fs=200.0, output="sos")
    bandpass_hip(x[:1, :1], lo, hi, 0)  # load the library and the kernel once
    t0 = time.perf_counter()
    got = bandpass_hip(x, lo, hi, 0)
    t_gpu = time.perf_counter() - t0
    sample = [(0, 0), (shape[0] - 1, shape[1] - 1), (shape[0] // 2, shape[1] // 3)]
    t0 = time.perf_counter()
    want = {ij: signal.sosfiltfilt(hi, signal.sosfiltfilt(lo, x[ij])).astype(np.float32) for ij in sample}
    t_cpu = (time.perf_counter() - t0) / len(sample) * shape[0] * shape[1]
    print(f"GPU {t_gpu * 1e3:.1f} ms (host to host) for {shape[0] * shape[1]} traces; SciPy {t_cpu * 1e3:.1f} ms "
          "(scaled from 3 traces)", flush=True)
    if shape[-1] < 1000:
        for i in range(shape[0]):
            for j in range(shape[1]):
                w = signal.sosfiltfilt(hi, signal.sosfiltfilt(lo, x[i, j])).astype(np.float32)
                assert np.array_equal(got[i, j], w), (i, j)
    for ij, w in want.items():
        assert np.array_equal(got[ij], w), ij


def test_refuses_traces_shorter_than_the_padding(torch_mod):
    from scipy import signal

    from toporecon_hip.preprocessing import bandpass_hip
    lo = signal.butter(10, 2.0, "low", fs=200.0, output="sos")
    hi = signal.butter(5, 0.1, "high", fs=200.0, output="sos")
    with pytest.raises(ValueError, match="padlen"):
        bandpass_hip(np.zeros((2, 33), np.float32), lo, hi, 0)


def test_dataset_entry_point_and_cli_on_the_gpu(torch_mod, gold, tmp_path):
    from toporecon_hip.data import cfl
    from toporecon_hip.mains import cli_main
    x = gold["centres"]
    E, C, T, _ = x.shape
    raw = tmp_path / "raw"
    raw.mkdir()
    k = np.concatenate([x, np.zeros_like(x)], axis=3)
    cfl.write_cfl(raw / "ksp", k.reshape(1, 1, 1, E, 1, C, T, 2, 1))
    cfl.write_cfl(raw / "ktraj", np.zeros((1, 1, 1, 1, 1, 1, T, 2, 3), np.complex64))
    cfl.write_cfl(raw / "dens", np.ones((1, 1, 1, 1, 1, 1, T, 2, 1), np.complex64))
    (raw / "imageDim.txt").write_text("8:8:8\n")
    (raw / "voxelSize.txt").write_text("1:1:1\n")
    (raw / "tr.txt").write_text(f"{float(gold['tr_seconds'])}\n")
    assert cli_main(["resp", str(raw), "--work-dir", str(tmp_path / "w"), "--device", "0"]) == 0
    resp = cfl.read_cfl(tmp_path / "w" / "resp")
    assert resp.shape == (T, 1)
    assert np.array_equal(resp[:, 0].real, gold["resp_3echo"])
