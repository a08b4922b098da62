"""Respiratory self-gating (CPU): parity with the reference's own outputs on the golden vectors
(tests/golden/respiratory_vectors.npz, written by tests/golden/make_respiratory_golden.py from
src/toporecon/preprocessing/respiratory.py), a known answer (the recovered trace follows the
breathing modulation), the validation errors, and the dataset entry point / CLI writing resp.cfl."""

import os

import numpy as np
import pytest

from toporecon_hip.data import cfl
from toporecon_hip.mains import cli_main
from toporecon_hip.preprocessing import RespiratoryConfig, estimate_respiratory_from_dataset, \
    estimate_respiratory_signal

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "respiratory_vectors.npz")


@pytest.fixture(scope="module")
def gold():
    return np.load(GOLDEN)


def _close(got, want, tol=2e-5):
    return float(np.max(np.abs(got.astype(np.float64) - want))) < tol


@pytest.mark.parametrize("case", ["3echo", "1echo", "custom"])
def test_matches_reference_vectors(gold, case):
    x, tr = gold["centres"], float(gold["tr_seconds"])
    # readout samples past the first are ignored: append a random one
    extra = np.random.default_rng(1).standard_normal(x.shape[:3] + (1,)).astype(np.complex64)
    k = np.concatenate([x, extra], axis=3)
    if case == "3echo":
        got = estimate_respiratory_signal(k, tr)
    elif case == "1echo":
        got = estimate_respiratory_signal(k[:1], tr)
    else:
        lo, hi, olp, ohp = gold["custom"]
        got = estimate_respiratory_signal(k, tr, config=RespiratoryConfig(float(lo), float(hi), int(olp), int(ohp)))
    want = gold[f"resp_{case}"]
    assert got.dtype == np.float32 and got.shape == want.shape
    assert got.min() == 0.0 and got.max() == 1.0
    # the same NumPy / SciPy calls in the same order: identical bits to the reference's outputs
    assert np.array_equal(got, want), float(np.max(np.abs(got - want)))


def test_recovers_breathing_trace(gold):
    """Known answer: the trace follows the breathing modulation of the magnitudes (up to the
    principal component's sign; 0.984 here, where 10 s hold 2.7 cycles and the 0.1 Hz high-pass
    bends them slightly), and a band that excludes it (3-20 Hz) recovers none of it."""
    tr = float(gold["tr_seconds"])
    t = np.arange(gold["centres"].shape[2]) * tr
    breath = np.sin(2 * np.pi * 0.27 * t + 0.4) + 0.3 * np.sin(2 * np.pi * 0.54 * t + 1.1)
    got = estimate_respiratory_signal(gold["centres"], tr)
    r = np.corrcoef(got[200:-200], breath[200:-200])[0, 1]
    assert abs(r) > 0.97, r
    off = estimate_respiratory_signal(gold["centres"], tr, config=RespiratoryConfig(low_cut_hz=3.0, high_cut_hz=20.0))
    assert abs(np.corrcoef(off[200:-200], breath[200:-200])[0, 1]) < 0.2


def test_validation_errors():
    k = np.ones((1, 2, 50, 3), np.complex64)
    with pytest.raises(ValueError, match=r"\[echo,coil,trajectory,readout\]"):
        estimate_respiratory_signal(k[0], 0.005)
    with pytest.raises(ValueError, match="tr_seconds must be positive"):
        estimate_respiratory_signal(k, 0.0)
    with pytest.raises(ValueError, match="Nyquist"):
        estimate_respiratory_signal(k, 0.5)  # 1 Hz sampling: 2 Hz cut-off above Nyquist
    with pytest.raises(ValueError, match="greater than low_cut_hz"):
        estimate_respiratory_signal(k, 0.005, config=RespiratoryConfig(low_cut_hz=1.0, high_cut_hz=0.5))
    with pytest.raises(ValueError, match="low_cut_hz must be positive"):
        estimate_respiratory_signal(k, 0.005, config=RespiratoryConfig(low_cut_hz=0.0))
    with pytest.raises(ValueError, match="constant"):
        estimate_respiratory_signal(np.ones((2, 3, 400, 1), np.complex64), 0.005)


def test_dataset_entry_point_and_cli(gold, tmp_path):
    x = gold["centres"]
    E, C, T, _ = x.shape
    raw = tmp_path / "raw"
    raw.mkdir()
    k = np.concatenate([x, np.zeros_like(x)], axis=3)  # (E, C, T, 2)
    cfl.write_cfl(raw / "ksp", k.reshape(1, 1, 1, E, 1, C, T, 2, 1))
    cfl.write_cfl(raw / "ktraj", np.zeros((1, 1, 1, 1, 1, 1, T, 2, 3), np.complex64))
    cfl.write_cfl(raw / "dens", np.ones((1, 1, 1, 1, 1, 1, T, 2, 1), np.complex64))
    (raw / "imageDim.txt").write_text("8:8:8\n")
    (raw / "voxelSize.txt").write_text("1:1:1\n")
    (raw / "tr.txt").write_text(f"{float(gold['tr_seconds'])}\n")
    out = estimate_respiratory_from_dataset(raw, tmp_path / "work")
    resp = cfl.read_cfl(out)
    assert resp.shape == (T, 1)
    assert _close(resp[:, 0].real, gold["resp_3echo"])
    assert cli_main(["resp", str(raw), "--work-dir", str(tmp_path / "work2")]) == 0
    assert np.array_equal(cfl.read_cfl(tmp_path / "work2" / "resp"), resp)
    (raw / "tr.txt").unlink()
    with pytest.raises(FileNotFoundError, match="missing required files"):
        estimate_respiratory_from_dataset(raw, tmp_path / "work3")
