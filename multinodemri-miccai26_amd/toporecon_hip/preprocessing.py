"""Respiratory self-gating: the respiratory trace ``resp`` that the solvers bin by.

Restates ``toporecon.preprocessing.respiratory`` (src/toporecon/preprocessing/respiratory.py:15-133):
the magnitude of each trajectory's first readout sample (the k-space centre of a centre-out
trajectory) per echo and coil, mean removed, band-passed forward and backward with Butterworth
second-order sections (low-pass order 10 at 2 Hz, then high-pass order 5 at 0.1 Hz), the leading
principal component over coils per echo (singular value x right singular vector), then over
echoes, normalised to [0, 1].  Host code (NumPy / SciPy), as in the reference: a few thousand
traces of one value each, milliseconds of work, once per dataset.  The sign of the principal
component is the one LAPACK's SVD returns, as in the reference.
"""

from __future__ import annotations

from dataclasses import dataclass
from pathlib import Path

import numpy as np

from .data import cfl


@dataclass(frozen=True)
class RespiratoryConfig:
    """Band-pass settings (respiratory.py:15-33)."""

    low_cut_hz: float = 0.1
    high_cut_hz: float = 2.0
    low_pass_order: int = 10
    high_pass_order: int = 5

    def validate(self, sampling_rate_hz: float) -> None:
        if self.low_cut_hz <= 0:
            raise ValueError("low_cut_hz must be positive.")
        if self.high_cut_hz <= self.low_cut_hz:
            raise ValueError("high_cut_hz must be greater than low_cut_hz.")
        nyquist = sampling_rate_hz / 2
        if self.high_cut_hz >= nyquist:
            raise ValueError(f"high_cut_hz={self.high_cut_hz} must be below the Nyquist frequency {nyquist}.")


def _leading_component(a: np.ndarray) -> np.ndarray:
    """s_0 * v_0 of a (rows x T) float32 matrix: the trace the rows share most."""
    _, s, vt = np.linalg.svd(a, full_matrices=False)
    return s[0] * vt[0]


def estimate_respiratory_signal(kspace, tr_seconds: float, *, config: RespiratoryConfig | None = None) -> np.ndarray:
    """Normalised respiratory trace (float32, one value per trajectory) from canonical k-space
    ``[echo, coil, trajectory, readout]`` (respiratory.py:36-114)."""
    try:
        from scipy import signal
    except ImportError as exc:
        raise RuntimeError("SciPy is required for respiratory signal estimation.") from exc
    k = np.asarray(kspace)
    if k.ndim != 4:
        raise ValueError(f"k-space must have shape [echo,coil,trajectory,readout], got {k.shape}.")
    if tr_seconds <= 0:
        raise ValueError("tr_seconds must be positive.")
    cfg = config or RespiratoryConfig()
    fs = 1.0 / float(tr_seconds)
    cfg.validate(fs)
    low = signal.butter(cfg.low_pass_order, cfg.high_cut_hz, "low", fs=fs, output="sos")
    high = signal.butter(cfg.high_pass_order, cfg.low_cut_hz, "high", fs=fs, output="sos")
    E, C, T = k.shape[:3]
    centre = np.abs(k[..., 0])  # (E, C, T) float32
    band = np.empty((E, C, T), dtype=np.float32)
    for e in range(E):
        for c in range(C):
            x = centre[e, c] - np.mean(centre[e, c])
            band[e, c] = signal.sosfiltfilt(high, signal.sosfiltfilt(low, x))
    per_echo = np.empty((E, T), dtype=np.float32)
    for e in range(E):
        per_echo[e] = _leading_component(band[e])
    resp = per_echo[0].copy() if E == 1 else _leading_component(per_echo)
    lo, hi = float(np.min(resp)), float(np.max(resp))
    if hi - lo <= np.finfo(np.float32).eps:
        raise ValueError("Respiratory signal is constant and cannot be normalized.")
    return np.asarray((resp - lo) / (hi - lo), dtype=np.float32)


def read_tr_seconds(input_directory) -> float:
    """The repetition time from the raw dataset's ``tr.txt`` (src/toporecon/data/dataset.py:169)."""
    return float((Path(input_directory) / "tr.txt").read_text(encoding="utf-8").strip())


def estimate_respiratory_from_dataset(input_directory, output_directory, *,
                                      config: RespiratoryConfig | None = None) -> Path:
    """Estimate and write ``resp.hdr/.cfl`` (trajectories x 1) into ``output_directory`` without
    touching the raw dataset (respiratory.py:117-133)."""
    root = Path(input_directory)
    names = ("ksp.hdr", "ksp.cfl", "ktraj.hdr", "ktraj.cfl", "dens.hdr", "dens.cfl", "imageDim.txt",
             "voxelSize.txt", "tr.txt")
    missing = [str(root / n) for n in names if not (root / n).is_file()]
    if missing:
        raise FileNotFoundError(f"Dataset is missing required files: {', '.join(missing)}")
    shape = cfl.read_cfl_header(root / "ksp")
    E, C, T, R = cfl.kspace_layout(shape)
    ksp = cfl.read_cfl(root / "ksp", mmap=True).reshape(E, C, T, R)
    resp = estimate_respiratory_signal(ksp, read_tr_seconds(root), config=config)
    out = Path(output_directory) / "resp"
    cfl.write_cfl(out, resp[:, None])
    return out
