"""Respiratory self-gating: the respiratory trace ``resp`` that the solvers bin by.

Restates ``toporecon.preprocessing.respiratory`` (src/toporecon/preprocessing/respiratory.py:15-133):
the magnitude of each trajectory's first readout sample (the k-space centre of a centre-out
trajectory) per echo and coil, mean removed, band-passed forward and backward with Butterworth
second-order sections (low-pass order 10 at 2 Hz, then high-pass order 5 at 0.1 Hz), the leading
principal component over coils per echo (singular value x right singular vector), then over
echoes, normalised to [0, 1].  The sign of the principal component is the one LAPACK's SVD returns,
as in the reference.

With ``device`` set, the band-pass of all E x C traces runs on the GPU (``tr_resp_bandpass``, one
thread per trace, float64 in SciPy's operation order, so the filtered traces carry the same bits);
the centre magnitudes, their means, the filter design and the two small SVDs stay on the host,
where the reference computes them.  Without it everything runs in NumPy / SciPy as in the reference.
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


def _filter_setup(sos):
    """(sos, zi, sections, padding) of one SciPy ``sosfiltfilt`` call with its default odd padding:
    3 * ntaps samples, ntaps = 2 * sections + 1 less the trailing zero coefficients of a first-order
    section."""
    from scipy import signal
    sos = np.ascontiguousarray(sos, dtype=np.float64)
    n = sos.shape[0]
    ntaps = 2 * n + 1 - min(int((sos[:, 2] == 0).sum()), int((sos[:, 5] == 0).sum()))
    zi = np.ascontiguousarray(signal.sosfilt_zi(sos), dtype=np.float64)
    return sos, zi, n, 3 * ntaps


def bandpass_hip(traces: np.ndarray, low_sos, high_sos, device=0) -> np.ndarray:
    """Zero-phase low-pass then high-pass (SciPy's ``sosfiltfilt`` twice) of float32 traces
    (..., T) on the GPU; returns float32 (..., T), the bits of the host path."""
    import torch
    from . import _lib
    x = np.ascontiguousarray(traces, dtype=np.float32)
    T = x.shape[-1]
    lo, hi = _filter_setup(low_sos), _filter_setup(high_sos)
    for _, _, _, pad in (lo, hi):
        if T <= pad:
            raise ValueError(f"The length of the input vector x must be greater than padlen, which is {pad}.")
    dev = torch.device("cuda", device) if not isinstance(device, torch.device) else device
    with torch.cuda.device(dev):
        xt = torch.from_numpy(x.reshape(-1, T)).to(dev)
        n = xt.shape[0]
        lo_t = [torch.from_numpy(a).to(dev) for a in lo[:2]]
        hi_t = [torch.from_numpy(a).to(dev) for a in hi[:2]]
        work = torch.empty(int(_lib.load().tr_resp_work_count(n, T, lo[3], hi[3])), dtype=torch.float64,
                           device=dev)
        out = torch.empty((n, T), dtype=torch.float32, device=dev)
        _lib.call("tr_resp_bandpass", xt.data_ptr(), n, T, lo_t[0].data_ptr(), lo_t[1].data_ptr(), lo[2], lo[3],
                  hi_t[0].data_ptr(), hi_t[1].data_ptr(), hi[2], hi[3], work.data_ptr(), out.data_ptr(),
                  _lib.stream_handle(dev))
        return out.cpu().numpy().reshape(x.shape)


def estimate_respiratory_signal(kspace, tr_seconds: float, *, config: RespiratoryConfig | None = None,
                                device=None) -> np.ndarray:
    """Normalised respiratory trace (float32, one value per trajectory) from canonical k-space
    ``[echo, coil, trajectory, readout]`` (respiratory.py:36-114).  device: a GPU index for the
    band-pass (None: SciPy on the host, as the reference)."""
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
    if device is not None:
        for e in range(E):
            for c in range(C):
                band[e, c] = centre[e, c] - np.mean(centre[e, c])
        band = bandpass_hip(band, low, high, device)
    else:
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
                                      config: RespiratoryConfig | None = None, device=None) -> Path:
    """Estimate and write ``resp.hdr/.cfl`` (trajectories x 1) into ``output_directory`` without
    touching the raw dataset (respiratory.py:117-133); device as in estimate_respiratory_signal."""
    root = Path(input_directory)
    names = ("ksp.hdr", "ksp.cfl", "ktraj.hdr", "ktraj.cfl", "dens.hdr", "dens.cfl", "imageDim.txt",
             "voxelSize.txt", "tr.txt")
    missing = [str(root / n) for n in names if not (root / n).is_file()]
    if missing:
        raise FileNotFoundError(f"Dataset is missing required files: {', '.join(missing)}")
    shape = cfl.read_cfl_header(root / "ksp")
    E, C, T, R = cfl.kspace_layout(shape)
    ksp = cfl.read_cfl(root / "ksp", mmap=True).reshape(E, C, T, R)
    resp = estimate_respiratory_signal(ksp, read_tr_seconds(root), config=config, device=device)
    out = Path(output_directory) / "resp"
    cfl.write_cfl(out, resp[:, None])
    return out
