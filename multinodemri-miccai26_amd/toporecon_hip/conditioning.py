"""Host-side input conditioning that defines every hot-path input.

Restates the single-process steps of the reference ``main`` functions:

* readout fraction, 2 %/98 % respiratory outlier drop and ``int(Ntr/acc)``
  truncation (TVM/TVME only): tvm.py:747-790, tvme.py:725-780;
* DCF normalisation by its maximum (TVM: again after bin selection,
  tvm.py:803-805, :956-959; TVME: once, tvme.py:806-809; TVMW: after bin
  selection only, tvmw.py:1130-1133);
* percentile bin edges over the kept respiratory trace and membership
  ``edge_b <= r < edge_{b+1}`` (tvme.py:856-881, tvm.py:159-168);
* ``ksp /= max|ksp|`` and ``mps /= RSS`` (TVME floors the RSS at 1e-12,
  tvme.py:938-949; TVM/TVMW replace a zero RSS by 1, tvm.py:945-950).
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np


def drop_respiratory_outliers(resp: np.ndarray, acceleration: int = 1, margin: float = 2):
    """Indices of kept readouts: p_margin < r < p_(100-margin), then first int(N/acc)."""
    resp = np.asarray(resp).reshape(-1)
    n_total = int(resp.size)
    keep = np.arange(n_total, dtype=np.int64)
    lo = np.percentile(resp, 0 + margin)
    hi = np.percentile(resp, 100 - margin)
    keep = keep[(resp > lo) & (resp < hi)]
    n_target = int(n_total / max(int(acceleration), 1))
    return keep[:n_target]


def readout_count(frac: float, n_readout: int) -> int:
    return int(frac * n_readout)


def bin_edges(resp: np.ndarray, num_bins: int, margin: float = 0.0) -> np.ndarray:
    return np.percentile(resp, np.linspace(0 + margin, 100 - margin, num_bins + 1))


def bin_members(resp: np.ndarray, edges: np.ndarray, b: int) -> np.ndarray:
    """Boolean mask of readouts in bin b (strict upper edge, so max(resp) is in no bin)."""
    return (resp >= edges[b]) & (resp < edges[b + 1])


def normalize_dcf(dcf: np.ndarray) -> np.ndarray:
    m = float(np.max(dcf)) if dcf.size else 0.0
    if m > 0:
        dcf = dcf / np.float32(m)
    return dcf.astype(np.float32, copy=False)


def normalize_kspace(ksp: np.ndarray) -> np.ndarray:
    m = float(np.max(np.abs(ksp))) if ksp.size else 0.0
    if m > 0:
        ksp = ksp / np.float32(m)
    return ksp.astype(np.complex64, copy=False)


def normalize_maps(mps: np.ndarray, algorithm: str = "tvme") -> np.ndarray:
    sos = np.sqrt(np.sum(np.abs(mps) ** 2, axis=0, dtype=np.float32), dtype=np.float32)
    if algorithm == "tvme":
        sos = np.maximum(sos, np.float32(1e-12))
    else:
        sos = np.where(sos == 0, np.float32(1.0), sos)
    return (mps / sos[None]).astype(np.complex64, copy=False)


@dataclass
class Conditioned:
    """Conditioned inputs of one solver run (all readouts of all local bins)."""

    ksp: np.ndarray      # (E, C, Ntr, Nro) complex64
    coord: np.ndarray    # (Ntr, Nro, 3) float32, [z, y, x] SigPy units
    dcf: np.ndarray      # (Ntr, Nro) float32
    resp: np.ndarray     # (Ntr,)
    mps: np.ndarray      # (C, Z, Y, X) complex64
    bin_edges: np.ndarray
