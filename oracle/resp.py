"""Scalar restatement of the band-pass of respiratory self-gating (test infrastructure only).

The reference filters every (echo, coil) centre trace with SciPy 1.15's ``sosfiltfilt``, low-pass
then high-pass (src/toporecon/preprocessing/respiratory.py:86-92).  This module restates that call
one operation at a time, in the order the HIP kernel ``k_resp_bandpass`` evaluates it:

* padding: 3 * ntaps samples of odd extension at each end, ntaps = 2 * sections + 1 less the
  trailing zero coefficients of a first-order section; the extension 2 x[0] - x[pad .. 1] and
  2 x[-1] - x[-2 .. -1-pad] is computed in the input's precision (float32 for the first filter);
* a forward pass of the second-order sections from the state zi * ext[0], then a backward pass
  from zi * y[-1], in float64: per sample and section, x_new = b0 x + z0,
  z0 = (b1 x - a1 x_new) + z1, z1 = b2 x - a2 x_new, each product and sum rounded;
* the middle T samples are the result.

Pinned by tests/test_oracle_resp.py: equal, bit for bit, to ``scipy.signal.sosfiltfilt`` on
float32 and float64 traces for even- and odd-order Butterworth filters (the odd orders end in a
first-order section, which shortens the padding), and a constant trace passes a low-pass unchanged
while a high-pass maps it to zero.
"""

from __future__ import annotations

import numpy as np


def padding(sos: np.ndarray) -> int:
    sos = np.asarray(sos, dtype=np.float64)
    n = sos.shape[0]
    return 3 * (2 * n + 1 - min(int((sos[:, 2] == 0).sum()), int((sos[:, 5] == 0).sum())))


def _pass(sos: np.ndarray, zi: np.ndarray, buf: list, backward: bool) -> None:
    x0 = buf[-1] if backward else buf[0]
    z = [[float(zi[s, 0]) * x0, float(zi[s, 1]) * x0] for s in range(sos.shape[0])]
    order = range(len(buf) - 1, -1, -1) if backward else range(len(buf))
    for n in order:
        xc = buf[n]
        for s in range(sos.shape[0]):
            b0, b1, b2, _, a1, a2 = (float(v) for v in sos[s])
            xn = b0 * xc + z[s][0]
            z[s][0] = (b1 * xc - a1 * xn) + z[s][1]
            z[s][1] = b2 * xc - a2 * xn
            xc = xn
        buf[n] = xc


def filtfilt(sos: np.ndarray, zi: np.ndarray, x: np.ndarray) -> np.ndarray:
    """One zero-phase pass of the cascade over a 1-D trace (float32 or float64); float64 out."""
    sos = np.asarray(sos, dtype=np.float64)
    e = padding(sos)
    T = x.shape[0]
    if T <= e:
        raise ValueError(f"The length of the input vector x must be greater than padlen, which is {e}.")
    two = x.dtype.type(2)
    left = [float(two * x[0] - x[e - j]) for j in range(e)]
    right = [float(two * x[T - 1] - x[T - 2 - j]) for j in range(e)]
    buf = left + [float(v) for v in x] + right
    _pass(sos, zi, buf, backward=False)
    _pass(sos, zi, buf, backward=True)
    return np.array(buf[e:e + T], dtype=np.float64)


def bandpass(trace32: np.ndarray, low, low_zi, high, high_zi) -> np.ndarray:
    """float32 trace -> float32 band-passed trace (low-pass, then high-pass)."""
    return filtfilt(high, high_zi, filtfilt(low, low_zi, np.asarray(trace32, dtype=np.float32))).astype(np.float32)
