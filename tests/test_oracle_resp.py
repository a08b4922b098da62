"""The band-pass restatement of respiratory self-gating (oracle/resp.py) against SciPy's own
``sosfiltfilt``, bit for bit, and against what the filters must do to a constant trace."""

import numpy as np
import pytest

from oracle import resp as orr

signal = pytest.importorskip("scipy.signal")


@pytest.mark.parametrize("order_lo,order_hi", [(10, 5), (6, 3), (4, 4), (1, 2)])
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_filtfilt_matches_scipy_bitwise(order_lo, order_hi, dtype):
    rng = np.random.default_rng(order_lo * 10 + order_hi)
    fs = 200.0
    x = (5.0 + rng.standard_normal(400)).astype(dtype)
    lo = signal.butter(order_lo, 2.0, "low", fs=fs, output="sos")
    hi = signal.butter(order_hi, 0.1, "high", fs=fs, output="sos")
    for sos in (lo, hi):
        got = orr.filtfilt(sos, signal.sosfilt_zi(sos), x)
        want = signal.sosfiltfilt(sos, x)
        assert want.dtype == np.float64
        assert np.array_equal(got, want), float(np.max(np.abs(got - want)))


def test_padding_follows_section_zeros():
    # an odd order ends in a first-order section (b2 = a2 = 0): one tap fewer
    assert orr.padding(signal.butter(5, 0.1, "high", fs=200.0, output="sos")) == 3 * (2 * 3 + 1 - 1)
    assert orr.padding(signal.butter(10, 2.0, "low", fs=200.0, output="sos")) == 3 * (2 * 5 + 1)


def test_bandpass_chain_matches_reference_calls():
    """float32 trace -> low-pass -> high-pass -> float32, as respiratory.py:86-92 chains them."""
    rng = np.random.default_rng(3)
    fs = 200.0
    x = (rng.standard_normal(1000) * 0.3 + 2.0).astype(np.float32)
    x = x - np.mean(x)
    lo = signal.butter(10, 2.0, "low", fs=fs, output="sos")
    hi = signal.butter(5, 0.1, "high", fs=fs, output="sos")
    want = np.empty(x.shape, np.float32)
    want[:] = signal.sosfiltfilt(hi, signal.sosfiltfilt(lo, x))
    got = orr.bandpass(x, lo, signal.sosfilt_zi(lo), hi, signal.sosfilt_zi(hi))
    assert np.array_equal(got, want)


def test_constant_trace():
    """Steady-state start: a constant passes a low-pass unchanged (to rounding) and a high-pass
    removes it; too short a trace is refused as SciPy refuses it."""
    x = np.full(300, 3.25, np.float64)
    lo = signal.butter(10, 2.0, "low", fs=200.0, output="sos")
    hi = signal.butter(5, 0.1, "high", fs=200.0, output="sos")
    assert np.max(np.abs(orr.filtfilt(lo, signal.sosfilt_zi(lo), x) - 3.25)) < 1e-9
    assert np.max(np.abs(orr.filtfilt(hi, signal.sosfilt_zi(hi), x))) < 1e-9
    with pytest.raises(ValueError):
        orr.filtfilt(lo, signal.sosfilt_zi(lo), x[:33])
