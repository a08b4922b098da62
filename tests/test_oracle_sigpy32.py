"""The single-precision CPU baseline path (oracle/sigpy32.py) against the float64 oracle.

The float64 oracle is pinned independently (brute-force NUDFT, centred FFT on grid
coordinates, exact adjointness: tests/test_oracle_nufft.py); here the complex64 restatement
timed as the CPU baseline must agree with it to single precision, including samples on grid
lines (5 taps per axis) and samples whose taps wrap around the grid edge.
"""

import numpy as np

from oracle import nufft as on
from oracle import sigpy32 as s32


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def _coords(rng, shape, m):
    c = rng.uniform(-0.5, 0.5, (m, 3)) * np.array(shape)
    c[:40] = np.round(c[:40])                    # on-grid on every axis: 5 taps each
    c[40:80, 1] = np.round(c[40:80, 1])          # on-grid on one axis only
    c[80:90] = np.array(shape) / 2.0 - 0.25       # taps wrap past the grid edge
    return c.astype(np.float32)


def test_forward_matches_float64_oracle():
    rng = np.random.default_rng(11)
    shape = (11, 14, 16)
    img = (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(np.complex64)
    coord = _coords(rng, shape, 3000)
    ref = on.nufft(img.astype(np.complex128), coord)
    out = s32.nufft(img, coord)
    assert out.dtype == np.complex64
    assert _rel(out, ref) < 2e-6


def test_adjoint_matches_float64_oracle_and_is_adjoint():
    rng = np.random.default_rng(12)
    shape = (9, 12, 10)
    coord = _coords(rng, shape, 2500)
    k = (rng.standard_normal(2500) + 1j * rng.standard_normal(2500)).astype(np.complex64)
    ref = on.nufft_adjoint(k.astype(np.complex128), coord, shape)
    out = s32.nufft_adjoint(k, coord, shape)
    assert out.dtype == np.complex64
    assert _rel(out, ref) < 2e-6
    x = (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(np.complex64)
    lhs = np.vdot(s32.nufft(x, coord).astype(np.complex128), k.astype(np.complex128))
    rhs = np.vdot(x.astype(np.complex128), out.astype(np.complex128))
    assert abs(lhs - rhs) / abs(lhs) < 1e-5


def test_data_channel_matches_float64_step():
    rng = np.random.default_rng(13)
    shape = (8, 10, 12)
    coord = _coords(rng, shape, 1500)
    u = (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(np.complex64)
    s = (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(np.complex64)
    y = (rng.standard_normal(1500) + 1j * rng.standard_normal(1500)).astype(np.complex64)
    q = (rng.standard_normal(1500) + 1j * rng.standard_normal(1500)).astype(np.complex64)
    dcf = rng.uniform(0.1, 1.0, 1500).astype(np.float32)
    g = np.zeros(shape, dtype=np.complex64)
    sig = 1.0 / 6.0
    q_ref = (q.astype(np.complex128) + sig * (on.nufft(u.astype(np.complex128) * s, coord) - y)) / (1 + sig)
    g_ref = np.conj(s.astype(np.complex128)) * on.nufft_adjoint(dcf * q_ref, coord, shape)
    s32.data_channel(u, s, y, q, dcf, coord, g, sig)
    assert _rel(q, q_ref) < 2e-6
    assert _rel(g, g_ref) < 2e-6
