"""Pins the float64 SigPy NUFFT restatement (oracle/nufft.py) independently of its own formula."""

import math

import numpy as np
import pytest
import scipy.special

from oracle import nufft as on


def test_i0_within_abramowitz_stegun_bounds():
    # A&S 9.8.1: |eps| < 1.6e-7 (relative, |x| <= 3.75); 9.8.2: |eps| < 1.9e-7 on x^0.5 e^-x I0.
    x = np.linspace(0.0, 3.75, 2001)
    rel = np.abs(on.i0_as(x) / scipy.special.i0(x) - 1.0)
    assert rel.max() < 1.6e-7
    x = np.linspace(3.75, 8.0, 2001)
    scaled_err = np.abs(on.i0_as(x) - scipy.special.i0(x)) * np.sqrt(x) * np.exp(-x)
    assert scaled_err.max() < 1.9e-7 * 1.01


def test_beta_closed_form():
    assert on.kb_beta() == pytest.approx(math.pi * math.sqrt(4.96), rel=1e-15)
    assert on.kb_beta() == pytest.approx(6.996659, abs=1e-6)


def test_integral_coordinate_has_five_taps_with_unit_end_weights():
    beta32 = float(np.float32(on.kb_beta()))
    idx, w = on._axis_taps(np.array([10.0], dtype=np.float32), 40, beta32)
    assert np.count_nonzero(w) == 5
    assert w[0, 0] == pytest.approx(1.0) and w[0, 4] == pytest.approx(1.0)
    assert list(idx[0]) == [8, 9, 10, 11, 12]
    idx, w = on._axis_taps(np.array([10.3], dtype=np.float32), 40, beta32)
    assert np.count_nonzero(w) == 4


def test_oversampled_shape_and_coordinate_scaling():
    assert on.oversamp_shape((80, 192, 192)) == [100, 240, 240]
    assert on.oversamp_shape((48, 64, 64)) == [60, 80, 80]
    assert on.oversamp_shape((13, 15, 7)) == [17, 19, 9]
    c = np.array([[-24.0, 0.0, 32.0]], dtype=np.float32)
    s = on.scale_coord(c, (48, 64, 64))
    assert s.dtype == np.float32
    assert list(s[0]) == [0.0, 40.0, 80.0]


def _rand_problem(shape, M, seed):
    rng = np.random.default_rng(seed)
    img = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    half = np.array(shape) / 2.0
    coord = (rng.uniform(-1, 1, (M, 3)) * half).astype(np.float32)
    return img, coord


@pytest.mark.parametrize("shape", [(8, 8, 8), (6, 10, 7), (9, 8, 12)])
def test_forward_matches_brute_force_nudft(shape):
    # Width-4 / 1.25x Kaiser-Bessel NUFFT approximates the exact centred NUDFT to ~1e-2;
    # a sign, shift, scale or axis-order slip gives O(1) errors.
    img, coord = _rand_problem(shape, 300, 1)
    a = on.nufft(img, coord)
    b = on.nudft(img, coord.astype(np.float64))
    rel = np.linalg.norm(a - b) / np.linalg.norm(b)
    assert rel < 2e-2, rel


def test_forward_on_grid_matches_centred_fft():
    shape = (8, 10, 12)
    rng = np.random.default_rng(3)
    img = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    kz, ky, kx = np.meshgrid(*[np.arange(-(n // 2), n - n // 2) for n in shape], indexing="ij")
    coord = np.stack([kz, ky, kx], axis=-1).reshape(-1, 3).astype(np.float32)
    F = np.fft.fftshift(np.fft.fftn(np.fft.ifftshift(img))) / math.sqrt(img.size)
    a = on.nufft(img, coord).reshape(shape)
    rel = np.linalg.norm(a - F) / np.linalg.norm(F)
    assert rel < 2e-2, rel


def test_adjointness_is_exact():
    shape = (7, 10, 9)
    img, coord = _rand_problem(shape, 500, 5)
    rng = np.random.default_rng(6)
    y = rng.standard_normal(500) + 1j * rng.standard_normal(500)
    lhs = np.vdot(on.nufft(img, coord), y)
    rhs = np.vdot(img, on.nufft_adjoint(y, coord, shape))
    assert abs(lhs - rhs) / abs(lhs) < 1e-12


def test_vectorised_gridding_matches_scalar_loops():
    shape = (6, 8, 10)
    gshape = on.oversamp_shape(shape)
    _, coord = _rand_problem(shape, 40, 9)
    coord[0] = [0.0, 0.0, 0.0]            # integral position (5 taps per axis)
    coord[1] = [-3.0, -4.0, -5.0]         # edge, wraps around
    cs = on.scale_coord(coord, shape)
    rng = np.random.default_rng(10)
    v = rng.standard_normal(40) + 1j * rng.standard_normal(40)
    beta32 = float(np.float32(on.kb_beta()))
    a = on.gridding(v, cs, gshape, beta32)
    b = on.gridding_loops(v, cs, gshape, beta32)
    assert np.abs(a - b).max() < 1e-12 * np.abs(b).max()


def test_delta_image_flat_spectrum():
    # A centred delta has |F(k)| = 1/sqrt(N) for every k; the NUFFT reproduces it up to KB ripple.
    shape = (8, 8, 8)
    img = np.zeros(shape, complex)
    img[4, 4, 4] = 1.0
    _, coord = _rand_problem(shape, 200, 11)
    a = np.abs(on.nufft(img, coord)) * math.sqrt(img.size)
    assert np.abs(a - 1.0).max() < 2e-2
    assert np.sqrt(np.mean((a - 1.0) ** 2)) < 5e-3
