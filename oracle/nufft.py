"""Float64 restatement of SigPy 0.1.27 ``nufft`` / ``nufft_adjoint`` (test infrastructure only).

The reference calls ``sp.nufft(image, coord)`` and
``sp.nufft_adjoint(kspace, coord, oshape=img_shape)`` with SigPy's defaults
(oversamp 1.25, width 4; src/toporecon/nufft_backend.py:104-107, :112-116).
SigPy is not vendored in the reference tree; its published algorithm is
restated here:

1. forward: apodise, divide by sqrt(prod n), centred zero-pad to
   g = ceil(1.25 n), centred unnormalised FFT, Kaiser-Bessel interpolation,
   divide by width**3;
2. adjoint: Kaiser-Bessel gridding, divide by width**3, centred unnormalised
   inverse FFT (NumPy ``ifftn`` divides by prod g), centred crop, multiply by
   prod g / sqrt(prod n), apodise.

Precision mirrors what both SigPy paths share: the trajectory arrives as
float32, so the coordinate scaling ``k * (g/n) + g//2`` is done in float32,
and the Kaiser-Bessel ``beta`` is cast to the coordinate dtype.  Weights,
FFT and accumulation are float64 here (SigPy's CPU path evaluates the
weights in float64 too).

Pinned by (tests/test_oracle_nufft.py): the I0 polynomial against
``scipy.special.i0`` within Abramowitz & Stegun's stated error bounds; the
forward operator against a brute-force NUDFT on tiny grids (the width-4 /
1.25x Kaiser-Bessel approximation error bounds the gap); on-grid
coordinates against the centred FFT; exact adjointness <A x, y> = <x, A^H y>;
beta = pi*sqrt(4.96); 5 taps with end weight 1 for an integral coordinate; a
pure-Python loop restatement of SigPy's scalar gridding loops on tiny inputs.
"""

from __future__ import annotations

import math

import numpy as np

OVERSAMP = 1.25
WIDTH = 4

# Abramowitz & Stegun 9.8.1 (|x| <= 3.75, t = x/3.75) and 9.8.2 (x >= 3.75).
_AS_981 = (1.0, 3.5156229, 3.0899424, 1.2067492, 0.2659732, 0.0360768, 0.0045813)
_AS_982 = (0.39894228, 0.01328592, 0.00225319, -0.00157565, 0.00916281,
           -0.02057706, 0.02635537, -0.01647633, 0.00392377)


def kb_beta(oversamp: float = OVERSAMP, width: int = WIDTH) -> float:
    """Kaiser-Bessel shape parameter used by SigPy's nufft (float64 value)."""
    return math.pi * (((width / oversamp) * (oversamp - 0.5)) ** 2 - 0.8) ** 0.5


def oversamp_shape(shape, oversamp: float = OVERSAMP):
    return [int(math.ceil(oversamp * int(n))) for n in shape]


def i0_as(x):
    """Modified Bessel I0 by the A&S 9.8.1 / 9.8.2 polynomials (float64)."""
    x = np.asarray(x, dtype=np.float64)
    t = x / 3.75
    small = 0.0
    for k, c in enumerate(_AS_981):
        small = small + c * t ** (2 * k)
    ts = np.where(x >= 3.75, t, 1.0)
    large = 0.0
    for k, c in enumerate(_AS_982):
        large = large + c * ts ** (-k)
    xs = np.where(x >= 3.75, x, 1.0)
    large = xs ** -0.5 * np.exp(xs) * large
    return np.where(x < 3.75, small, large)


def kb_kernel(x, beta):
    """SigPy's Kaiser-Bessel kernel: I0(beta*sqrt(1-x^2)) for |x|<=1, else 0."""
    x = np.asarray(x, dtype=np.float64)
    inside = np.abs(x) <= 1.0
    arg = float(beta) * np.sqrt(np.where(inside, 1.0 - x * x, 0.0))
    return np.where(inside, i0_as(arg), 0.0)


def apodization(n: int, g: int, beta: float, width: int = WIDTH) -> np.ndarray:
    """Per-axis apodisation a(j) = s/sinh(s), s = sqrt(beta^2 - (pi W (j - n//2)/g)^2)."""
    j = np.arange(n, dtype=np.float64)
    s = np.sqrt(beta ** 2 - (np.pi * width * (j - n // 2) / g) ** 2)
    return s / np.sinh(s)


def scale_coord(coord, shape, oversamp: float = OVERSAMP) -> np.ndarray:
    """k' = k * (g/n) + g//2 per axis, evaluated in float32 like SigPy's _scale_coord."""
    c = np.asarray(coord, dtype=np.float32).copy()
    ndim = c.shape[-1]
    for i in range(-ndim, 0):
        n = int(shape[i])
        g = int(math.ceil(oversamp * n))
        c[..., i] = c[..., i] * np.float32(g / n)
        c[..., i] = c[..., i] + np.float32(g // 2)
    return c


def _axis_taps(kp: np.ndarray, g: int, beta32: float, width: int = WIDTH):
    """Grid indices (mod g) and float64 weights of the width+1 candidate taps."""
    k64 = kp.astype(np.float64)
    t0 = np.ceil(k64 - width / 2).astype(np.int64)
    t = t0[:, None] + np.arange(width + 1)[None, :]
    w = kb_kernel((t - k64[:, None]) / (width / 2), beta32)
    return np.mod(t, g), w


def _tap_tables(coord_scaled: np.ndarray, gshape, beta32: float):
    kz, ky, kx = (coord_scaled[:, i] for i in range(3))
    iz, wz = _axis_taps(kz, gshape[0], beta32)
    iy, wy = _axis_taps(ky, gshape[1], beta32)
    ix, wx = _axis_taps(kx, gshape[2], beta32)
    # (M, 5, 5, 5) linear indices and weights, z outer, x inner (SigPy loop order).
    lin = ((iz[:, :, None, None] * gshape[1] + iy[:, None, :, None]) * gshape[2]
           + ix[:, None, None, :])
    w = wz[:, :, None, None] * wy[:, None, :, None] * wx[:, None, None, :]
    return lin.reshape(len(kz), -1), w.reshape(len(kz), -1)


def interpolate(grid: np.ndarray, coord_scaled: np.ndarray, beta32: float) -> np.ndarray:
    """SigPy interp.interpolate (kaiser_bessel, width 4) on a batch of grids.

    grid: (..., gz, gy, gx); coord_scaled: (..., 3) already scaled. Returns
    (..., npts) without the final /width**3 (done by the caller, as SigPy does).
    """
    gshape = grid.shape[-3:]
    batch = grid.shape[:-3]
    pts = coord_scaled.shape[:-1]
    c = coord_scaled.reshape(-1, 3)
    flat = grid.reshape(-1, int(np.prod(gshape)))
    out = np.empty((flat.shape[0], c.shape[0]), dtype=np.result_type(grid.dtype, np.complex128))
    step = 4096
    for s in range(0, c.shape[0], step):
        lin, w = _tap_tables(c[s:s + step], gshape, beta32)
        out[:, s:s + step] = np.einsum("mt,bmt->bm", w, flat[:, lin])
    return out.reshape(tuple(batch) + tuple(pts))


def gridding(values: np.ndarray, coord_scaled: np.ndarray, gshape, beta32: float) -> np.ndarray:
    """SigPy interp.gridding (kaiser_bessel, width 4): adjoint of ``interpolate``."""
    pts = coord_scaled.shape[:-1]
    npts = int(np.prod(pts))
    batch = values.shape[: values.ndim - len(pts)]
    v = values.reshape(-1, npts)
    lin, w = _tap_tables(coord_scaled.reshape(-1, 3), gshape, beta32)
    G = int(np.prod(gshape))
    lin_f = lin.ravel()
    out = np.empty((v.shape[0], G), dtype=np.complex128)
    for b in range(v.shape[0]):
        contrib = (w * v[b][:, None]).ravel()
        out[b] = (np.bincount(lin_f, weights=contrib.real, minlength=G)
                  + 1j * np.bincount(lin_f, weights=contrib.imag, minlength=G))
    return out.reshape(tuple(batch) + tuple(gshape))


def _resize_center(x: np.ndarray, oshape) -> np.ndarray:
    """sigpy.util.resize restricted to the last three axes (centred pad or crop)."""
    ishape = x.shape[-3:]
    out = np.zeros(x.shape[:-3] + tuple(oshape), dtype=x.dtype)
    isl, osl = [], []
    for i, o in zip(ishape, oshape):
        si = max(i // 2 - o // 2, 0)
        so = max(o // 2 - i // 2, 0)
        c = min(i - si, o - so)
        isl.append(slice(si, si + c))
        osl.append(slice(so, so + c))
    out[(Ellipsis,) + tuple(osl)] = x[(Ellipsis,) + tuple(isl)]
    return out


def _apodize(x: np.ndarray, shape, gshape, beta: float) -> np.ndarray:
    out = x
    for a, (n, g) in enumerate(zip(shape, gshape)):
        ap = apodization(n, g, beta)
        sh = [1] * 3
        sh[a] = n
        out = out * ap.reshape(sh)
    return out


def nufft(image: np.ndarray, coord: np.ndarray) -> np.ndarray:
    """sp.nufft(image, coord): (..., Z, Y, X) -> (..., *coord.shape[:-1])."""
    image = np.asarray(image, dtype=np.complex128)
    shape = image.shape[-3:]
    gshape = oversamp_shape(shape)
    beta = kb_beta()
    beta32 = float(np.float32(beta))
    x = _apodize(image, shape, gshape, beta)
    x = x / math.sqrt(float(np.prod(shape)))
    x = _resize_center(x, gshape)
    axes = (-3, -2, -1)
    x = np.fft.fftshift(np.fft.fftn(np.fft.ifftshift(x, axes=axes), axes=axes), axes=axes)
    cs = scale_coord(coord, shape)
    return interpolate(x, cs, beta32) / WIDTH ** 3


def nufft_adjoint(kspace: np.ndarray, coord: np.ndarray, oshape) -> np.ndarray:
    """sp.nufft_adjoint(kspace, coord, oshape): (..., *pts) -> (..., Z, Y, X)."""
    kspace = np.asarray(kspace, dtype=np.complex128)
    shape = tuple(int(v) for v in oshape[-3:])
    gshape = oversamp_shape(shape)
    beta = kb_beta()
    beta32 = float(np.float32(beta))
    cs = scale_coord(coord, shape)
    x = gridding(kspace, cs, gshape, beta32) / WIDTH ** 3
    axes = (-3, -2, -1)
    x = np.fft.fftshift(np.fft.ifftn(np.fft.ifftshift(x, axes=axes), axes=axes), axes=axes)
    x = _resize_center(x, shape)
    x = x * (float(np.prod(gshape)) / math.sqrt(float(np.prod(shape))))
    return _apodize(x, shape, gshape, beta)


def nudft(image: np.ndarray, coord: np.ndarray) -> np.ndarray:
    """Exact centred non-uniform DFT that SigPy's nufft approximates.

    F(k) = sum_j x[j] exp(-2 pi i sum_d k_d (j_d - n_d//2) / n_d) / sqrt(prod n).
    Used only on tiny grids to bound gross errors of ``nufft``.
    """
    image = np.asarray(image, dtype=np.complex128)
    shape = image.shape[-3:]
    pos = [np.arange(n) - n // 2 for n in shape]
    c = np.asarray(coord, dtype=np.float64).reshape(-1, 3)
    ez = np.exp(-2j * np.pi * np.outer(c[:, 0], pos[0]) / shape[0])
    ey = np.exp(-2j * np.pi * np.outer(c[:, 1], pos[1]) / shape[1])
    ex = np.exp(-2j * np.pi * np.outer(c[:, 2], pos[2]) / shape[2])
    out = np.einsum("...zyx,mz,my,mx->...m", image, ez, ey, ex)
    return out.reshape(image.shape[:-3] + coord.shape[:-1]) / math.sqrt(float(np.prod(shape)))


def gridding_loops(values, coord_scaled, gshape, beta32):
    """Scalar-loop restatement of SigPy's 3-D gridding loop (tiny inputs only).

    Mirrors the loop structure of SigPy's ``_gridding3``: per sample, integer
    taps from ceil(k - W/2) to floor(k + W/2) per axis, weight product z*y*x,
    periodic wrap.  An independent formulation of ``gridding`` used to pin
    its vectorised indexing.
    """
    v = np.asarray(values, dtype=np.complex128).reshape(-1)
    c = np.asarray(coord_scaled, dtype=np.float32).reshape(-1, 3)
    gz, gy, gx = gshape
    out = np.zeros((gz, gy, gx), dtype=np.complex128)
    half = WIDTH / 2
    for i in range(len(v)):
        kz, ky, kx = (float(c[i, 0]), float(c[i, 1]), float(c[i, 2]))
        for z in range(math.ceil(kz - half), math.floor(kz + half) + 1):
            wz = float(kb_kernel((z - kz) / half, beta32))
            for y in range(math.ceil(ky - half), math.floor(ky + half) + 1):
                wy = wz * float(kb_kernel((y - ky) / half, beta32))
                for x in range(math.ceil(kx - half), math.floor(kx + half) + 1):
                    w = wy * float(kb_kernel((x - kx) / half, beta32))
                    out[z % gz, y % gy, x % gx] += w * v[i]
    return out
