"""Float64 restatement of the TVMW spatial db6 wavelet (test infrastructure only).

The reference applies ``ptwt.wavedec3(x, "db6", mode="zero", level=None,
axes=(-3,-2,-1))`` and ``ptwt.waverec3`` per motion bin on the (E, 2, Z, Y, X)
float32 view of the complex image (src/toporecon/algorithms/tvmw.py:196-206,
303-310).  ptwt is not vendored and not installed; its published algorithm
(the PyWavelets zero-mode DWT: full convolution with the decomposition filter,
keep odd-indexed outputs; level=None means ``pywt.dwtn_max_level``) is
restated here on complex arrays (real and imaginary parts see the same real
filters, so this equals the 2-channel real transform).

``waverec3`` is implemented as the exact transpose of ``wavedec3`` for the
image shape; for the orthogonal db6 filters this is PyWavelets' ``waverecn``
(upsample, convolve with the reconstruction filters, crop the approximation
to the next detail shape, crop to the image shape).

The db6 filter is derived here from Daubechies' construction (spectral
factorisation of P(y) = sum_k C(5+k, k) y^k, minimum-phase roots) rather than
copied from a table.

Pinned by (tests/test_oracle_wavelet.py): the derived filter and the
coefficients of ``wavedecn`` / ``waverecn`` against golden vectors written by
PyWavelets 1.1.1 (tests/golden/make_pywt_golden.py); orthogonality of the
filter; perfect reconstruction and adjointness of the 3-D transform.  ptwt's
own boundary handling beyond PyWavelets equality is "parity unpinned".
"""

from __future__ import annotations

import math
from functools import lru_cache

import numpy as np

FILTER_LEN = 12


@lru_cache(maxsize=None)
def daubechies_rec_lo(n_moments: int = 6) -> np.ndarray:
    """Daubechies extremal-phase scaling filter (sum = sqrt 2), length 2N."""
    N = n_moments
    # P(y) = sum_{k<N} C(N-1+k, k) y^k ; roots y_k.
    coeffs = [math.comb(N - 1 + k, k) for k in range(N)]
    yroots = np.roots(coeffs[::-1])
    zroots = []
    for y in yroots:
        # (2 - z - 1/z)/4 = y  ->  z^2 - (2 - 4y) z + 1 = 0 ; keep |z| < 1
        r = np.roots([1.0, -(2.0 - 4.0 * y), 1.0])
        zroots.append(r[np.argmin(np.abs(r))])
    poly = np.array([1.0 + 0j])
    for _ in range(N):
        poly = np.convolve(poly, [1.0, 1.0])
    for z in zroots:
        poly = np.convolve(poly, [1.0, -z])
    h = np.real(poly)
    h = h * (math.sqrt(2.0) / h.sum())
    return h


def filters():
    """(dec_lo, dec_hi) in PyWavelets orientation."""
    rec_lo = daubechies_rec_lo(6)
    dec_lo = rec_lo[::-1].copy()
    L = len(rec_lo)
    rec_hi = np.array([(-1) ** k * rec_lo[L - 1 - k] for k in range(L)])
    dec_hi = rec_hi[::-1].copy()
    return dec_lo, dec_hi


def coeff_len(n: int, L: int = FILTER_LEN) -> int:
    return (n + L - 1) // 2


def max_level(shape, L: int = FILTER_LEN) -> int:
    """pywt.dwtn_max_level: min over axes of floor(log2(n / (L - 1)))."""
    lv = []
    for n in shape:
        if n < L - 1:
            lv.append(0)
        else:
            lv.append(int(math.floor(math.log2(n / (L - 1)))))
    return min(lv)


def _dwt_axis(x, filt, axis):
    x = np.moveaxis(x, axis, -1)
    n = x.shape[-1]
    L = len(filt)
    full = np.zeros(x.shape[:-1] + (n + L - 1,), dtype=np.result_type(x, filt))
    for j in range(L):
        full[..., j:j + n] += filt[j] * x
    out = full[..., 1::2][..., :coeff_len(n, L)]
    return np.moveaxis(out, -1, axis)


def _dwt_axis_T(c, filt, axis, n):
    """Transpose of ``_dwt_axis`` for an input of length n."""
    c = np.moveaxis(c, axis, -1)
    L = len(filt)
    full = np.zeros(c.shape[:-1] + (n + L - 1,), dtype=np.result_type(c, filt))
    full[..., 1::2][..., :c.shape[-1]] = c
    out = np.zeros(c.shape[:-1] + (n,), dtype=full.dtype)
    for j in range(L):
        out += filt[j] * full[..., j:j + n]
    return np.moveaxis(out, -1, axis)


_KEYS = ["aad", "ada", "add", "daa", "dad", "dda", "ddd"]


def dwt3(x):
    dec_lo, dec_hi = filters()
    out = {}
    for kz, fz in (("a", dec_lo), ("d", dec_hi)):
        tz = _dwt_axis(x, fz, -3)
        for ky, fy in (("a", dec_lo), ("d", dec_hi)):
            ty = _dwt_axis(tz, fy, -2)
            for kx, fx in (("a", dec_lo), ("d", dec_hi)):
                out[kz + ky + kx] = _dwt_axis(ty, fx, -1)
    return out


def dwt3_T(coeffs, shape):
    dec_lo, dec_hi = filters()
    out = None
    for kz, fz in (("a", dec_lo), ("d", dec_hi)):
        for ky, fy in (("a", dec_lo), ("d", dec_hi)):
            for kx, fx in (("a", dec_lo), ("d", dec_hi)):
                c = coeffs[kz + ky + kx]
                t = _dwt_axis_T(c, fx, -1, shape[2])
                t = _dwt_axis_T(t, fy, -2, shape[1])
                t = _dwt_axis_T(t, fz, -3, shape[0])
                out = t if out is None else out + t
    return out


def level_shapes(shape, level):
    shapes = [tuple(shape)]
    for _ in range(level):
        shapes.append(tuple(coeff_len(n) for n in shapes[-1]))
    return shapes


def wavedec3(x, level=None):
    """[cA_n, details_n, ..., details_1]; details are dicts keyed 'aad'..'ddd'."""
    x = np.asarray(x)
    if level is None:
        level = max_level(x.shape[-3:])
    details = []
    a = x
    for _ in range(level):
        c = dwt3(a)
        a = c.pop("aaa")
        details.append(c)
    return [a] + details[::-1]


def waverec3(coeffs, img_shape):
    """Transpose of ``wavedec3`` for an image of spatial shape img_shape."""
    level = len(coeffs) - 1
    shapes = level_shapes(img_shape, level)
    a = coeffs[0]
    for i in range(level):
        det = dict(coeffs[1 + i])
        det["aaa"] = a
        target = shapes[level - 1 - i]
        a = dwt3_T(det, target)
    return a


def tree_leaves(coeffs):
    leaves = [coeffs[0]]
    for d in coeffs[1:]:
        leaves.extend(d[k] for k in _KEYS)
    return leaves


def zeros_like_coeffs(coeffs):
    return [np.zeros_like(coeffs[0])] + [{k: np.zeros_like(v) for k, v in d.items()} for d in coeffs[1:]]


def tree_axpy(dst, a, src):
    return [dst[0] + a * src[0]] + [{k: dst[i][k] + a * src[i][k] for k in dst[i]}
                                   for i in range(1, len(dst))]


def tree_project_linf(coeffs, lam):
    """t * min(1, lam / sqrt(re^2 + im^2 + 1e-12)); zero if lam <= 0 (tvmw.py:162-193)."""
    def proj(t):
        if lam <= 0:
            return np.zeros_like(t)
        mag = np.sqrt(t.real ** 2 + t.imag ** 2 + 1e-12)
        return t * np.minimum(1.0, lam / mag)
    return [proj(coeffs[0])] + [{k: proj(v) for k, v in d.items()} for d in coeffs[1:]]
