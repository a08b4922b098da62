"""Float64 restatement of JSENSE coil-sensitivity estimation (test infrastructure only).

The reference estimates coil maps with SigPy 0.1.27's ``sigpy.mri.app.JsenseRecon`` on the
first echo (src/toporecon/preprocessing/jsense.py:33-97, call at :85), passing the trajectory,
the density as weights and no other options, so SigPy's defaults apply: map-kernel width 16,
calibration width 24, lambda 0, 10 alternating iterations of 10 conjugate-gradient steps each.
SigPy is not vendored in the reference and not installed here; its published algorithm is
restated:

1. ``img_shape`` = ``estimate_shape(coord)`` = int(max - min) per axis of the full trajectory;
   the calibration samples are those with max_axis |k| < 24 / 2; y keeps them (all coils),
   y <- sqrt(w) y / max|y|; the calibration grid is ``estimate_shape`` of their coordinates.
2. Model, per coil c: y_c = sqrt(w) NUFFT(IFFT(mps_ker_c (*) img_ker)), where (*) is the
   'valid' N-D convolution of the image kernel (calibration grid + 15 per axis, a centred
   Dirac at the start) with the 16^3 map kernel (zeros at the start), IFFT is SigPy's centred
   orthonormal inverse FFT on the calibration grid and NUFFT is sp.nufft on that grid.
3. Alternating minimisation: 10 x (10 CG steps of the normal equations in mps_ker, warm
   started; then 10 in img_ker); SigPy's ConjugateGradient (x0 = current value, real inner
   products, the residual not updated on the last step).
4. Maps: mps_c = IFFT(resize(mps_ker_c, img_shape)), divided by their root sum of squares;
   the reference then resizes each map through k-space to the requested image shape,
   IFFT(resize(FFT(mps_c), shape)).

Pinned by (tests/test_oracle_jsense.py): the 'valid' convolution against a scalar loop and
SciPy's convolve; adjoint identities of both convolution adjoints, of the centred (I)FFT and of
the whole forward operator; the objective never increases over the alternating half-steps; on
noiseless data generated inside the model the objective falls by orders of magnitude and the
recovered maps match the generating ones up to a global phase.  SigPy's own floating-point
behaviour is "parity unpinned" (SigPy is not installed; nothing was denied).
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from . import nufft as onufft

MPS_KER_WIDTH = 16
KSP_CALIB_WIDTH = 24
MAX_ITER = 10
MAX_INNER_ITER = 10


def estimate_shape(coord):
    """sp.estimate_shape: int(max - min) of the coordinates per axis."""
    c = np.asarray(coord)
    return [int(c[..., i].max() - c[..., i].min()) for i in range(c.shape[-1])]


def fft_c(x, axes=(-3, -2, -1)):
    """SigPy's centred orthonormal FFT."""
    return np.fft.fftshift(np.fft.fftn(np.fft.ifftshift(x, axes=axes), axes=axes, norm="ortho"), axes=axes)


def ifft_c(x, axes=(-3, -2, -1)):
    """SigPy's centred orthonormal inverse FFT."""
    return np.fft.fftshift(np.fft.ifftn(np.fft.ifftshift(x, axes=axes), axes=axes, norm="ortho"), axes=axes)


def resize(x, oshape):
    """sp.resize on the last three axes (centred crop / zero pad)."""
    return onufft._resize_center(np.asarray(x), oshape)


def conv_valid(img, ker):
    """'valid' N-D convolution (signal-processing definition, the kernel flipped):
    out[n] = sum_k ker[k] img[n + W - 1 - k] over the three last axes."""
    W = ker.shape[-3:]
    I = img.shape[-3:]
    G = tuple(i - w + 1 for i, w in zip(I, W))
    out = np.zeros(ker.shape[:-3] + G, dtype=np.result_type(img, ker))
    for kz in range(W[0]):
        for ky in range(W[1]):
            for kx in range(W[2]):
                sl = img[..., W[0] - 1 - kz:W[0] - 1 - kz + G[0], W[1] - 1 - ky:W[1] - 1 - ky + G[1],
                         W[2] - 1 - kx:W[2] - 1 - kx + G[2]]
                out += ker[..., kz, ky, kx, None, None, None] * sl
    return out


def conv_valid_adj_img(out, ker, img_shape):
    """Adjoint of img -> sum over coils of conv_valid(img, ker_c)."""
    W = ker.shape[-3:]
    G = out.shape[-3:]
    adj = np.zeros(tuple(img_shape), dtype=np.result_type(out, ker))
    for kz in range(W[0]):
        for ky in range(W[1]):
            for kx in range(W[2]):
                contrib = np.sum(np.conj(ker[:, kz, ky, kx])[:, None, None, None] * out, axis=0)
                adj[W[0] - 1 - kz:W[0] - 1 - kz + G[0], W[1] - 1 - ky:W[1] - 1 - ky + G[1],
                    W[2] - 1 - kx:W[2] - 1 - kx + G[2]] += contrib
    return adj


def conv_valid_adj_ker(out, img, ker_shape):
    """Adjoint of ker -> conv_valid(img, ker) (per coil)."""
    W = ker_shape[-3:]
    G = out.shape[-3:]
    adj = np.zeros(tuple(ker_shape), dtype=np.result_type(out, img))
    for kz in range(W[0]):
        for ky in range(W[1]):
            for kx in range(W[2]):
                sl = img[W[0] - 1 - kz:W[0] - 1 - kz + G[0], W[1] - 1 - ky:W[1] - 1 - ky + G[1],
                         W[2] - 1 - kx:W[2] - 1 - kx + G[2]]
                adj[:, kz, ky, kx] = np.sum(np.conj(sl)[None] * out, axis=(-3, -2, -1))
    return adj


@dataclass
class JsenseProblem:
    y: np.ndarray        # (C, Mcal) weighted, normalised calibration data
    coord: np.ndarray    # (Mcal, 3) calibration coordinates
    sqrt_w: np.ndarray   # (Mcal,)
    img_shape: list      # estimate_shape of the full trajectory
    grd_shape: list      # estimate_shape of the calibration coordinates
    ker_width: int


def calibration(y, coord, weights, mps_ker_width=MPS_KER_WIDTH, ksp_calib_width=KSP_CALIB_WIDTH):
    """JsenseRecon._get_data: y (C, ...) and coord (..., 3) of the first echo."""
    coord = np.asarray(coord)
    img_shape = estimate_shape(coord)
    sel = np.amax(np.abs(coord), axis=-1) < ksp_calib_width / 2
    c = coord[sel]
    yc = np.asarray(y)[:, sel]
    w = np.asarray(weights)[sel]
    yc = (w ** 0.5) * yc / np.abs(yc).max()
    return JsenseProblem(yc, c, w ** 0.5, img_shape, estimate_shape(c), mps_ker_width)


def forward(prob, img_ker, mps_ker):
    """sqrt(w) NUFFT(IFFT(mps_ker_c (*) img_ker)) for every coil -> (C, Mcal)."""
    z = conv_valid(img_ker[None], mps_ker)
    return onufft.nufft(ifft_c(z), prob.coord) * prob.sqrt_w[None]


def adjoint_img(prob, r, mps_ker, img_shape):
    z = fft_c(onufft.nufft_adjoint(r * prob.sqrt_w[None], prob.coord, prob.grd_shape))
    return conv_valid_adj_img(z, mps_ker, img_shape)


def adjoint_ker(prob, r, img_ker, ker_shape):
    z = fft_c(onufft.nufft_adjoint(r * prob.sqrt_w[None], prob.coord, prob.grd_shape))
    return conv_valid_adj_ker(z, img_ker, ker_shape)


def objective(prob, img_ker, mps_ker):
    r = forward(prob, img_ker, mps_ker) - prob.y
    return float(np.vdot(r, r).real)


def cg(normal, b, x, max_iter, lamda=0.0):
    """SigPy ConjugateGradient on (A^H A + lamda I) x = b from x."""
    x = x.copy()
    r = b - (normal(x) + lamda * x)
    p = r.copy()
    rzold = float(np.vdot(r, r).real)
    for it in range(max_iter):
        if rzold ** 0.5 <= 0:
            break
        ap = normal(p) + lamda * p
        pap = float(np.vdot(p, ap).real)
        if pap <= 0:
            break
        alpha = rzold / pap
        x += alpha * p
        if it < max_iter - 1:
            r -= alpha * ap
            rznew = float(np.vdot(r, r).real)
            p = r + (rznew / rzold) * p
            rzold = rznew
    return x


def init_vars(prob, num_coils, dtype=np.complex128):
    W = prob.ker_width
    img_ker_shape = [g + W - 1 for g in prob.grd_shape]
    img_ker = np.zeros(img_ker_shape, dtype)
    img_ker[tuple(s // 2 for s in img_ker_shape)] = 1.0  # sp.dirac
    mps_ker = np.zeros((num_coils, W, W, W), dtype)
    return img_ker, mps_ker


def jsense(prob, max_iter=MAX_ITER, max_inner_iter=MAX_INNER_ITER, lamda=0.0, history=None):
    """Alternating minimisation; returns (img_ker, mps_ker)."""
    C = prob.y.shape[0]
    img_ker, mps_ker = init_vars(prob, C)
    for _ in range(max_iter):
        b = adjoint_ker(prob, prob.y, img_ker, mps_ker.shape)
        mps_ker = cg(lambda v: adjoint_ker(prob, forward(prob, img_ker, v), img_ker, mps_ker.shape), b, mps_ker,
                     max_inner_iter, lamda)
        if history is not None:
            history.append(objective(prob, img_ker, mps_ker))
        b = adjoint_img(prob, prob.y, mps_ker, img_ker.shape)
        img_ker = cg(lambda v: adjoint_img(prob, forward(prob, v, mps_ker), mps_ker, img_ker.shape), b, img_ker,
                     max_inner_iter, lamda)
        if history is not None:
            history.append(objective(prob, img_ker, mps_ker))
    return img_ker, mps_ker


def maps_from_kernel(mps_ker, img_shape):
    """JsenseRecon._output: IFFT of the kernel resized to img_shape, divided by the RSS."""
    mps = np.stack([ifft_c(resize(k, img_shape)) for k in mps_ker])
    return mps / np.sqrt(np.sum(np.abs(mps) ** 2, axis=0))


def to_shape(mps, shape):
    """The reference's final resize of each map through k-space (jsense.py:93-96)."""
    return np.stack([ifft_c(resize(fft_c(m), shape)) for m in mps])


def estimate_sensitivity_maps(kspace, coordinates_zyx, density, image_shape_zyx, **kw):
    """estimate_sensitivity_maps (jsense.py:33-97) in float64: kspace (E, C, Ntr, Nro)."""
    prob = calibration(np.asarray(kspace)[0], coordinates_zyx, density)
    _, mps_ker = jsense(prob, **kw)
    return to_shape(maps_from_kernel(mps_ker, prob.img_shape), tuple(image_shape_zyx))
