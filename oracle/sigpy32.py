"""Single-precision NumPy restatement of the reference's CPU data path (test infrastructure only).

This is the CPU baseline timed beside the HIP path: the arithmetic the
reference runs with ``--device -1`` (SigPy 0.1.27 on NumPy), in the dtypes
that path uses (complex64 images, k-space and grids, float32 trajectory and
density), on one thread.

* ``nufft`` / ``nufft_adjoint`` follow SigPy's ``sp.nufft`` /
  ``sp.nufft_adjoint`` as called at src/toporecon/nufft_backend.py:104-107 and
  :112-116 (apodise, scale by 1/sqrt(prod n), centred pad to ceil(1.25 n),
  ifftshift / fftn / fftshift, Kaiser-Bessel width-4 interpolation over the
  taps ceil(k - 2) .. floor(k + 2), divide by 4^3; the adjoint in reverse).
  SigPy's NumPy path runs ``numpy.fft`` (single-threaded pocketfft) and
  numba loops without ``parallel``; here the tap loops are vectorised over
  samples (64 taps per sample, plus the 5th tap of samples on a grid line).
* ``data_channel`` is the per-(bin, coil, echo) unit of the data term of
  TvmeReconstructor.pdhg: the forward product and dual prox of
  src/toporecon/algorithms/tvme.py:257-267 and the weighted adjoint and coil
  accumulation of tvme.py:383-394.
* ``tvme_elementwise`` is the TV part of one TVME iteration
  (tvme.py:218-337 duals and prox, tvme.py:421-480 divergences, primal step
  and extrapolation), via the float64 oracle's step functions applied to
  complex64 arrays (NumPy keeps complex64 under Python-float scalars).

Pinned by (tests/test_oracle_sigpy32.py): agreement with the float64 oracle
(``oracle.nufft``), which is itself pinned by a brute-force NUDFT, the centred
FFT on grid coordinates and exact adjointness, to single-precision accuracy on
random and on-grid coordinates (the 5-tap case), and the adjoint identity
<A x, y> = <x, A^H y> to single-precision accuracy.
"""

from __future__ import annotations

import math

import numpy as np

from . import nufft as onufft
from . import pdhg as opdhg

WIDTH = 4
_TAP_CHUNK = 1 << 17  # samples per vectorised block (bounds the temporaries)


def _apod32(shape, gshape, beta):
    """Per-axis apodisation vectors in float32 (SigPy's _apodize)."""
    return [onufft.apodization(n, g, beta).astype(np.float32) for n, g in zip(shape, gshape)]


def _apodize32(x, vecs):
    z, y, xx = vecs
    return x * z[:, None, None] * y[None, :, None] * xx[None, None, :]


def _axis_taps32(kp, g, beta32):
    """First tap ceil(k - 2), float32 weights of the 5 candidate taps (zero past floor(k + 2)),
    grid offsets (t mod g) of the 5 candidates, and whether the 5th tap exists."""
    k64 = kp.astype(np.float64)
    t0 = np.ceil(k64 - WIDTH / 2).astype(np.int64)
    t = t0[:, None] + np.arange(WIDTH + 1)[None, :]
    w = onufft.kb_kernel((t - k64[:, None]) / (WIDTH / 2), beta32).astype(np.float32)
    has5 = np.floor(k64 + WIDTH / 2).astype(np.int64) - t0 == WIDTH
    w[~has5, WIDTH] = 0.0
    return np.mod(t, g), w, has5


def _taps(cs, gshape, beta32):
    gz, gy, gx = gshape
    iz, wz, fz = _axis_taps32(cs[:, 0], gz, beta32)
    iy, wy, fy = _axis_taps32(cs[:, 1], gy, beta32)
    ix, wx, fx = _axis_taps32(cs[:, 2], gx, beta32)
    return (iz * (gy * gx), wz), (iy * gx, wy), (ix, wx), fz | fy | fx


def _interp_block(flat, cs, gshape, beta32):
    (oz, wz), (oy, wy), (ox, wx), five = _taps(cs, gshape, beta32)
    out = np.zeros(cs.shape[0], dtype=np.complex64)
    for a in range(WIDTH):
        for b in range(WIDTH):
            zy = oz[:, a] + oy[:, b]
            wzy = wz[:, a] * wy[:, b]
            for c in range(WIDTH):
                out += (wzy * wx[:, c]) * flat[zy + ox[:, c]]
    sel = np.nonzero(five)[0]
    if sel.size:  # samples on a grid line: all 5 x 5 x 5 candidates (zero weight past the taps)
        acc = np.zeros(sel.size, dtype=np.complex64)
        for a in range(WIDTH + 1):
            for b in range(WIDTH + 1):
                zy = oz[sel, a] + oy[sel, b]
                wzy = wz[sel, a] * wy[sel, b]
                for c in range(WIDTH + 1):
                    acc += (wzy * wx[sel, c]) * flat[zy + ox[sel, c]]
        out[sel] = acc
    return out


def _grid_block(re, im, vals, cs, gshape, beta32):
    (oz, wz), (oy, wy), (ox, wx), five = _taps(cs, gshape, beta32)
    keep = ~five
    idx, wts = [], []
    for a in range(WIDTH):
        for b in range(WIDTH):
            zy = oz[keep, a] + oy[keep, b]
            wzy = wz[keep, a] * wy[keep, b]
            for c in range(WIDTH):
                idx.append(zy + ox[keep, c])
                wts.append(wzy * wx[keep, c])
    sel = np.nonzero(five)[0]
    vk = vals[keep]
    if sel.size:
        vk = np.concatenate([vk] * (WIDTH ** 3) + [vals[sel]] * ((WIDTH + 1) ** 3))
        for a in range(WIDTH + 1):
            for b in range(WIDTH + 1):
                zy = oz[sel, a] + oy[sel, b]
                wzy = wz[sel, a] * wy[sel, b]
                for c in range(WIDTH + 1):
                    idx.append(zy + ox[sel, c])
                    wts.append(wzy * wx[sel, c])
    else:
        vk = np.tile(vk, WIDTH ** 3)
    lin = np.concatenate(idx)
    contrib = np.concatenate(wts) * vk
    G = re.size
    re += np.bincount(lin, weights=contrib.real, minlength=G)
    im += np.bincount(lin, weights=contrib.imag, minlength=G)


def nufft(image, coord):
    """sp.nufft(image (Z, Y, X) complex64, coord (M, 3) float32) -> (M,) complex64."""
    image = np.asarray(image, dtype=np.complex64)
    shape = image.shape
    gshape = onufft.oversamp_shape(shape)
    beta = onufft.kb_beta()
    beta32 = float(np.float32(beta))
    x = _apodize32(image, _apod32(shape, gshape, beta))
    x = x / np.float32(math.sqrt(float(np.prod(shape))))
    x = onufft._resize_center(x, gshape)
    x = np.fft.fftshift(np.fft.fftn(np.fft.ifftshift(x)))
    flat = x.reshape(-1)
    cs = onufft.scale_coord(coord, shape).reshape(-1, 3)
    out = np.empty(cs.shape[0], dtype=np.complex64)
    for s in range(0, cs.shape[0], _TAP_CHUNK):
        out[s:s + _TAP_CHUNK] = _interp_block(flat, cs[s:s + _TAP_CHUNK], gshape, beta32)
    return out / np.float32(WIDTH ** 3)


def nufft_adjoint(kspace, coord, oshape):
    """sp.nufft_adjoint(kspace (M,) complex64, coord (M, 3), oshape (Z, Y, X)) -> complex64 image."""
    vals = np.asarray(kspace, dtype=np.complex64).reshape(-1)
    shape = tuple(int(v) for v in oshape)
    gshape = onufft.oversamp_shape(shape)
    beta = onufft.kb_beta()
    beta32 = float(np.float32(beta))
    cs = onufft.scale_coord(coord, shape).reshape(-1, 3)
    G = int(np.prod(gshape))
    re = np.zeros(G)
    im = np.zeros(G)
    for s in range(0, cs.shape[0], _TAP_CHUNK):
        _grid_block(re, im, vals[s:s + _TAP_CHUNK], cs[s:s + _TAP_CHUNK], gshape, beta32)
    x = (re + 1j * im).astype(np.complex64).reshape(gshape) / np.float32(WIDTH ** 3)
    x = np.fft.fftshift(np.fft.ifftn(np.fft.ifftshift(x)))
    x = onufft._resize_center(x, shape)
    x = x * np.float32(float(np.prod(gshape)) / math.sqrt(float(np.prod(shape))))
    return _apodize32(x, _apod32(shape, gshape, beta)).astype(np.complex64)


def data_channel(ubar_be, map_c, y_c, q_c, dcf_b, coord_b, g_be, sigma=1.0 / 6.0):
    """One (bin, coil, echo) of the TVME data term, in place on q_c and g_be:
    q <- (q + sigma (nufft(ubar S_c) - y)) / (1 + sigma);  g += conj(S_c) nufft^H(dcf q)."""
    sig = np.float32(sigma)
    r = nufft(ubar_be * map_c, coord_b) - y_c
    q_c += sig * r
    q_c /= np.float32(1.0 + sigma)
    g_be += np.conj(map_c) * nufft_adjoint(dcf_b * q_c, coord_b, ubar_be.shape)
    return q_c, g_be


def tvme_elementwise(u, ubar, g, p_m, p_ex, p_ey, p_ez, lam_m, lam_e, l2_coupling=True,
                     sigma=1.0 / 6.0, tau=1.0 / 6.0):
    """The TV part of one TVME iteration on complex64 (B, E, Z, Y, X) arrays, given the data
    term's adjoint g.  Returns (u, ubar, p_m, p_ex, p_ey, p_ez)."""
    u_old = u.copy()
    d_e = opdhg.echo_diff(ubar)
    p_m = p_m + sigma * opdhg.motion_diff(ubar)
    p_ex = p_ex + sigma * opdhg.fdiff(d_e, -1)
    p_ey = p_ey + sigma * opdhg.fdiff(d_e, -2)
    p_ez = p_ez + sigma * opdhg.fdiff(d_e, -3)
    p_ex, p_ey, p_ez = opdhg.prox_iso3(p_ex, p_ey, p_ez, lam_e)
    p_m = opdhg.prox_motion(p_m, lam_m, l2_coupling)
    divp = opdhg.fdiff_adj(p_ex, -1) + opdhg.fdiff_adj(p_ey, -2) + opdhg.fdiff_adj(p_ez, -3)
    divp_e = opdhg.echo_div(divp)
    divp_m = opdhg.motion_div_tvme(p_m)
    u = u.copy()
    u -= tau * divp_m
    u -= tau * divp_e
    u -= tau * g
    ubar = 2 * u - u_old
    return u, ubar, p_m, p_ex, p_ey, p_ez


def tvm_iteration(prob, st, lam, l2_coupling, sigma=1.0 / 6.0, tau=1.0 / 6.0):
    """One TVM iteration (tvm.py:227-380) in complex64 over every bin, coil and echo.

    prob: ``oracle.pdhg.BinnedProblem`` with complex64 k-space; st: dict of complex64
    arrays u, ubar, p_m (B, E, Z, Y, X) and q (list of (E, C, M_b)); updated in place."""
    B, E, C = prob.B, prob.E, prob.C
    u_old = st["u"].copy()
    g = np.zeros_like(st["u"])
    for b in range(B):
        for e in range(E):
            for c in range(C):
                data_channel(st["ubar"][b, e], prob.mps[c], prob.ksp[b][e, c], st["q"][b][e, c],
                             prob.dcf[b], prob.coord[b], g[b, e], sigma)
    p_m = st["p_m"] + sigma * opdhg.motion_diff(st["ubar"])
    st["p_m"] = opdhg.prox_motion(p_m, lam, l2_coupling).astype(np.complex64)
    div = opdhg.motion_div_tvm(st["p_m"])
    st["u"] = (st["u"] - tau * (g - div)).astype(np.complex64)
    st["ubar"] = (2 * st["u"] - u_old).astype(np.complex64)
    return st
