"""FFT-pass parity: the HIP pad/apodise/FFT passes against the centred FFT of SigPy's nufft."""

import math

import numpy as np
import pytest

from oracle import nufft as on

pytestmark = pytest.mark.gpu


def reference_grid(img):
    """Apodised, scaled, centred zero-padded image -> centred FFT, in FFT-natural order."""
    shape = img.shape[-3:]
    g = on.oversamp_shape(shape)
    beta = on.kb_beta()
    x = on._apodize(img.astype(np.complex128), shape, g, beta) / math.sqrt(float(np.prod(shape)))
    x = on._resize_center(x, g)
    axes = (-3, -2, -1)
    return np.fft.fftn(np.fft.ifftshift(x, axes=axes), axes=axes)  # natural order = fftshifted F


# grid lengths beyond the compile-time plans: 17, 47 and 31 are prime factors above 13 (Bluestein:
# 170 = 2 5 17, 188 = 4 47, 1023 = 3 11 31, 513 = 27 19), 275 = 25 11, 320 and 2000 are smooth
# lengths above 256 (smaller channel chunks)
GENERAL = [(13, 9, 7), (136, 8, 8), (8, 150, 8), (8, 8, 220), (6, 256, 6), (4, 410, 4), (4, 4, 818),
           (4, 4, 1600)]


@pytest.mark.parametrize("shape", [(12, 16, 16), (9, 11, 7), (8, 20, 6), (7, 9, 11), (16, 16, 24)] + GENERAL)
def test_forward_grid_matches_centred_fft(shape):
    import torch

    from toporecon_hip import _lib
    from toporecon_hip.nufft_backend import HipNufftBackend
    rng = np.random.default_rng(4)
    img = (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(np.complex64)
    be = HipNufftBackend(shape, [np.zeros((4, 3), np.float32)], 0)
    g = be.plan.grid_shape
    out = torch.empty(g, dtype=torch.complex64, device="cuda")
    x = torch.from_numpy(img).cuda()
    _lib.call("tr_nufft_forward_grid", be.plan.handle, x.data_ptr(), None, 1, 1, out.data_ptr(),
              _lib.stream_handle())
    hip = out.cpu().numpy()
    ref = reference_grid(img)
    err = np.abs(hip - ref)
    bad = np.argwhere(err > 1e-4 * np.abs(ref).max())
    print("shape", shape, "grid", g, "bad positions", len(bad), bad[:20].tolist(), flush=True)
    assert np.linalg.norm(hip - ref) / np.linalg.norm(ref) < 2e-6


@pytest.mark.parametrize("shape", [(12, 16, 16), (9, 11, 7), (8, 20, 6)] + GENERAL)
def test_adjoint_grid_matches_centred_ifft(shape):
    import torch

    from toporecon_hip import _lib
    from toporecon_hip.nufft_backend import HipNufftBackend
    rng = np.random.default_rng(5)
    be = HipNufftBackend(shape, [np.zeros((4, 3), np.float32)], 0)
    g = be.plan.grid_shape
    grid = (rng.standard_normal(g) + 1j * rng.standard_normal(g)).astype(np.complex64)
    out = torch.empty(shape, dtype=torch.complex64, device="cuda")
    gt = torch.from_numpy(grid).cuda()
    _lib.call("tr_nufft_adjoint_grid", be.plan.handle, gt.data_ptr(), None, 1, 1, out.data_ptr(),
              _lib.stream_handle())
    hip = out.cpu().numpy()
    # reference: natural-order grid -> unnormalised inverse FFT -> fftshift -> crop -> / sqrt(N) -> apodise
    beta = on.kb_beta()
    y = np.fft.ifftn(grid.astype(np.complex128)) * np.prod(g)
    y = np.fft.fftshift(y)
    y = on._resize_center(y, shape) / math.sqrt(float(np.prod(shape)))
    ref = on._apodize(y, shape, g, beta)
    assert np.linalg.norm(hip - ref) / np.linalg.norm(ref) < 2e-6
