"""The warm-start network oracle (oracle/cnn.py) pinned by torch's CPU conv3d and
by central finite differences."""

import numpy as np
import pytest

from oracle import cnn


def _torch_conv(x, w, b):
    import torch
    xt = torch.from_numpy(np.ascontiguousarray(np.moveaxis(x, -1, 0)))[None]
    wt = torch.from_numpy(np.ascontiguousarray(
        np.moveaxis(w.reshape(3, 3, 3, w.shape[1], w.shape[2]), (3, 4), (1, 0))))
    y = torch.nn.functional.conv3d(xt, wt, torch.from_numpy(b), padding=1)
    return np.moveaxis(y[0].numpy(), 0, -1)


def test_conv3_matches_torch_cpu():
    rng = np.random.default_rng(1)
    x = rng.standard_normal((5, 6, 7, 3))
    w = rng.standard_normal((27, 3, 4))
    b = rng.standard_normal(4)
    np.testing.assert_allclose(cnn.conv3(x, w, b), _torch_conv(x, w, b), rtol=1e-12, atol=1e-12)


def test_conv3_tap_numbering():
    # a single tap (dz, dy, dx) = (1, -1, 0) reads x[z + 1, y - 1, x]
    x = np.zeros((4, 4, 4, 1))
    x[2, 1, 3, 0] = 1.0
    w = np.zeros((27, 1, 1))
    w[9 * 2 + 3 * 0 + 1, 0, 0] = 1.0
    y = cnn.conv3(x, w)
    assert y[1, 2, 3, 0] == 1.0 and np.count_nonzero(y) == 1


def test_conv3_backward_is_the_adjoint():
    rng = np.random.default_rng(2)
    x = rng.standard_normal((4, 5, 6, 3))
    w = rng.standard_normal((27, 3, 2))
    dy = rng.standard_normal((4, 5, 6, 2))
    dx, dw, db = cnn.conv3_backward(x, w, dy)
    # <conv(x), dy> = <x, dx>  and  d<conv(x; w), dy>/dw = dw
    assert np.isclose(np.sum(cnn.conv3(x, w) * dy), np.sum(x * dx), rtol=1e-12)
    for t in (0, 13, 26):
        e = np.zeros_like(w)
        e[t, 1, 0] = 1.0
        assert np.isclose(np.sum(cnn.conv3(x, e) * dy), dw[t, 1, 0], rtol=1e-12)
    assert np.allclose(db, dy.sum(axis=(0, 1, 2)))


def test_network_gradients_by_finite_differences():
    rng = np.random.default_rng(3)
    params = cnn.init_params(seed=4, channels=(2, 4, 4, 2), scale=3.0)
    x = rng.standard_normal((3, 4, 5, 2))
    target = rng.standard_normal((3, 4, 5, 2))
    pred, cache = cnn.forward(params, x)
    loss, dpred = cnn.mse(pred, target)
    grads = cnn.backward(params, cache, dpred)
    h = 1e-6
    for li, (t, ci, co) in [(0, (4, 1, 2)), (1, (13, 3, 0)), (2, (26, 2, 1)), (1, (0, 0, 3))]:
        w, b = params[li]
        wp, wm = w.copy(), w.copy()
        wp[t, ci, co] += h
        wm[t, ci, co] -= h
        pp = [(wp, b) if i == li else p for i, p in enumerate(params)]
        pm = [(wm, b) if i == li else p for i, p in enumerate(params)]
        fd = (cnn.mse(cnn.forward(pp, x)[0], target)[0] - cnn.mse(cnn.forward(pm, x)[0], target)[0]) / (2 * h)
        assert np.isclose(grads[li][0][t, ci, co], fd, rtol=1e-5, atol=1e-9)
    w, b = params[1]
    bp, bm = b.copy(), b.copy()
    bp[2] += h
    bm[2] -= h
    fd = (cnn.mse(cnn.forward([params[0], (w, bp), params[2]], x)[0], target)[0]
          - cnn.mse(cnn.forward([params[0], (w, bm), params[2]], x)[0], target)[0]) / (2 * h)
    assert np.isclose(grads[1][1][2], fd, rtol=1e-5, atol=1e-9)


def test_adam_first_step_moves_by_lr():
    p, g = np.array([1.0, -2.0]), np.array([0.5, -3.0])
    p1, m, v = cnn.adam(p, g, np.zeros(2), np.zeros(2), 1, lr=0.01)
    np.testing.assert_allclose(p1, p - 0.01 * np.sign(g), rtol=1e-6)


def test_bf16_rounding():
    vals = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, -2.5, 3.14159], np.float32)
    r = cnn.bf16(vals)
    assert r[0] == 1.0 and r[1] == 1.0 and r[2] == 1.0 + 2 ** -6 and r[3] == -2.5
    assert abs(r[4] - 3.140625) < 1e-12
