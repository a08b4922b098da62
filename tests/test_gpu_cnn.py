"""Warm-start network kernels (MFMA bf16, fp32 accumulation) against the float64 oracle.

Inputs and weights are rounded to bf16 before the oracle sees them, so the
only differences are the fp32 accumulation order (relative ~1e-6) and, where
a kernel stores bf16, one rounding of its output (relative 2^-8)."""

import numpy as np
import pytest

from oracle import cnn

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_mod():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def _bf16_tensor(torch, a):
    return torch.from_numpy(np.asarray(a, np.float32)).to(torch.bfloat16).cuda()


def _lib():
    from toporecon_hip import _lib as lib
    return lib


def _frag(torch, w_pad, cin, flip=0):
    lib = _lib()
    wt = torch.from_numpy(np.asarray(w_pad, np.float32)).cuda()
    ncb = (32 if flip else cin) // 8
    frag = torch.empty(27 * ncb * 64 * 4, dtype=torch.bfloat16, device="cuda")
    lib.call("tr_cnn_pack_weights", wt.data_ptr(), cin, flip, frag.data_ptr(), lib.stream_handle(0))
    return frag


def _rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


@pytest.mark.parametrize("cin,shape", [(8, (2, 3, 5, 37)), (32, (1, 4, 6, 33)), (32, (2, 2, 4, 32))])
def test_conv_forward_modes(torch_mod, cin, shape):
    torch = torch_mod
    lib = _lib()
    rng = np.random.default_rng(cin + shape[-1])
    n, Z, Y, X = shape
    x = cnn.bf16(rng.standard_normal((n, Z, Y, X, cin)))
    w = cnn.bf16(rng.standard_normal((27, cin, 32)) * 0.2)
    b = rng.standard_normal(32).astype(np.float32)
    xt = _bf16_tensor(torch, x)
    frag = _frag(torch, w, cin)
    bt = torch.from_numpy(b).cuda()
    s = lib.stream_handle(0)
    ref = np.stack([cnn.conv3(x[i], w, b.astype(np.float64)) for i in range(n)])
    out1 = torch.empty((n, Z, Y, X, 5), dtype=torch.float32, device="cuda")
    lib.call("tr_cnn_conv", xt.data_ptr(), n, Z, Y, X, cin, frag.data_ptr(), bt.data_ptr(), 1, None,
             out1.data_ptr(), 5, s)
    assert _rel(out1.cpu().numpy(), ref[..., :5]) < 2e-6
    out0 = torch.empty((n, Z, Y, X, 32), dtype=torch.bfloat16, device="cuda")
    lib.call("tr_cnn_conv", xt.data_ptr(), n, Z, Y, X, cin, frag.data_ptr(), bt.data_ptr(), 0, None,
             out0.data_ptr(), 0, s)
    got = out0.float().cpu().numpy()
    want = np.maximum(ref, 0.0)
    # one bf16 rounding of the fp32 result (half an ulp = 2^-9 relative) plus fp32 accumulation
    assert np.all(np.abs(got - want) <= 2.0 ** -8 * np.abs(want) + 1e-5)


def test_conv_data_gradient_with_mask(torch_mod):
    torch = torch_mod
    lib = _lib()
    rng = np.random.default_rng(7)
    n, Z, Y, X = 2, 3, 5, 35
    w = cnn.bf16(rng.standard_normal((27, 32, 32)) * 0.2)
    dy = cnn.bf16(rng.standard_normal((n, Z, Y, X, 32)))
    act = cnn.bf16(np.maximum(rng.standard_normal((n, Z, Y, X, 32)), 0.0))
    frag_t = _frag(torch, w, 32, flip=1)
    out = torch.empty((n, Z, Y, X, 32), dtype=torch.bfloat16, device="cuda")
    dyt, actt = _bf16_tensor(torch, dy), _bf16_tensor(torch, act)  # keep both alive across the launch
    lib.call("tr_cnn_conv", dyt.data_ptr(), n, Z, Y, X, 32, frag_t.data_ptr(), None, 2,
             actt.data_ptr(), out.data_ptr(), 0, lib.stream_handle(0))
    ref = np.stack([cnn.conv3_backward(np.zeros((Z, Y, X, 32)), w, dy[i])[0] for i in range(n)]) * (act > 0)
    got = out.float().cpu().numpy()
    assert np.all(got[act <= 0] == 0)
    assert _rel(got, ref) < 4e-3


@pytest.mark.parametrize("cin,nblocks", [(8, 7), (32, 64)])
def test_weight_gradient(torch_mod, cin, nblocks):
    torch = torch_mod
    lib = _lib()
    rng = np.random.default_rng(11 + cin)
    n, Z, Y, X = 2, 3, 6, 34
    x = cnn.bf16(rng.standard_normal((n, Z, Y, X, cin)))
    dy = cnn.bf16(rng.standard_normal((n, Z, Y, X, 32)))
    work = torch.empty(lib.load().tr_cnn_wgrad_work_size(nblocks), dtype=torch.float32, device="cuda")
    dw = torch.empty((27, cin, 32), dtype=torch.float32, device="cuda")
    db = torch.empty(32, dtype=torch.float32, device="cuda")
    xt, dyt = _bf16_tensor(torch, x), _bf16_tensor(torch, dy)  # keep both alive across the launch
    lib.call("tr_cnn_wgrad", xt.data_ptr(), dyt.data_ptr(), n, Z, Y, X, cin, nblocks, work.data_ptr(),
             dw.data_ptr(), db.data_ptr(), lib.stream_handle(0))
    ref_w = sum(cnn.conv3_backward(x[i], np.zeros((27, cin, 32)), dy[i])[1] for i in range(n))
    ref_b = dy.reshape(-1, 32).sum(axis=0)
    assert _rel(dw.cpu().numpy(), ref_w) < 1e-5
    assert _rel(db.cpu().numpy(), ref_b) < 1e-5


def test_adam_kernel_matches_oracle(torch_mod):
    torch = torch_mod
    lib = _lib()
    rng = np.random.default_rng(3)
    p, g = rng.standard_normal(1000), rng.standard_normal(1000)
    m, v = rng.standard_normal(1000) * 0.1, np.abs(rng.standard_normal(1000)) * 0.1
    ref = cnn.adam(p, g, m, v, 5, lr=1e-3)
    tp, tg, tm, tv = (torch.from_numpy(a.astype(np.float32)).cuda() for a in (p, g, m, v))
    lib.call("tr_cnn_adam", tp.data_ptr(), tg.data_ptr(), tm.data_ptr(), tv.data_ptr(), 1000, 1e-3, 0.9, 0.999,
             1e-8, 5, lib.stream_handle(0))
    for got, want in zip((tp, tm, tv), ref):
        np.testing.assert_allclose(got.cpu().numpy(), want, rtol=2e-5, atol=1e-6)


def test_network_forward_and_gradients_match_oracle(torch_mod):
    """Whole network on the GPU (bf16 activations and gradients) vs the oracle with the
    same bf16-rounded weights and input.  Against the oracle that rounds activations and
    gradients to bf16 at the same points the kernels do, predictions and every parameter
    gradient agree to fp32-accumulation level; against the plain float64 oracle they
    agree to what bf16 storage allows."""
    torch = torch_mod
    from toporecon_hip.learned import WarmStartNet
    rng = np.random.default_rng(5)
    shape = (4, 6, 40)
    params = [(cnn.bf16(w), b) for w, b in cnn.init_params(seed=8)]
    params = [(w, rng.standard_normal(len(b)) * 0.05) for w, b in params]
    net = WarmStartNet(0, params=params)
    x2 = cnn.bf16(rng.standard_normal((*shape, 2)))
    target = rng.standard_normal((*shape, 2)) * 0.5
    x8 = np.zeros((1, *shape, 8))
    x8[0, ..., :2] = x2
    xt = _bf16_tensor(torch, x8)
    out, acts = net.forward(xt, 1, shape, keep=True)
    pred = x2 + out[0].cpu().numpy()
    resid_target = torch.from_numpy((target - x2)[None].astype(np.float32)).cuda()
    partial, dy, count = net.loss_grad(out, resid_target, 1, shape)
    loss = float(partial[:, 0].sum()) / count
    gw, gb = net.backward(acts, dy, 1, shape)
    for act_round, tol_pred, tol_grad in ((cnn.bf16, 1e-4, 2e-3), (None, 2e-2, 1e-1)):
        ref_pred, cache = cnn.forward(params, x2, act_round=act_round)
        assert _rel(pred, ref_pred) < tol_pred
        ref_loss, dpred = cnn.mse(ref_pred, target)
        assert abs(loss - ref_loss) / ref_loss < tol_pred
        ref_grads = cnn.backward(params, cache, dpred, grad_round=act_round)
        for li in range(5):
            ci, co = params[li][0].shape[1:]
            assert _rel(gw[li].cpu().numpy()[:, :ci, :co], ref_grads[li][0]) < tol_grad, (li, act_round)
            assert _rel(gb[li].cpu().numpy()[:co], ref_grads[li][1]) < tol_grad, (li, act_round)
    for li in range(5):
        ci, co = params[li][0].shape[1:]
        # padded rows / columns get exactly zero gradient
        assert np.all(gw[li].cpu().numpy()[:, ci:, :] == 0) and np.all(gw[li].cpu().numpy()[:, :, co:] == 0)
