"""CNN warm start for a HIP PDHG solver.

For every (bin, echo) image the network sees the density-compensated adjoint
g = A^H (dcf * y) divided by its RMS s and predicts the converged image up to
scale, v = s (x + f(x)).  The scale is then fitted to the data: per (b, e),

    c = Re <A v, y>_dcf / ||A v||^2_dcf      (float64 sums on the GPU)

and the warm start is u0 = c v.  PDHG then runs from u0 to the same
tolerance as the plain solve, so the network only changes where the
iterations start, never what they converge to; with no weights, or a shape
the network was not built for, the caller runs plain PDHG.

``cg_refine`` optionally follows with a few conjugate-gradient steps on the
data normal equations A^H W A u = A^H W y from u0 (float64 step sizes and
residual sums on the GPU): CG removes the well-sampled part of the remaining
error much faster than PDHG's fixed steps do.  ``warm_start`` chains the two.

DEFAULT_CG_STEPS = 4 was chosen on held-out problems: the training size
(64x96x96, 16 coils, 4 echoes, 6 bins, seeds 9000-9003) and the full size
(80x192x192, 32 coils, 6 echoes, 6 bins, generator seeds 2 and 3).  It is
the largest step count whose final objective stayed within +-0.1 % of the
plain solve on every one of them (4 steps: 2.5-3.3x faster to tol 1e-3 at
the training size, +4e-4 to +8e-4; 3.9x at full size, -1.4e-4; 5 steps ends
up to 0.14 % *below* the plain objective).
"""

from __future__ import annotations

import numpy as np

from .. import _lib

FIT_BLOCKS = 512
CG_BLOCKS = 1024
DEFAULT_CG_STEPS = 4


def dc_adjoint(solver, out=None):
    """g_b = A^H (dcf * y_b) for every bin, (B, E, Z, Y, X) complex64."""
    import torch
    if out is None:
        out = torch.empty(solver.vol_dims(), dtype=torch.complex64, device=solver.device)
    s = solver.stream
    for b in range(solver.B):
        _lib.call("tr_data_adjoint", solver.plan.handle, b, solver.y[b].data_ptr(), solver.dcf[b].data_ptr(),
                  solver.maps.data_ptr(), solver.E, solver.C, out[b].data_ptr(), 0, s)
    return out


def image_rms(img, n, V, device, stream):
    """Per-image RMS (float32, computed in float64) of n images of V voxels."""
    import torch
    ss = torch.empty(n, dtype=torch.float64, device=device)
    _lib.call("tr_cnn_sumsq", img.data_ptr(), V, n, ss.data_ptr(), stream)
    return torch.sqrt(ss / V).to(torch.float32)


def predict(net, img, shape, chunk=6):
    """Network prediction v = s (x + f(x)) for images (n, Z, Y, X) complex64; returns (v, x, out, s)."""
    import torch
    n = img.shape[0]
    V = int(np.prod(shape))
    dev = img.device
    s = net.stream
    scale = image_rms(img, n, V, dev, s)
    x = torch.empty((n, *shape, 8), dtype=torch.bfloat16, device=dev)
    _lib.call("tr_cnn_pack_input", img.data_ptr(), scale.data_ptr(), V, n, x.data_ptr(), s)
    out = torch.empty((n, *shape, 2), dtype=torch.float32, device=dev)
    for i0 in range(0, n, chunk):
        i1 = min(n, i0 + chunk)
        out[i0:i1] = net.forward(x[i0:i1], i1 - i0, shape)
    v = torch.empty_like(img)
    _lib.call("tr_cnn_unpack", x.data_ptr(), out.data_ptr(), scale.data_ptr(), V, n, v.data_ptr(), s)
    return v, x, out, scale


def fit_scales(solver, v, keep_av=False):
    """Least-squares scale per (bin, echo) of v (B, E, Z, Y, X) against the data (float64); with
    keep_av, also A v per bin ((M_b, E*C) sorted k-space), which the caller may reuse."""
    import torch
    s = solver.stream
    K = solver.E * solver.C
    c = np.zeros((solver.B, solver.E))
    work = torch.empty((FIT_BLOCKS, K, 2), dtype=torch.float64, device=solver.device)
    sums = torch.empty((K, 2), dtype=torch.float64, device=solver.device)
    avs = []
    for b in range(solver.B):
        M = solver.y[b].shape[0]
        av = torch.empty_like(solver.y[b])
        if keep_av:
            avs.append(av)
        if M == 0:
            continue
        _lib.call("tr_data_simulate", solver.plan.handle, b, v[b].data_ptr(), solver.maps.data_ptr(), solver.E,
                  solver.C, av.data_ptr(), s)
        _lib.call("tr_data_fit_sums", av.data_ptr(), solver.y[b].data_ptr(), solver.dcf[b].data_ptr(), M, K,
                  work.data_ptr(), FIT_BLOCKS, sums.data_ptr(), s)
        h = sums.cpu().numpy().reshape(solver.E, solver.C, 2).sum(axis=1)
        c[b] = np.where(h[:, 1] > 0, h[:, 0] / np.where(h[:, 1] > 0, h[:, 1], 1.0), 0.0)
        del av
    return (c, avs) if keep_av else c


def cnn_warm_start(solver, net, chunk=6, keep_rhs=False, keep_ax=False):
    """u0 (B, E, Z, Y, X) complex64 for ``solver.run(init=u0, dual_init="residual")``; with
    keep_rhs, also A^H W y (the network's input, which ``cg_refine`` can reuse as its rhs); with
    keep_ax, also A u0 per bin (sorted k-space), formed as c * A v from the scale fit (A is linear)
    instead of another forward NUFFT.  Returns u0, or a tuple (u0[, rhs][, ax])."""
    import torch
    shape = tuple(solver.img_shape)
    B, E, C = solver.B, solver.E, solver.C
    V = int(np.prod(shape))
    g = dc_adjoint(solver)
    imgs = g.reshape(B * E, *shape)
    v, x, out, scale = predict(net, imgs, shape, chunk=chunk)
    if not keep_rhs:
        del g
    fit = fit_scales(solver, v.reshape(B, E, *shape), keep_av=keep_ax)
    c, avs = fit if keep_ax else (fit, None)
    scale2 = scale * torch.from_numpy(c.reshape(-1).astype(np.float32)).to(scale.device)
    _lib.call("tr_cnn_unpack", x.data_ptr(), out.data_ptr(), scale2.data_ptr(), V, B * E, v.data_ptr(),
              net.stream)
    u0 = v.reshape(B, E, *shape)
    if keep_ax:
        for b in range(B):
            cb = torch.from_numpy(np.repeat(c[b], C).astype(np.float32)).to(solver.device)  # channel e*C + c
            _lib.call("tr_data_scale_channels", avs[b].data_ptr(), int(avs[b].shape[0]), E * C, cb.data_ptr(),
                      solver.stream)
    res = (u0,) + ((g,) if keep_rhs else ()) + ((avs,) if keep_ax else ())
    return res if len(res) > 1 else u0


def _host_sum(partial):
    return float(partial.cpu().numpy().sum())  # block sums added in block order on the host


def _normal(solver, x, out, qtmp):
    """out = A^H W A x over all bins."""
    s = solver.stream
    for b in range(solver.B):
        _lib.call("tr_data_simulate", solver.plan.handle, b, x[b].data_ptr(), solver.maps.data_ptr(), solver.E,
                  solver.C, qtmp[b].data_ptr(), s)
        _lib.call("tr_data_adjoint", solver.plan.handle, b, qtmp[b].data_ptr(), solver.dcf[b].data_ptr(),
                  solver.maps.data_ptr(), solver.E, solver.C, out[b].data_ptr(), 0, s)


def _forward_weighted_sq(solver, x, qtmp, work, sums):
    """qtmp[b] = A x per bin; returns <A x, W A x> = sum over bins, channels and samples of
    dcf |A x|^2 (float64 per channel on the device, added in order on the host).  It equals
    <x, A^H W A x> without the adjoint NUFFT."""
    s = solver.stream
    K = solver.E * solver.C
    for b in range(solver.B):
        M = int(solver.y[b].shape[0])
        _lib.call("tr_data_simulate", solver.plan.handle, b, x[b].data_ptr(), solver.maps.data_ptr(), solver.E,
                  solver.C, qtmp[b].data_ptr(), s)
        if M == 0:
            sums[b].zero_()
            continue
        _lib.call("tr_data_fit_sums", qtmp[b].data_ptr(), qtmp[b].data_ptr(), solver.dcf[b].data_ptr(), M, K,
                  work.data_ptr(), FIT_BLOCKS, sums[b].data_ptr(), s)
    return float(sums.cpu().numpy()[:, :, 1].sum())


def _data_adjoint_all(solver, q, out):
    """out = A^H W q over all bins, q per bin in sorted order."""
    s = solver.stream
    for b in range(solver.B):
        _lib.call("tr_data_adjoint", solver.plan.handle, b, q[b].data_ptr(), solver.dcf[b].data_ptr(),
                  solver.maps.data_ptr(), solver.E, solver.C, out[b].data_ptr(), 0, s)


def cg_refine(solver, x0, steps: int, rhs=None, x0_zero: bool = False, ax0=None, track_ax: bool = False):
    """`steps` conjugate-gradient iterations on A^H W A u = A^H W y from x0 (returns a new array).

    rhs: A^H W y if the caller has it (the buffer is consumed as the residual); x0_zero: x0 is
    known to be zero, so the initial residual is the rhs itself; ax0: A x0 per bin (sorted k-space)
    if the caller has it, so the initial residual needs only the adjoint.  track_ax: keep A x in
    k-space along the steps (A x += alpha A p, with the A p each step computes anyway) and return
    (x, A x per bin) -- ax0 (updated in place) or zeros when x0 is zero.  The last step runs the
    forward NUFFT only: its step length needs <A p, W A p>, and its residual update (the adjoint) is
    never used."""
    import torch
    if steps <= 0 and not track_ax:
        return x0
    s = solver.stream
    n = x0.numel()
    dev = x0.device
    x = torch.empty_like(x0)
    _lib.call("tr_cscale", x0.data_ptr(), n, 1.0, x.data_ptr(), s)
    ax = None
    if track_ax:
        ax = ax0 if ax0 is not None else [torch.zeros_like(solver.y[b]) for b in range(solver.B)]
    if steps <= 0:
        return x, ax
    r = rhs if rhs is not None else dc_adjoint(solver)
    qtmp = [torch.empty_like(solver.y[b]) for b in range(solver.B)]
    ap = torch.empty_like(x)
    p = torch.empty_like(x)
    partial = torch.empty(CG_BLOCKS, dtype=torch.float64, device=dev)
    if x0_zero:
        _lib.call("tr_cdot_re", r.data_ptr(), r.data_ptr(), n, partial.data_ptr(), CG_BLOCKS, s)
    else:
        if ax0 is not None:
            _data_adjoint_all(solver, ax0, ap)
        else:
            _normal(solver, x, ap, qtmp)
        # r = A^H W y - A^H W A x (p is scratch for the x-side update here)
        _lib.call("tr_cg_update", p.data_ptr(), r.data_ptr(), ap.data_ptr(), ap.data_ptr(), n, 1.0,
                  partial.data_ptr(), CG_BLOCKS, s)
    rr = _host_sum(partial)
    _lib.call("tr_cscale", r.data_ptr(), n, 1.0, p.data_ptr(), s)
    K = solver.E * solver.C
    fit_work = torch.empty((FIT_BLOCKS, K, 2), dtype=torch.float64, device=dev)
    fit_sums = torch.empty((solver.B, K, 2), dtype=torch.float64, device=dev)
    for i in range(steps):
        if rr <= 0.0:
            break
        last = i == steps - 1
        if last:
            # the last step needs only alpha, and <p, A^H W A p> = <A p, W A p> is summed in k-space:
            # the adjoint NUFFT, which would only update the residual, is not run
            pap = _forward_weighted_sq(solver, p, qtmp, fit_work, fit_sums)
        else:
            _normal(solver, p, ap, qtmp)
            _lib.call("tr_cdot_re", p.data_ptr(), ap.data_ptr(), n, partial.data_ptr(), CG_BLOCKS, s)
            pap = _host_sum(partial)
        if not pap > 0.0:
            break
        alpha = rr / pap
        if last:
            _lib.call("tr_js_axpy", x.data_ptr(), p.data_ptr(), n, alpha, s)  # x += alpha p
        else:
            _lib.call("tr_cg_update", x.data_ptr(), r.data_ptr(), p.data_ptr(), ap.data_ptr(), n, alpha,
                      partial.data_ptr(), CG_BLOCKS, s)
        if ax is not None:  # A x += alpha A p (qtmp holds A p of this step)
            for b in range(solver.B):
                _lib.call("tr_js_axpy", ax[b].data_ptr(), qtmp[b].data_ptr(), qtmp[b].numel(), alpha, s)
        if last:
            break
        rr_new = _host_sum(partial)
        _lib.call("tr_cg_direction", p.data_ptr(), r.data_ptr(), n, rr_new / rr, s)
        rr = rr_new
    del qtmp, ap, p, r
    return (x, ax) if track_ax else x


def warm_start(solver, net=None, cg_steps: int = DEFAULT_CG_STEPS, chunk=6, with_dual: bool = False):
    """Warm start u0: the network's estimate (zeros without a network), then cg_steps CG steps.

    with_dual: return (u0, q0) with q0 = A u0 - y per bin (sorted), the data dual's fixed point, for
    ``solver.run(init=u0, dual_init=q0)``; A u0 comes from the scale fit's A v and the CG steps' A p
    (A is linear), so no forward NUFFT is spent on it."""
    import torch
    if net is not None:
        if not with_dual:
            if cg_steps <= 0:
                return cnn_warm_start(solver, net, chunk=chunk)
            u0, rhs, ax0 = cnn_warm_start(solver, net, chunk=chunk, keep_rhs=True, keep_ax=True)
            return cg_refine(solver, u0, cg_steps, rhs=rhs, ax0=ax0)
        u0, rhs, ax0 = cnn_warm_start(solver, net, chunk=chunk, keep_rhs=True, keep_ax=True)
        u, ax = cg_refine(solver, u0, cg_steps, rhs=rhs, ax0=ax0, track_ax=True)
    else:
        u0 = torch.zeros(solver.vol_dims(), dtype=torch.complex64, device=solver.device)
        if not with_dual:
            return cg_refine(solver, u0, cg_steps, x0_zero=True)
        u, ax = cg_refine(solver, u0, cg_steps, x0_zero=True, track_ax=True)
    for b in range(solver.B):  # q0 = A u - y
        _lib.call("tr_js_axpy", ax[b].data_ptr(), solver.y[b].data_ptr(), ax[b].numel(), -1.0, solver.stream)
    return u, ax
