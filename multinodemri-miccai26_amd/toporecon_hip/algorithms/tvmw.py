"""Motion-TV + db1 echo wavelet + db6 spatial wavelet PDHG (TVMW) on one MI355X.

Same constructor, state and result as ``TvmwReconstructor`` in
src/toporecon/algorithms/tvmw.py:213-786 on a single process.  One iteration
(tvmw.py:385-685): the db6 zero-mode wavelet of every bin x echo image in one
batched launch chain plus the projected dual update of all sub-bands
(tvmw.py:447-452, 162-193), the fused forward/data-dual chain per bin, one
kernel for the motion and db1 echo duals, the adjoint chain per bin, the
synthesis of the wavelet duals, and one primal kernel with the TVMW stopping
sums on ubar (tvmw.py:717-769).
"""

from __future__ import annotations

import math

import torch

from .. import _lib
from ._base import HipPdhgBase

FILTER_LEN = 12


def db6_max_level(shape) -> int:
    """PyWavelets' dwtn_max_level for a 12-tap filter: min over axes of floor(log2(n / 11))."""
    levels = []
    for n in shape:
        n = int(n)
        levels.append(0 if n < FILTER_LEN - 1 else int(math.floor(math.log2(n / (FILTER_LEN - 1)))))
    return min(levels)


class TvmwReconstructor(HipPdhgBase):
    """TVMW PDHG reconstruction with the HIP NUFFT and wavelets."""

    algorithm = "tvmw"

    def __init__(self, ksp, coord, dcf, mps, resp, dual_q, B_local, l2_coupling=False, lambda1=1e-5,
                 lambda2=1e-5, lambda3=1e-5, sigma=0.1, tau=0.1, max_iter=300, tol=1e-3, device=0,
                 margin=2, E_total=None, e0=None, e1=None, echo_left_peer=None, echo_right_peer=None,
                 B_total=None, b0=None, b1=None, motion_left_peer=None, motion_right_peer=None,
                 bin_edges=None, world_comm=None, group_comm=None, leader_comm=None, nccl_group=None,
                 dist_grid=None, tagger=None, show_pbar=True, nufft_backend="hip", echoes=None,
                 level_w2=None):
        self.lambda1 = lambda1
        self.lambda2 = lambda2
        self.lambda3 = lambda3
        self.l2_coupling = bool(l2_coupling)
        E = int(ksp.shape[0]) if ksp is not None else int(echoes)
        self.E_total = int(E_total) if E_total is not None else E
        self.e0 = int(e0) if e0 is not None else 0
        self.e1 = int(e1) if e1 is not None else self.e0 + E
        if self.e0 != 0 or self.e1 != self.E_total:
            raise NotImplementedError("echo-axis sharding needs halo exchanges; one GPU owns all echoes")
        self._setup(ksp, coord, dcf, mps, resp, dual_q, B_local, margin=margin, device=device,
                    B_total=B_total, b0=b0, b1=b1, bin_edges=bin_edges,
                    comms=(world_comm, group_comm, leader_comm), show_pbar=show_pbar,
                    nufft_backend=nufft_backend, sigma=sigma, tau=tau, max_iter=max_iter, tol=tol,
                    echoes=echoes)
        Z, Y, X = self.img_shape
        self.level_w2 = db6_max_level(self.img_shape) if level_w2 is None else int(level_w2)
        lib = _lib.load()
        self.ncoef = int(lib.tr_w2_coeff_count(Z, Y, X, self.level_w2))
        self.nwork = int(lib.tr_w2_work_count(Z, Y, X, self.level_w2))
        if self.ncoef < 0 or self.nwork < 0:
            raise ValueError(f"invalid wavelet level {self.level_w2} for {self.img_shape}")

    def pdhg_init(self, init=None):
        shape = self.vol_dims()
        dev = self.device
        B, E = shape[:2]
        self.ubar = torch.zeros(shape, dtype=torch.complex64, device=dev)
        self.u = torch.zeros_like(self.ubar)
        self.p_m = torch.zeros_like(self.ubar)
        self.p_w1 = torch.zeros_like(self.ubar)
        self.p_w2 = torch.zeros((B * E, self.ncoef), dtype=torch.complex64, device=dev)
        self.w2 = torch.empty_like(self.p_w2)
        self.divw2 = torch.empty_like(self.ubar)
        self.w2_work = torch.empty((max(B * E * self.nwork, 1),), dtype=torch.complex64, device=dev)
        self.g = torch.empty_like(self.ubar)
        if init is not None:
            self.ubar.copy_(init)
            self.u.copy_(init)
        self._alloc_reduction()

    def w2_forward(self, x, out):
        B, E, Z, Y, X = self.vol_dims()
        _lib.call("tr_w2_forward", x.data_ptr(), B * E, Z, Y, X, self.level_w2, out.data_ptr(),
                  self.w2_work.data_ptr(), self.stream)

    def w2_adjoint(self, c, out):
        B, E, Z, Y, X = self.vol_dims()
        _lib.call("tr_w2_adjoint", c.data_ptr(), B * E, Z, Y, X, self.level_w2, out.data_ptr(),
                  self.w2_work.data_ptr(), self.stream)

    def _tv_normal(self, x, u, ubar_scr, g, scratch):
        B, E, Z, Y, X = dims = self.vol_dims()
        s = self.stream
        if not scratch:
            scratch["p_m"] = torch.zeros_like(x)
            scratch["p_w1"] = torch.zeros_like(x)
            scratch["p_w2"] = torch.zeros((B * E, self.ncoef), dtype=torch.complex64, device=self.device)
            scratch["w2"] = torch.empty_like(scratch["p_w2"])
            scratch["divw2"] = torch.empty_like(x)
            scratch["work"] = torch.empty((max(B * E * self.nwork, 1),), dtype=torch.complex64,
                                          device=self.device)
        for k in ("p_m", "p_w1", "p_w2"):
            scratch[k].zero_()
        work = scratch["work"].data_ptr()
        _lib.call("tr_w2_forward", x.data_ptr(), B * E, Z, Y, X, self.level_w2, scratch["w2"].data_ptr(), work, s)
        _lib.call("tr_w2_dual_update", scratch["p_w2"].data_ptr(), scratch["w2"].data_ptr(),
                  scratch["p_w2"].numel(), 1.0, 1e30, s)
        _lib.call("tr_tvmw_dual", *dims, x.data_ptr(), scratch["p_m"].data_ptr(), scratch["p_w1"].data_ptr(),
                  1.0, 1e30, 1e30, int(self.l2_coupling), s)
        _lib.call("tr_w2_adjoint", scratch["p_w2"].data_ptr(), B * E, Z, Y, X, self.level_w2,
                  scratch["divw2"].data_ptr(), work, s)
        ubar_scr.zero_()
        _lib.call("tr_tvmw_primal", *dims, u.data_ptr(), ubar_scr.data_ptr(), g.data_ptr(),
                  scratch["p_m"].data_ptr(), scratch["p_w1"].data_ptr(), scratch["divw2"].data_ptr(), -1.0,
                  self._partial.data_ptr(), s)
        # the TVMW stop sums compare ubar before (0) and after (2 u): num = 4 |u|^2
        return self._reduce()[0] / 4.0

    def pdhg_launch(self, it):
        """Launch one PDHG iteration (tvmw.py:385-685); the float64 stopping sums
        (sum|dubar|^2, sum|ubar_prev|^2) land in self._red."""
        dims = self.vol_dims()
        s = self.stream
        self.w2_forward(self.ubar, self.w2)
        _lib.call("tr_w2_dual_update", self.p_w2.data_ptr(), self.w2.data_ptr(), self.p_w2.numel(),
                  float(self.sigma), float(self.lambda3), s)
        self.data_dual_update(self.ubar)
        _lib.call("tr_tvmw_dual", *dims, self.ubar.data_ptr(), self.p_m.data_ptr(), self.p_w1.data_ptr(),
                  float(self.sigma), float(self.lambda1), float(self.lambda2), int(self.l2_coupling), s)
        self.data_adjoint(self.g)
        self.w2_adjoint(self.p_w2, self.divw2)
        _lib.call("tr_tvmw_primal", *dims, self.u.data_ptr(), self.ubar.data_ptr(), self.g.data_ptr(),
                  self.p_m.data_ptr(), self.p_w1.data_ptr(), self.divw2.data_ptr(), float(self.tau),
                  self._partial.data_ptr(), s)
        self._reduce_launch()

    def _converged_value(self, num, den):
        """TVMW rule: ||ubar_prev - ubar|| / max(||ubar_prev||, 1e-12), inf if not finite."""
        if math.isfinite(num) and math.isfinite(den):
            return math.sqrt(max(num, 0.0)) / max(math.sqrt(max(den, 0.0)), 1e-12)
        return float("inf")
