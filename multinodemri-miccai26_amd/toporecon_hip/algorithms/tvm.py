"""Motion-TV PDHG (TVM) on one MI355X.

Same constructor, state and result as ``TvmReconstructor`` in
src/toporecon/algorithms/tvm.py:63-491 on a single process; the iteration
(tvm.py:227-380) runs as one fused forward/dual chain per motion bin, one
motion-dual kernel, one adjoint chain per bin and one primal kernel that also
produces the float64 stopping sums (tvm.py:416-463).
"""

from __future__ import annotations

import torch

from .. import _lib
from ._base import HipPdhgBase


class TvmReconstructor(HipPdhgBase):
    """Motion-TV PDHG reconstruction (L1 or echo-L2 coupling) with the HIP NUFFT."""

    algorithm = "tvm"

    def __init__(self, ksp, coord, dcf, mps, resp, dual_q, B_local, l2_coupling=False, lamda=1e-5,
                 sigma=0.01, tau=0.01, max_iter=300, tol=1e-3, margin=10, device=0, B_total=None,
                 b0=None, b1=None, motion_left_peer=None, motion_right_peer=None, bin_edges=None,
                 world_comm=None, group_comm=None, leader_comm=None, nccl_group=None, dist_grid=None,
                 tagger=None, show_pbar=True, reduce_per_bin=False, nufft_backend="hip", echoes=None):
        self.lamda = lamda
        self.l2_coupling = bool(l2_coupling)
        self.reduce_per_bin = reduce_per_bin
        self._setup(ksp, coord, dcf, mps, resp, dual_q, B_local, margin=margin, device=device,
                    B_total=B_total, b0=b0, b1=b1, bin_edges=bin_edges,
                    comms=(world_comm, group_comm, leader_comm), show_pbar=show_pbar,
                    nufft_backend=nufft_backend, sigma=sigma, tau=tau, max_iter=max_iter, tol=tol,
                    echoes=echoes)

    def pdhg_init(self, init=None):
        shape = self.vol_dims()
        dev = self.device
        self.ubar = torch.zeros(shape, dtype=torch.complex64, device=dev)
        self.u = torch.zeros_like(self.ubar)
        self.p_m = torch.zeros_like(self.ubar)
        self.g = torch.empty_like(self.ubar)
        if init is not None:
            self.ubar.copy_(init)
            self.u.copy_(init)
        self._alloc_reduction()

    def _tv_normal(self, x, u, ubar_scr, g, scratch):
        dims = self.vol_dims()
        s = self.stream
        if "p_m" in scratch:
            scratch["p_m"].zero_()
        else:
            scratch["p_m"] = torch.zeros_like(x)
        _lib.call("tr_tvm_dual", *dims, x.data_ptr(), scratch["p_m"].data_ptr(), 1.0, 1e30, int(self.l2_coupling), s)
        _lib.call("tr_tvm_primal", *dims, u.data_ptr(), ubar_scr.data_ptr(), g.data_ptr(), scratch["p_m"].data_ptr(),
                  -1.0, self._partial.data_ptr(), s)
        return self._reduce()[0]

    def pdhg_launch(self, it):
        """Launch one PDHG iteration (tvm.py:227-380); the float64 stopping sums
        (sum|du|^2, sum|u|^2) land in self._red."""
        dims = self.vol_dims()
        s = self.stream
        self.data_dual_update(self.ubar)
        _lib.call("tr_tvm_dual", *dims, self.ubar.data_ptr(), self.p_m.data_ptr(), float(self.sigma),
                  float(self.lamda), int(self.l2_coupling), s)
        self.data_adjoint(self.g)
        _lib.call("tr_tvm_primal", *dims, self.u.data_ptr(), self.ubar.data_ptr(), self.g.data_ptr(),
                  self.p_m.data_ptr(), float(self.tau), self._partial.data_ptr(), s)
        self._reduce_launch()
