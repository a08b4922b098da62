"""Motion-TV + echo-difference isotropic spatial TV PDHG (TVME) on one MI355X.

Same constructor, state and result as ``TvmeReconstructor`` in
src/toporecon/algorithms/tvme.py:56-575 on a single process; the iteration
(tvme.py:208-482) runs as one fused forward/dual chain per motion bin, one
dual kernel for the motion and echo/spatial TV duals, one adjoint chain per
bin, and one primal kernel with the divergences, the three in-place updates in
the reference's order, the extrapolation and the float64 stopping sums
(tvme.py:520-557).
"""

from __future__ import annotations

import torch

from .. import _lib
from ._base import HipPdhgBase


class TvmeReconstructor(HipPdhgBase):
    """TVME PDHG reconstruction with the HIP NUFFT."""

    algorithm = "tvme"

    def __init__(self, ksp, coord, dcf, mps, resp, dual_q, B_local, l2_coupling=False, lamda_m=1e-5,
                 lamda_e=1e-5, sigma=0.01, tau=0.01, max_iter=300, tol=1e-3, margin=10, device=0,
                 E_total=None, e0=None, e1=None, echo_left_peer=None, echo_right_peer=None,
                 B_total=None, b0=None, b1=None, motion_left_peer=None, motion_right_peer=None,
                 bin_edges=None, world_comm=None, group_comm=None, leader_comm=None, nccl_group=None,
                 dist_grid=None, tagger=None, show_pbar=True, nufft_backend="hip", echoes=None):
        self.lamda_m = lamda_m
        self.lamda_e = lamda_e
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

    def pdhg_init(self, init=None):
        shape = self.vol_dims()
        dev = self.device
        self.ubar = torch.zeros(shape, dtype=torch.complex64, device=dev)
        self.u = torch.zeros_like(self.ubar)
        self.p_m = torch.zeros_like(self.ubar)
        self.p_ex = torch.zeros_like(self.ubar)
        self.p_ey = torch.zeros_like(self.ubar)
        self.p_ez = torch.zeros_like(self.ubar)
        self.g = torch.empty_like(self.ubar)
        if init is not None:
            self.ubar.copy_(init)
            self.u.copy_(init)
        self._alloc_reduction()

    def _tv_normal(self, x, u, ubar_scr, g, scratch):
        dims = self.vol_dims()
        s = self.stream
        for k in ("p_m", "p_ex", "p_ey", "p_ez"):
            if k in scratch:
                scratch[k].zero_()
            else:
                scratch[k] = torch.zeros_like(x)
        p = [scratch[k].data_ptr() for k in ("p_m", "p_ex", "p_ey", "p_ez")]
        _lib.call("tr_tvme_dual", *dims, x.data_ptr(), *p, 1.0, 1e30, 1e30, int(self.l2_coupling), s)
        _lib.call("tr_tvme_primal", *dims, u.data_ptr(), ubar_scr.data_ptr(), g.data_ptr(), *p, -1.0,
                  self._partial.data_ptr(), s)
        return self._reduce()[0]

    def pdhg_launch(self, it):
        """Launch one PDHG iteration (tvme.py:208-482); the float64 stopping sums
        (sum|du|^2, sum|u|^2) land in self._red."""
        dims = self.vol_dims()
        s = self.stream
        self.data_dual_update(self.ubar)
        _lib.call("tr_tvme_dual", *dims, self.ubar.data_ptr(), self.p_m.data_ptr(), self.p_ex.data_ptr(),
                  self.p_ey.data_ptr(), self.p_ez.data_ptr(), float(self.sigma), float(self.lamda_m),
                  float(self.lamda_e), int(self.l2_coupling), s)
        self.data_adjoint(self.g)
        _lib.call("tr_tvme_primal", *dims, self.u.data_ptr(), self.ubar.data_ptr(), self.g.data_ptr(),
                  self.p_m.data_ptr(), self.p_ex.data_ptr(), self.p_ey.data_ptr(), self.p_ez.data_ptr(),
                  float(self.tau), self._partial.data_ptr(), s)
        self._reduce_launch()
