"""Shared single-GPU machinery of the HIP PDHG reconstructors.

Mirrors the constructor work common to TvmReconstructor (tvm.py:66-178),
TvmeReconstructor (tvme.py:57-149) and TvmwReconstructor (tvmw.py:214-310):
motion binning with ``edge_b <= r < edge_{b+1}`` (tvm.py:147-168), per-bin
k-space / trajectory / DCF / dual slices, and the NUFFT plan.  On the device,
each bin's k-space ``y`` and data dual ``q`` live as (M_b, E*C) complex64 in the
plan's sorted sample order, so one launch serves every echo and coil.

One process drives one GPU: the MPI / NCCL arguments of the reference are
accepted for signature compatibility, but a communicator with more than one
rank raises ``NotImplementedError``.
"""

from __future__ import annotations

import logging
import math
import time

import numpy as np

from .. import _lib
from ..nufft_backend import NufftPlan, resolve_device


def _comm_size(comm):
    if comm is None:
        return 1
    for attr in ("Get_size", "size"):
        v = getattr(comm, attr, None)
        if v is not None:
            return int(v() if callable(v) else v)
    return 1


class HipPdhgBase:
    """Binned data, plan and run loop shared by the TVM / TVME / TVMW HIP solvers."""

    algorithm = "base"

    def _setup(self, ksp, coord, dcf, mps, resp, dual_q, B_local, *, margin, device, B_total, b0,
               b1, bin_edges, comms, show_pbar, nufft_backend, sigma, tau, max_iter, tol,
               echoes=None):
        import torch
        if str(nufft_backend).strip().lower() != "hip":
            raise ValueError(f"the HIP reconstructors run the 'hip' NUFFT backend, got {nufft_backend!r}")
        for c in comms:
            if _comm_size(c) > 1:
                raise NotImplementedError("the HIP reconstructor drives one GPU per process; "
                                          "multi-rank communicators are not supported")
        self.device = resolve_device(device)
        self.B = int(B_local)
        self.B_total = int(B_total) if B_total is not None else self.B
        self.b0 = int(b0) if b0 is not None else 0
        self.b1 = int(b1) if b1 is not None else self.b0 + self.B
        if self.b0 != 0 or self.b1 != self.B_total or self.b1 - self.b0 != self.B:
            raise NotImplementedError("motion-axis sharding needs halo exchanges; one GPU owns all bins")
        self.sigma = sigma
        self.tau = tau
        self.max_iter = int(max_iter)
        self.tol = tol
        self.show_pbar = show_pbar
        self.C = int(mps.shape[0])
        if ksp is None and echoes is None:
            raise ValueError("ksp is required (or echoes= for a solver whose k-space is simulated)")
        self.E = int(ksp.shape[0]) if ksp is not None else int(echoes)
        self.K = self.E * self.C
        self.img_shape = [int(v) for v in mps.shape[-3:]]

        resp = np.asarray(resp if not isinstance(resp, torch.Tensor) else resp.cpu().numpy()).reshape(-1)
        if bin_edges is None:
            bins = np.percentile(resp, np.linspace(0 + margin, 100 - margin, self.B_total + 1))
        else:
            bins = np.asarray(bin_edges)
            assert bins.ndim == 1 and bins.size == self.B_total + 1, \
                f"bin_edges must have shape ({self.B_total + 1},), got {bins.shape}"
        self.bin_edges = bins

        dev = self.device
        self.maps = self._tensor(mps, torch.complex64)
        coords, self.y, self.q, self.dcf, self.q0 = [], [], [], [], []
        members = []
        for gb in range(self.b0, self.b1):
            idx = np.nonzero((resp >= bins[gb]) & (resp < bins[gb + 1]))[0]
            members.append(idx)
            c = self._take(coord, idx, axis=0, dtype=torch.float32).reshape(-1, 3)
            coords.append(c)
        self.members = members
        self.plan = NufftPlan(self.img_shape, coords, dev)
        stream = _lib.stream_handle(dev)
        for b, idx in enumerate(members):
            M = self.plan.samples(b)
            if ksp is None:
                y_s = torch.zeros((M, self.K), dtype=torch.complex64, device=dev)
            else:
                k_b = self._take(ksp, idx, axis=2, dtype=torch.complex64).reshape(self.K, M)
                y_s = torch.empty((M, self.K), dtype=torch.complex64, device=dev)
                _lib.call("tr_plan_gather_samples", self.plan.handle, b, k_b.data_ptr(), self.K,
                          y_s.data_ptr(), stream)
                del k_b
            q_s = torch.zeros((M, self.K), dtype=torch.complex64, device=dev)
            q0 = None
            if dual_q is not None and np.any(np.asarray(dual_q) if not isinstance(dual_q, torch.Tensor)
                                             else bool(torch.any(dual_q != 0))):
                q_b = self._take(dual_q, idx, axis=2, dtype=torch.complex64).reshape(self.K, M)
                _lib.call("tr_plan_gather_samples", self.plan.handle, b, q_b.data_ptr(), self.K,
                          q_s.data_ptr(), stream)
                del q_b
                q0 = q_s.clone()
            self.q0.append(q0)
            d_b = self._take(dcf, idx, axis=0, dtype=torch.float32).reshape(M)
            d_s = torch.empty((M,), dtype=torch.float32, device=dev)
            _lib.call("tr_plan_gather_weights", self.plan.handle, b, d_b.data_ptr(), d_s.data_ptr(),
                      stream)
            self.y.append(y_s)
            self.q.append(q_s)
            self.dcf.append(d_s)
        self.iterations = 0
        self.elapsed = 0.0
        self.history = []

    # -- helpers ----------------------------------------------------------------
    def _tensor(self, a, dtype):
        import torch
        if isinstance(a, torch.Tensor):
            return a.to(device=self.device, dtype=dtype).contiguous()
        npd = np.complex64 if dtype == torch.complex64 else np.float32
        return torch.from_numpy(np.ascontiguousarray(np.asarray(a), dtype=npd)).to(self.device)

    def _take(self, a, idx, axis, dtype):
        import torch
        if isinstance(a, torch.Tensor):
            it = torch.as_tensor(idx, device=a.device, dtype=torch.long)
            return a.index_select(axis, it).to(device=self.device, dtype=dtype).contiguous()
        return self._tensor(np.take(np.asarray(a), idx, axis=axis), dtype)

    @property
    def stream(self):
        return _lib.stream_handle(self.device)

    def vol_dims(self):
        Z, Y, X = self.img_shape
        return (self.B, self.E, Z, Y, X)

    def _alloc_reduction(self):
        import torch
        B, E, Z, Y, X = self.vol_dims()
        n = int(_lib.load().tr_partial_count(B, Z, Y, X))
        self._npartial = n
        self._partial = torch.empty((n, 2), dtype=torch.float64, device=self.device)
        self._red = torch.empty((2,), dtype=torch.float64, device=self.device)

    def _reduce_launch(self):
        _lib.call("tr_reduce_partials", self._partial.data_ptr(), self._npartial, self._red.data_ptr(),
                  self.stream)

    def _reduce(self):
        self._reduce_launch()
        num, den = self._red.tolist()  # device -> host, synchronises the stream
        return num, den

    def pdhg(self, it):
        """One PDHG iteration; returns the float64 stopping sums (num, den).  The launches
        (pdhg_launch) contain no host synchronisation, allocation or memset once the
        workspaces exist, so one iteration can be captured into a HIP graph."""
        self.pdhg_launch(it)
        num, den = self._red.tolist()  # device -> host, synchronises the stream
        return num, den

    # -- data term -----------------------------------------------------------------
    def reset_dual(self):
        """Data duals back to their constructor value (dual_q, or 0), as at the start of run()."""
        for b in range(self.B):
            if self.q0[b] is None:
                self.q[b].zero_()
            else:
                self.q[b].copy_(self.q0[b])

    def set_bin_streams(self, n: int):
        """Issue the data term of bin b on stream b % n with its own grid workspace (n = 1: one
        stream).  Results are bit-identical for every n; n > 1 lets bins overlap on the GPU."""
        import torch
        n = max(1, min(int(n), 8, max(self.B, 1)))
        self.plan.set_param("workspace_slots", n)
        self._bstreams = [torch.cuda.Stream(self.device) for _ in range(n)] if n > 1 else None

    def _bin_streams(self):
        """Per-bin stream handles, forked from the current stream (None: use the current one)."""
        import torch
        bs = getattr(self, "_bstreams", None)
        if not bs:
            return None
        main = torch.cuda.current_stream(self.device)
        for st in bs:
            st.wait_stream(main)
        return [bs[b % len(bs)].cuda_stream for b in range(self.B)]

    def _join_bin_streams(self):
        import torch
        bs = getattr(self, "_bstreams", None)
        if bs:
            main = torch.cuda.current_stream(self.device)
            for st in bs:
                main.wait_stream(st)

    def data_dual_update(self, ubar):
        """q_b <- (q_b + sigma (A ubar_b - y_b)) / (1 + sigma) for every bin (one launch chain per bin)."""
        s = self.stream
        hs = self._bin_streams()
        sig = float(self.sigma)
        ops = float(np.float32(1.0 + self.sigma))
        for b in range(self.B):
            _lib.call("tr_data_dual_update", self.plan.handle, b, ubar[b].data_ptr(), self.maps.data_ptr(),
                      self.E, self.C, self.y[b].data_ptr(), self.q[b].data_ptr(), sig, ops,
                      hs[b] if hs else s)
        self._join_bin_streams()

    def data_adjoint(self, g):
        """g_b = sum_c conj(S_c) NUFFT_b^H(dcf_b q_b,c) for every bin."""
        s = self.stream
        hs = self._bin_streams()
        for b in range(self.B):
            _lib.call("tr_data_adjoint", self.plan.handle, b, self.q[b].data_ptr(), self.dcf[b].data_ptr(),
                      self.maps.data_ptr(), self.E, self.C, g[b].data_ptr(), 0, hs[b] if hs else s)
        self._join_bin_streams()

    def simulate_kspace(self, truth, snr_db=30.0, seed=0, normalize=True):
        """Replace y by A truth + complex Gaussian noise at snr_db, then y /= max|y|."""
        import torch
        x = self._tensor(truth, torch.complex64)
        s = self.stream
        ss, cnt = 0.0, 0
        for b in range(self.B):
            _lib.call("tr_data_simulate", self.plan.handle, b, x[b].data_ptr(), self.maps.data_ptr(),
                      self.E, self.C, self.y[b].data_ptr(), s)
            ss += float(torch.sum(torch.abs(self.y[b]) ** 2))
            cnt += self.y[b].numel()
        rms = math.sqrt(ss / max(cnt, 1))
        sig = rms * 10.0 ** (-snr_db / 20.0) / math.sqrt(2.0)
        gen = torch.Generator(device=self.device)
        gen.manual_seed(int(seed))
        mx = 0.0
        K = self.E * self.C
        for b in range(self.B):
            M = self.y[b].shape[0]
            # noise drawn per (channel, sample) in trajectory order and moved to the sorted layout,
            # so the realisation does not depend on the plan's internal sample order
            raw = torch.randn((K, M, 2), generator=gen, device=self.device, dtype=torch.float32)
            noise = torch.empty_like(self.y[b])
            if M:
                _lib.call("tr_plan_gather_samples", self.plan.handle, b, torch.view_as_complex(raw).data_ptr(), K,
                          noise.data_ptr(), s)
                self.y[b].add_(noise, alpha=sig)
                mx = max(mx, float(torch.max(torch.abs(self.y[b]))))
            del raw, noise
        if normalize and mx > 0:
            for b in range(self.B):
                self.y[b].div_(np.float32(mx))
        return mx

    def set_geometry(self, sort_tile: int = 8, spread_segment: int = 1024):
        """Rebuild the NUFFT plan with another sample geometry (sort tile edge, heavy-cluster
        segment length) and move the sorted per-sample data (k-space, data duals, DCF) to its
        sample order; launch parameters carry over."""
        import torch
        from ..nufft_backend import NufftPlan
        old = self.plan
        if (old.get_param("sort_tile"), old.get_param("spread_segment")) == (int(sort_tile), int(spread_segment)):
            return
        new = NufftPlan(self.img_shape, old._coords, self.device, sort_tile=sort_tile,
                        spread_segment=spread_segment)
        for k in NufftPlan.LAUNCH_PARAMS:
            new.set_param(k, old.get_param(k))
        s = self.stream
        with torch.cuda.device(self.device):
            for b in range(self.B):
                M = self.y[b].shape[0]
                if M == 0:
                    continue
                orig = torch.empty((self.K, M), dtype=torch.complex64, device=self.device)
                for t in (self.y[b], self.q[b], self.q0[b]):
                    if t is None:
                        continue
                    _lib.call("tr_plan_scatter_samples", old.handle, b, t.data_ptr(), self.K, orig.data_ptr(), s)
                    _lib.call("tr_plan_gather_samples", new.handle, b, orig.data_ptr(), self.K, t.data_ptr(), s)
                d = torch.empty_like(self.dcf[b])
                _lib.call("tr_plan_scatter_weights", old.handle, b, self.dcf[b].data_ptr(), d.data_ptr(), s)
                _lib.call("tr_plan_gather_weights", new.handle, b, d.data_ptr(), self.dcf[b].data_ptr(), s)
                del orig, d
            torch.cuda.synchronize(self.device)
        old.close()
        self.plan = new

    def autotune(self, reps: int = 2, cache_path=None, force: bool = False, bin_streams: bool = True,
                 workspace_budget: float = 64 * 2.0 ** 30, geometry: bool = False) -> dict:
        """Pick the plan's gridding launch parameters by timing them on the largest bin, then the
        number of bin streams (each with its own grid workspace, within workspace_budget bytes)
        by timing the adjoint data term.

        Only the workgroup order and stream assignment change, so run() gives bit-identical
        results whatever is chosen.  With geometry=True the sample geometry (gridding sort tile
        and heavy-cluster segment length) is tuned first on plans of the largest bin and the
        solver's plan rebuilt with it (set_geometry); that choice changes float32 rounding.
        """
        import time as _time

        import torch
        from .. import autotune as _autotune
        if self.B == 0:
            return {}
        b = max(range(self.B), key=self.plan.samples)
        if self.plan.samples(b) == 0:
            return {}
        img = torch.zeros((self.E, *self.img_shape), dtype=torch.complex64, device=self.device)
        ksp = torch.zeros_like(self.y[b])
        out = _autotune.tune(self.plan, image=img, maps=self.maps, E=self.E, C=self.C, ksp=ksp,
                             dcf=self.dcf[b], stream=self.stream, bins=(b,), reps=reps,
                             cache_path=cache_path, force=force)
        if geometry:
            launch = {k: self.plan.get_param(k) for k in ("interp_xcd_chunk", "interp_kgroups",
                                                          "spread_xcd_chunk")}
            geo = _autotune.tune_geometry(self.img_shape, self.plan._coords[b], image=img, maps=self.maps,
                                          E=self.E, C=self.C, dcf=self.dcf[b], device=self.device,
                                          stream=self.stream, reps=reps, launch=launch)
            self.set_geometry(geo["sort_tile"], geo["spread_segment"])
            out = dict(out, sort_tile=geo["sort_tile"], spread_segment=geo["spread_segment"],
                       timings_ms=dict(out.get("timings_ms", {}), geometry=geo["timings_ms"]))
        del img, ksp
        if bin_streams and self.B > 1:
            ws = float(np.prod(self.plan.grid_shape)) * self.K * 8.0
            cands = sorted({n for n in (1, 2, 3, self.B) if 1 <= n <= min(self.B, 8) and n * ws <= workspace_budget})
            g = torch.empty(self.vol_dims(), dtype=torch.complex64, device=self.device)
            times = {}
            for n in cands:
                self.set_bin_streams(n)
                self.data_adjoint(g)
                torch.cuda.synchronize(self.device)
                t0 = _time.perf_counter()
                for _ in range(max(1, reps)):
                    self.data_adjoint(g)
                torch.cuda.synchronize(self.device)
                times[n] = (_time.perf_counter() - t0) / max(1, reps) * 1e3
            best = min(times, key=times.get)
            self.set_bin_streams(best)
            if best < max(cands):
                self.plan.release_workspace()  # slots beyond `best` are no longer used
            del g
            out = dict(out, bin_streams=best, timings_ms=dict(out.get("timings_ms", {}),
                                                              bin_streams={str(k): round(v, 3) for k, v in times.items()}))
        return out

    # -- step-size check ------------------------------------------------------------
    def _tv_normal(self, x, u, ubar_scr, g, scratch) -> float:
        """u <- K^H K x from u = 0, given g = A^H W A x: the TV part of K through the
        algorithm's dual step (sigma = 1, no projection, duals from zero) and its adjoint
        through the primal step with tau = -1.  Returns sum |u|^2 (float64)."""
        raise NotImplementedError(f"{type(self).__name__} has no power iteration")

    def operator_norm(self, iters: int = 20, seed: int = 0) -> dict:
        """Power-iteration estimate of ||K|| for the stacked operator K = [W^1/2 A; TV parts]
        whose adjoint the primal step applies.  The normal operator K^H K x is built from the
        solver's own kernels (the data term through the NUFFT, the TV part through the dual step
        with sigma = 1 and no projection followed by the primal step with tau = -1 from zero),
        and every norm is a float64 reduction on the GPU.  Returns ||K||, sigma tau ||K||^2
        (PDHG needs < 1) and the per-iteration estimates."""
        import torch
        dims = self.vol_dims()
        dev = self.device
        s = self.stream
        n = int(np.prod(dims))
        gen = torch.Generator(device=dev)
        gen.manual_seed(int(seed))
        with torch.cuda.device(dev):
            x = torch.complex(torch.randn(dims, generator=gen, device=dev),
                              torch.randn(dims, generator=gen, device=dev))
            u = torch.zeros_like(x)
            ubar_scr = torch.empty_like(x)
            g = torch.empty_like(x)
            qtmp = [torch.empty_like(self.y[b]) for b in range(self.B)]
            self._alloc_reduction()
            scratch = {}
            # normalise the start vector (||x|| from the same float64 reduction)
            nrm = math.sqrt(self._sumsq(x))
            _lib.call("tr_cscale", x.data_ptr(), n, 1.0 / nrm, x.data_ptr(), s)
            hist = []
            for _ in range(iters):
                for b in range(self.B):
                    _lib.call("tr_data_simulate", self.plan.handle, b, x[b].data_ptr(), self.maps.data_ptr(),
                              self.E, self.C, qtmp[b].data_ptr(), s)
                    _lib.call("tr_data_adjoint", self.plan.handle, b, qtmp[b].data_ptr(), self.dcf[b].data_ptr(),
                              self.maps.data_ptr(), self.E, self.C, g[b].data_ptr(), 0, s)
                u.zero_()
                num = self._tv_normal(x, u, ubar_scr, g, scratch)  # u = K^H K x, num = ||u||^2
                lam = math.sqrt(max(num, 0.0))  # ||K^H K x|| with ||x|| = 1
                hist.append(lam)
                if lam == 0.0:
                    break
                _lib.call("tr_cscale", u.data_ptr(), n, 1.0 / lam, x.data_ptr(), s)
            del qtmp, g, u, ubar_scr, x, scratch
        k2 = hist[-1] if hist else 0.0
        return {"K_norm": math.sqrt(k2), "K_norm_sq": k2, "sigma_tau_K_norm_sq": self.sigma * self.tau * k2,
                "iterations": len(hist), "history_K_norm_sq": hist}

    def _sumsq(self, x):
        """sum |x|^2 of a (B, E, Z, Y, X) array, float64 on the GPU (per-image sums added on the host)."""
        import torch
        nimg = self.B * self.E
        out = torch.empty(nimg, dtype=torch.float64, device=self.device)
        _lib.call("tr_cnn_sumsq", x.data_ptr(), int(np.prod(self.img_shape)), nimg, out.data_ptr(), self.stream)
        return float(out.sum())

    # -- run loop ----------------------------------------------------------------
    def _converged_value(self, num, den):
        if math.isfinite(num) and math.isfinite(den) and den > 0.0:
            return math.sqrt(max(num, 0.0)) / math.sqrt(den)
        return float("inf")

    def init_dual_residual(self, x):
        """q_b = A x_b - y_b: the data dual's fixed point for the primal image x."""
        s = self.stream
        for b in range(self.B):
            _lib.call("tr_data_residual", self.plan.handle, b, x[b].data_ptr(), self.maps.data_ptr(), self.E,
                      self.C, self.y[b].data_ptr(), self.q[b].data_ptr(), s)

    def run(self, init=None, progress_every=0, dual_init=None):
        """PDHG to the stopping rule; returns ubar (B, E, Z, Y, X) complex64 on the GPU.

        init: starting image (B, E, Z, Y, X) instead of zeros (the reference
        always starts from zeros).  dual_init="residual" with an init also
        starts the data dual at its fixed point A init - y; a sequence of B
        arrays ((M_b, E*C) sorted, e.g. from warm_start(..., with_dual=True))
        starts it there; otherwise the data dual starts at the constructor's
        dual_q (or zero).  progress_every > 0 prints the iteration and
        stopping value to stderr every that many iterations.
        """
        import sys
        import torch
        given = isinstance(dual_init, (list, tuple))
        if not given and dual_init not in (None, "residual"):
            raise ValueError(f"unknown dual_init {dual_init!r}")
        if given and len(dual_init) != self.B:
            raise ValueError(f"dual_init has {len(dual_init)} bins, the solver {self.B}")
        with torch.cuda.device(self.device):
            self.pdhg_init(init)
            if given:
                for b in range(self.B):
                    self.q[b].copy_(dual_init[b])
            elif init is not None and dual_init == "residual":
                self.init_dual_residual(self.ubar)
            else:
                self.reset_dual()
            torch.cuda.synchronize(self.device)
            t0 = time.perf_counter()
            self.history = []
            it = -1
            for it in range(self.max_iter):
                num, den = self.pdhg(it)
                tol = self._converged_value(num, den)
                self.history.append(tol)
                if progress_every and (it + 1) % progress_every == 0:
                    print(f"[pdhg] it={it + 1} tol={tol:.3e} t={time.perf_counter() - t0:.1f}s",
                          file=sys.stderr, flush=True)
                if tol < self.tol:
                    logging.info("Converged in %s iterations.", it)
                    break
            torch.cuda.synchronize(self.device)
            self.elapsed = time.perf_counter() - t0
            self.iterations = it + 1
            if self.show_pbar:
                logging.info("[PDHG iterations only] %.2f s (max over ranks)", self.elapsed)
        return self.ubar
