"""Float64 restatement of the TVM / TVME / TVMW PDHG iterations (test infrastructure only).

One process, so every MPI peer is PROC_NULL (ghost slabs are zero,
src/toporecon/distributed/runtime.py:497-499), the coil all-reduce is a no-op
(tvm.py:188-190) and B, E are the global counts.  Each step function follows
the reference's order of operations:

* TVM  ``TvmReconstructor.pdhg``  src/toporecon/algorithms/tvm.py:227-380,
  stopping rule ``run`` tvm.py:406-471.
* TVME ``TvmeReconstructor.pdhg`` src/toporecon/algorithms/tvme.py:208-482,
  finite differences tvme.py:152-187, stopping rule tvme.py:512-557.
* TVMW ``TvmwReconstructor.pdhg`` src/toporecon/algorithms/tvmw.py:385-685,
  db1 echo wavelet tvmw.py:312-375, stopping rule tvmw.py:717-769.

Pinned by (tests/test_oracle_pdhg.py): adjoint identities <D u, p> = <u, D^T p>
for the motion, echo and spatial difference operators with the reference's
boundary rules and for the db1 echo wavelet; the prox maps are projections
(idempotent, |p| <= lambda); with lambda = 0 the iteration converges to the
weighted least-squares solution computed by a dense solve; the fixed-point
objective (``objective``) does not increase across late iterations.
"""

from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import List

import numpy as np

from . import nufft as onufft
from . import wavelet as owav


@dataclass
class BinnedProblem:
    """Inputs of one reconstructor after the reference's motion binning.

    ksp[b]: (E, C, M_b) complex; coord[b]: (M_b, 3) float32 [z, y, x] SigPy
    units; dcf[b]: (M_b,) float; mps: (C, Z, Y, X) complex.
    """

    ksp: List[np.ndarray]
    coord: List[np.ndarray]
    dcf: List[np.ndarray]
    mps: np.ndarray

    @property
    def B(self):
        return len(self.ksp)

    @property
    def E(self):
        return self.ksp[0].shape[0]

    @property
    def C(self):
        return self.mps.shape[0]

    @property
    def img_shape(self):
        return tuple(self.mps.shape[-3:])


def bin_problem(ksp, coord, dcf, mps, resp, bin_edges) -> BinnedProblem:
    """Motion binning of tvm.py:159-168: member iff edge_b <= r < edge_{b+1}."""
    ks, cs, ds = [], [], []
    for b in range(len(bin_edges) - 1):
        idx = (resp >= bin_edges[b]) & (resp < bin_edges[b + 1])
        ks.append(np.asarray(ksp)[:, :, idx].reshape(ksp.shape[0], ksp.shape[1], -1))
        cs.append(np.asarray(coord)[idx].reshape(-1, 3))
        ds.append(np.asarray(dcf)[idx].reshape(-1))
    return BinnedProblem(ks, cs, ds, np.asarray(mps))


# ----------------------------------------------------------------------------
# Data term
# ----------------------------------------------------------------------------

def forward_op(prob: BinnedProblem, u):
    """A u: per bin (E, C, M_b) = nufft(u[b, e] * S_c, coord_b)."""
    out = []
    for b in range(prob.B):
        img = u[b][:, None] * prob.mps[None]
        out.append(onufft.nufft(img, prob.coord[b]))
    return out


def adjoint_op(prob: BinnedProblem, q, weighted=True):
    """A^H W q: g[b, e] = sum_c conj(S_c) nufft_adjoint(dcf_b q[b][e, c])."""
    g = np.zeros((prob.B, prob.E) + prob.img_shape, dtype=np.complex128)
    for b in range(prob.B):
        qq = q[b] * (prob.dcf[b][None, None] if weighted else 1.0)
        x = onufft.nufft_adjoint(qq, prob.coord[b], prob.img_shape)  # (E, C, Z, Y, X)
        g[b] = np.sum(np.conj(prob.mps)[None] * x, axis=1)
    return g


def dual_data_update(prob, q, ubar, sigma):
    """q <- (q + sigma (A ubar - y)) / (1 + sigma)  (tvm.py:246-260)."""
    Au = forward_op(prob, ubar)
    return [(q[b] + sigma * (Au[b] - prob.ksp[b])) / (1.0 + sigma) for b in range(prob.B)]


# ----------------------------------------------------------------------------
# Finite differences with the reference's boundary rules (single process)
# ----------------------------------------------------------------------------

def motion_diff(u):
    """d[b] = u[b+1] - u[b], last bin 0 (tvm.py:262-269)."""
    d = np.zeros_like(u)
    if u.shape[0] > 1:
        d[:-1] = u[1:] - u[:-1]
    return d


def motion_div_tvm(p):
    """TVM 'divergence' = -D_m^T p (tvm.py:342-357)."""
    B = p.shape[0]
    div = np.zeros_like(p)
    if B > 1:
        div[1:] = p[1:] - p[:-1]
    div[0] = p[0]
    if B > 1:
        div[-1] = -p[-2]
    else:
        div[0] = 0
    return div


def motion_div_tvme(p):
    """TVME divp_m = +D_m^T p (tvme.py:421-439)."""
    B = p.shape[0]
    d = np.zeros_like(p)
    if B > 1:
        d[1:] = p[:-1] - p[1:]
    d[0] = -p[0]
    if B > 1:
        d[-1] = p[-2]
    else:
        d[0] = 0
    return d


def motion_div_tvmw(p):
    """TVMW divp_m (tvmw.py:661-671): no last-bin rule (p[B-1] stays 0)."""
    d = np.zeros_like(p)
    if p.shape[0] > 1:
        d[1:] = p[1:] - p[:-1]
    d[0] = p[0]
    return d


def echo_diff(u):
    """d_e[:, e] = u[:, e+1] - u[:, e], last echo 0 (tvme.py:233-237, 306-310)."""
    d = np.zeros_like(u)
    if u.shape[1] > 1:
        d[:, :-1] = u[:, 1:] - u[:, :-1]
    return d


def echo_div(divp):
    """divp_e (tvme.py:356-363, 443-454) = D_e^T divp."""
    E = divp.shape[1]
    d = np.zeros_like(divp)
    if E > 1:
        d[:, 1:] = divp[:, :-1] - divp[:, 1:]
    d[:, 0] = -divp[:, 0]
    if E > 1:
        d[:, -1] = divp[:, -2]
    return d


def fdiff(x, axis):
    """dxfc/dyfc/dzfc (tvme.py:152-168): forward difference, last index 0."""
    d = np.zeros_like(x)
    n = x.shape[axis]
    sl_lo = [slice(None)] * x.ndim
    sl_hi = [slice(None)] * x.ndim
    sl_lo[axis] = slice(0, n - 1)
    sl_hi[axis] = slice(1, n)
    d[tuple(sl_lo)] = x[tuple(sl_hi)] - x[tuple(sl_lo)]
    return d


def fdiff_adj(x, axis):
    """dxbc/dybc/dzbc (tvme.py:171-187) = adjoint of ``fdiff``."""
    n = x.shape[axis]
    tmp = np.zeros_like(x)
    sl = [slice(None)] * x.ndim
    sl[axis] = slice(0, n - 1)
    tmp[tuple(sl)] = -x[tuple(sl)]
    out = tmp.copy()
    sl_hi = [slice(None)] * x.ndim
    sl_hi[axis] = slice(1, n)
    out[tuple(sl_hi)] = -tmp[tuple(sl)] + tmp[tuple(sl_hi)]
    return out


def prox_motion(p, lam, l2_coupling):
    """Motion dual prox: L2 over echoes (tvm.py:276-306) or L1 (tvm.py:307-309)."""
    if l2_coupling:
        mag = np.sqrt(np.sum(np.abs(p) ** 2, axis=1, keepdims=True))
        return p / np.maximum(1.0, mag / lam)
    return p / np.maximum(1.0, np.abs(p) / lam)


def prox_iso3(px, py, pz, lam):
    """Isotropic TV prox on the three spatial duals (tvme.py:318-323)."""
    mag = np.sqrt(np.abs(px) ** 2 + np.abs(py) ** 2 + np.abs(pz) ** 2)
    s = np.maximum(1.0, mag / lam)
    return px / s, py / s, pz / s


# ----------------------------------------------------------------------------
# State and iterations
# ----------------------------------------------------------------------------

@dataclass
class State:
    ubar: np.ndarray
    u: np.ndarray
    q: list
    p_m: np.ndarray
    extra: dict = field(default_factory=dict)


def init_state(prob: BinnedProblem, algorithm: str) -> State:
    shape = (prob.B, prob.E) + prob.img_shape
    z = np.zeros(shape, dtype=np.complex128)
    st = State(z.copy(), z.copy(), [np.zeros_like(k, dtype=np.complex128) for k in prob.ksp], z.copy())
    if algorithm == "tvme":
        st.extra = {"p_ex": z.copy(), "p_ey": z.copy(), "p_ez": z.copy()}
    elif algorithm == "tvmw":
        st.extra = {"p_w1": z.copy(), "p_w2": None}
    return st


def tvm_step(prob, st: State, lam, sigma, tau, l2_coupling):
    """One TvmReconstructor.pdhg call (tvm.py:227-380)."""
    u_old = st.u.copy()
    q = dual_data_update(prob, st.q, st.ubar, sigma)
    p_m = st.p_m + sigma * motion_diff(st.ubar)
    p_m = prox_motion(p_m, lam, l2_coupling)
    g = adjoint_op(prob, q)
    div = motion_div_tvm(p_m)
    u = st.u - tau * (g - div)
    ubar = 2 * u - u_old
    return State(ubar, u, q, p_m, {})


def tvme_step(prob, st: State, lam_m, lam_e, sigma, tau, l2_coupling):
    """One TvmeReconstructor.pdhg call (tvme.py:208-482)."""
    u_old = st.u.copy()
    d_e = echo_diff(st.ubar)
    q = dual_data_update(prob, st.q, st.ubar, sigma)
    p_m = st.p_m + sigma * motion_diff(st.ubar)
    p_ex = st.extra["p_ex"] + sigma * fdiff(d_e, -1)
    p_ey = st.extra["p_ey"] + sigma * fdiff(d_e, -2)
    p_ez = st.extra["p_ez"] + sigma * fdiff(d_e, -3)
    p_ex, p_ey, p_ez = prox_iso3(p_ex, p_ey, p_ez, lam_e)
    p_m = prox_motion(p_m, lam_m, l2_coupling)
    divp = fdiff_adj(p_ex, -1) + fdiff_adj(p_ey, -2) + fdiff_adj(p_ez, -3)
    divp_e = echo_div(divp)
    g = adjoint_op(prob, q)
    divp_m = motion_div_tvme(p_m)
    u = st.u.copy()
    u -= tau * divp_m
    u -= tau * divp_e
    u -= tau * g
    ubar = 2 * u - u_old
    return State(ubar, u, q, p_m, {"p_ex": p_ex, "p_ey": p_ey, "p_ez": p_ez})


def w1_fwd(u):
    """db1 echo wavelet, global echo index (tvmw.py:312-342), one process."""
    E = u.shape[1]
    c = np.zeros_like(u)
    s = 1.0 / math.sqrt(2.0)
    for j in range(E):
        if j % 2 == 0:
            x1 = u[:, j + 1] if j + 1 < E else 0.0
            c[:, j] = (u[:, j] + x1) * s
        else:
            c[:, j] = (-u[:, j - 1] + u[:, j]) * s
    return c


def w1_adj(p):
    """Adjoint db1 echo wavelet (tvmw.py:344-375), one process."""
    E = p.shape[1]
    out = np.zeros_like(p)
    s = 1.0 / math.sqrt(2.0)
    for j in range(E):
        if j % 2 == 0:
            dn = p[:, j + 1] if j + 1 < E else 0.0
            out[:, j] = (p[:, j] - dn) * s
        else:
            out[:, j] = (p[:, j - 1] + p[:, j]) * s
    return out


def tvmw_step(prob, st: State, lam1, lam2, lam3, sigma, tau, l2_coupling):
    """One TvmwReconstructor.pdhg call (tvmw.py:385-685)."""
    ubar = st.ubar
    u_old = st.u.copy()
    B = prob.B
    p_w2 = st.extra["p_w2"]
    if p_w2 is None:
        p_w2 = [owav.zeros_like_coeffs(owav.wavedec3(ubar[b])) for b in range(B)]
    new_w2 = []
    for b in range(B):
        c = owav.wavedec3(ubar[b])
        t = owav.tree_axpy(p_w2[b], sigma, c)
        new_w2.append(owav.tree_project_linf(t, lam3))
    q = dual_data_update(prob, st.q, ubar, sigma)
    p_m = st.p_m + sigma * motion_diff(ubar)
    p_m = prox_motion(p_m, lam1, l2_coupling)
    p_w1 = st.extra["p_w1"] + sigma * w1_fwd(ubar)
    if lam2 > 0:
        p_w1 = p_w1 / np.maximum(1.0, np.abs(p_w1) / lam2)
    else:
        p_w1 = np.zeros_like(p_w1)
    g = adjoint_op(prob, q)
    divp_w2 = np.stack([owav.waverec3(new_w2[b], prob.img_shape) for b in range(B)])
    divp_w1 = w1_adj(p_w1)
    divp_m = motion_div_tvmw(p_m)
    u = st.u - tau * (g - divp_m + divp_w1 + divp_w2)
    ubar_new = 2 * u - u_old
    return State(ubar_new, u, q, p_m, {"p_w1": p_w1, "p_w2": new_w2})


def rel_change_tvm(u_new, u_old):
    """TVM/TVME stopping value (tvm.py:416-463): ||du|| / ||u||, inf if undefined."""
    num = float(np.vdot(u_new - u_old, u_new - u_old).real)
    den = float(np.vdot(u_new, u_new).real)
    if np.isfinite(num) and np.isfinite(den) and den > 0.0:
        return math.sqrt(max(num, 0.0)) / math.sqrt(den)
    return float("inf")


def rel_change_tvmw(ubar_prev, ubar_new):
    """TVMW stopping value (tvmw.py:723-769): ||ubar_prev - ubar|| / max(||ubar_prev||, 1e-12)."""
    d = ubar_prev - ubar_new
    num = float(np.vdot(d, d).real)
    den = float(np.vdot(ubar_prev, ubar_prev).real)
    if np.isfinite(num) and np.isfinite(den):
        return math.sqrt(max(num, 0.0)) / max(math.sqrt(max(den, 0.0)), 1e-12)
    return float("inf")


def run(prob: BinnedProblem, algorithm: str, *, lam_m=1e-5, lam_e=1e-5, lam1=None, lam2=0.0,
        lam3=0.0, sigma=1 / 6, tau=1 / 6, l2_coupling=False, max_iter=100, tol=1e-3,
        init: State | None = None, history=None):
    """Run a solver to its stopping rule; returns (ubar, iterations run, state)."""
    st = init if init is not None else init_state(prob, algorithm)
    iters = 0
    for it in range(max_iter):
        prev_u = st.u
        prev_ubar = st.ubar
        if algorithm == "tvm":
            st = tvm_step(prob, st, lam_m, sigma, tau, l2_coupling)
            t = rel_change_tvm(st.u, prev_u)
        elif algorithm == "tvme":
            st = tvme_step(prob, st, lam_m, lam_e, sigma, tau, l2_coupling)
            t = rel_change_tvm(st.u, prev_u)
        elif algorithm == "tvmw":
            st = tvmw_step(prob, st, lam_m if lam1 is None else lam1, lam2, lam3, sigma, tau,
                           l2_coupling)
            t = rel_change_tvmw(prev_ubar, st.ubar)
        else:
            raise ValueError(algorithm)
        iters = it + 1
        if history is not None:
            history.append(t)
        if t < tol:
            break
    return st.ubar, iters, st


def objective(prob: BinnedProblem, u, algorithm: str, *, lam_m=1e-5, lam_e=1e-5, lam1=None,
              lam2=0.0, lam3=0.0, l2_coupling=False):
    """Fixed-point objective Phi(u) = 1/2 sum ||W^1/2 (A u - y)||^2 + R(u).

    Derived from the PDHG fixed point (q* = A u* - y, 0 = A^H W q* + R^T p*).
    Not in the reference; used to compare final iterates.
    """
    Au = forward_op(prob, u)
    data = 0.0
    for b in range(prob.B):
        r = Au[b] - prob.ksp[b]
        data += 0.5 * float(np.sum(prob.dcf[b][None, None] * np.abs(r) ** 2))
    lm = lam_m if lam1 is None else lam1
    dm = motion_diff(u)
    if l2_coupling:
        reg = lm * float(np.sum(np.sqrt(np.sum(np.abs(dm) ** 2, axis=1))))
    else:
        reg = lm * float(np.sum(np.abs(dm)))
    if algorithm == "tvme":
        de = echo_diff(u)
        gm = np.sqrt(np.abs(fdiff(de, -1)) ** 2 + np.abs(fdiff(de, -2)) ** 2
                     + np.abs(fdiff(de, -3)) ** 2)
        reg += lam_e * float(np.sum(gm))
    elif algorithm == "tvmw":
        reg += lam2 * float(np.sum(np.abs(w1_fwd(u))))
        for b in range(u.shape[0]):
            for arr in owav.tree_leaves(owav.wavedec3(u[b])):
                reg += lam3 * float(np.sum(np.sqrt(np.abs(arr) ** 2 + 1e-12)))
    return data + reg
