"""Fixed-point objective of the TVM / TVME solvers, evaluated on the GPU in float64.

    Phi(u) = 1/2 sum_b sum_(samples, e, c) dcf |A_b u_b - y_b|^2 + R(u)

with R the motion TV (L2-coupled over echoes or L1) plus, for TVME, the
isotropic spatial TV of the echo differences.  The PDHG iterations of the
reference (``tvm.py:227-380``, ``tvme.py:152-480``) have this as their primal
objective: the data dual's fixed point is q* = A u* - y (``tvme.py:257-267``)
and the motion / echo duals are projections onto lambda-balls.

This is an evaluation utility (bench and tests compare final iterates of
different starts with it); the solvers never call it.  The forward operator
runs through the HIP library (``tr_data_simulate``); the sums are float64
reductions over device tensors.
"""

from __future__ import annotations

from . import _lib


def _fdiff(x, axis):
    import torch
    d = torch.zeros_like(x)
    n = x.shape[axis]
    if n > 1:
        d.narrow(axis, 0, n - 1).copy_(x.narrow(axis, 1, n - 1) - x.narrow(axis, 0, n - 1))
    return d


def regularizer(u, algorithm: str, lam_m: float, lam_e: float, l2_coupling: bool) -> float:
    """R(u) for u (B, E, Z, Y, X) complex64 on the GPU."""
    import torch
    u64 = u.to(torch.complex128)
    dm = _fdiff(u64, 0)
    if l2_coupling:
        reg = lam_m * float(torch.sum(torch.sqrt(torch.sum(torch.abs(dm) ** 2, dim=1))))
    else:
        reg = lam_m * float(torch.sum(torch.abs(dm)))
    del dm
    if algorithm == "tvme":
        de = _fdiff(u64, 1)
        gm = (torch.abs(_fdiff(de, -1)) ** 2 + torch.abs(_fdiff(de, -2)) ** 2
              + torch.abs(_fdiff(de, -3)) ** 2)
        reg += lam_e * float(torch.sum(torch.sqrt(gm)))
    return reg


def data_term(solver, u) -> float:
    """1/2 sum dcf |A u - y|^2 over all bins of a HIP solver (sorted sample order)."""
    import torch
    total = 0.0
    s = solver.stream
    for b in range(solver.B):
        if solver.y[b].numel() == 0:
            continue
        au = torch.empty_like(solver.y[b])
        ub = u[b].contiguous()  # held across the launch
        _lib.call("tr_data_simulate", solver.plan.handle, b, ub.data_ptr(),
                  solver.maps.data_ptr(), solver.E, solver.C, au.data_ptr(), s)
        r = (au - solver.y[b]).to(torch.complex128)
        total += 0.5 * float(torch.sum(solver.dcf[b].to(torch.float64)[:, None] * torch.abs(r) ** 2))
        del au, r
    return total


def objective(solver, u, algorithm: str = "tvme") -> float:
    lam_m = float(getattr(solver, "lamda_m", getattr(solver, "lamda", 0.0)))
    lam_e = float(getattr(solver, "lamda_e", 0.0))
    l2 = bool(getattr(solver, "l2_coupling", False))
    return data_term(solver, u) + regularizer(u, algorithm, lam_m, lam_e, l2)
