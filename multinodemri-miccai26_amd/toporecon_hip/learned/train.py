"""Train the warm-start network on the GPU from pairs the plain solver generates.

For every training seed the synthetic TVME acquisition generator (cone
trajectory, ellipsoid phantom with R2* / off-resonance decay, Gaussian-blob
coil maps, 30 dB noise) makes one problem; the plain HIP PDHG solves it to
the tolerance, and each (bin, echo) image contributes one pair

    input  x = g / s,            g = A^H (dcf * y), s = RMS(g)
    target t = u* / (alpha s),   alpha = Re<g, u*> / ||g||^2

(the least-squares scale alpha is removed so the network learns the shape;
at inference the scale is fitted to the data).  The network predicts
t - x (residual).  Pairs stay in HBM; training runs forward, loss, backward
and Adam through the library's MFMA kernels.  Held-out seeds are then solved
plainly and from the warm start to the same tolerance, and iterations,
time-to-tolerance and the final objective are reported.

    python -m toporecon_hip.learned.train --out weights.npz
"""

from __future__ import annotations

import argparse
import json
import math
import sys
import time

import numpy as np

from .. import _lib, synth
from .net import WarmStartNet, default_weights_path
from .warm import cnn_warm_start, dc_adjoint, image_rms

TRAIN_SHAPE = (64, 96, 96)


def training_config(seed, shape=TRAIN_SHAPE, coils=16, echoes=4, bins=6, interleaves=3000, readouts=384):
    cfg = synth.config_scaled(synth.config_full_tvme(), shape, coils=coils, echoes=echoes, bins=bins,
                              interleaves=interleaves, readouts=readouts, seed=seed)
    cfg.max_iter = 300
    return cfg


def make_solver(cfg, device):
    from ..algorithms import TvmeReconstructor
    acq = synth.make_acquisition(cfg)
    E = len(cfg.echo_times_ms)
    solver = TvmeReconstructor(None, acq.coord, acq.dcf, acq.mps, acq.resp, None, cfg.bins, cfg.l2_coupling,
                               lamda_m=cfg.lambda_motion, lamda_e=cfg.lambda_echo, sigma=1 / 6, tau=1 / 6,
                               max_iter=cfg.max_iter, tol=cfg.tol, margin=0, device=device,
                               bin_edges=acq.bin_edges, show_pbar=False, echoes=E)
    solver.simulate_kspace(acq.truth, cfg.snr_db, seed=cfg.seed + 104729)
    solver.autotune(reps=1)
    return solver


def make_pairs(solver):
    """(x bf16 (n, Z, Y, X, 8), residual target float32 (n, Z, Y, X, 2), plain iterations)."""
    import torch
    ustar = solver.run()
    its = solver.iterations
    shape = tuple(solver.img_shape)
    n = solver.B * solver.E
    V = int(np.prod(shape))
    g = dc_adjoint(solver).reshape(n, *shape)
    u = ustar.reshape(n, *shape)
    s = image_rms(g, n, V, solver.device, solver.stream)
    x = torch.empty((n, *shape, 8), dtype=torch.bfloat16, device=solver.device)
    _lib.call("tr_cnn_pack_input", g.data_ptr(), s.data_ptr(), V, n, x.data_ptr(), solver.stream)
    # least-squares scale per image and the residual target (host-side bookkeeping on device tensors)
    num = torch.sum((torch.conj(g) * u).real.reshape(n, -1).to(torch.float64), dim=1)
    den = torch.sum((torch.abs(g) ** 2).reshape(n, -1).to(torch.float64), dim=1)
    alpha = torch.where(den > 0, num / torch.where(den > 0, den, torch.ones_like(den)), torch.ones_like(den))
    t = u / (alpha.to(torch.float32) * s)[:, None, None, None]
    target = torch.stack([t.real, t.imag], dim=-1) - x[..., :2].float()
    return x, target.contiguous(), its


def train(net, xs, ts, shape, steps, lr, seed=0, log_every=100, log=print):
    import torch
    rng = np.random.default_rng(seed)
    n_pairs = sum(x.shape[0] for x in xs)
    index = [(i, j) for i, x in enumerate(xs) for j in range(x.shape[0])]
    work = torch.empty(_lib.load().tr_cnn_wgrad_work_size(512), dtype=torch.float32, device=net.device)
    partial = torch.empty((1024, 2), dtype=torch.float64, device=net.device)
    hist, acc, t0 = [], 0.0, time.perf_counter()
    for step in range(1, steps + 1):
        i, j = index[rng.integers(n_pairs)]
        x = xs[i][j:j + 1]
        tgt = ts[i][j:j + 1]
        out, acts = net.forward(x, 1, shape, keep=True)
        _, dy, count = net.loss_grad(out, tgt, 1, shape, partial=partial)
        gw, gb = net.backward(acts, dy, 1, shape, work=work)
        cur = lr * 0.5 * (1.0 + math.cos(math.pi * (step - 1) / steps))
        net.adam_step(gw, gb, max(cur, lr * 0.02))
        acc += float(partial[:, 0].sum()) / count
        if step % log_every == 0:
            hist.append((step, acc / log_every))
            log(f"[train] step {step}/{steps} loss {acc / log_every:.5f} lr {cur:.2e} "
                f"{time.perf_counter() - t0:.1f}s")
            acc = 0.0
    return hist


def relative_mse_of_identity(xs, ts):
    """Loss of the network f = 0 (pred = x): the baseline the training must beat."""
    tot, cnt = 0.0, 0
    for t in ts:
        tot += float((t.double() ** 2).sum())
        cnt += t.numel()
    return tot / cnt


def evaluate(net, seeds, device, log=print):
    """Plain vs warm-started PDHG on held-out seeds: iterations, seconds, objective."""
    import torch
    from .. import objective
    rows = []
    for seed in seeds:
        solver = make_solver(training_config(seed), device)
        torch.cuda.synchronize(device)
        t0 = time.perf_counter()
        u_plain = solver.run()
        torch.cuda.synchronize(device)
        t_plain, it_plain = time.perf_counter() - t0, solver.iterations
        phi_plain = objective.objective(solver, u_plain)
        del u_plain
        torch.cuda.synchronize(device)
        t0 = time.perf_counter()
        u0 = cnn_warm_start(solver, net)
        u_warm = solver.run(init=u0, dual_init="residual")
        torch.cuda.synchronize(device)
        t_warm, it_warm = time.perf_counter() - t0, solver.iterations
        phi_warm = objective.objective(solver, u_warm)
        row = {"seed": seed, "plain_iterations": it_plain, "plain_s": t_plain, "warm_iterations": it_warm,
               "warm_s": t_warm, "objective_plain": phi_plain, "objective_warm": phi_warm,
               "objective_rel_diff": (phi_warm - phi_plain) / phi_plain}
        log(f"[eval] {json.dumps(row)}")
        rows.append(row)
        del solver, u0, u_warm
    return rows


def main(argv=None):
    import torch
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--train-seeds", default="1000:1084", help="start:stop of the training seeds")
    ap.add_argument("--eval-seeds", default="9000:9002", help="start:stop of held-out seeds")
    ap.add_argument("--steps", type=int, default=4000)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--init-seed", type=int, default=0)
    ap.add_argument("--out", default=default_weights_path())
    ap.add_argument("--report", default=None, help="write the training / evaluation summary (JSON) here")
    args = ap.parse_args(argv)

    def log(msg):
        print(msg, file=sys.stderr, flush=True)

    dev = torch.device("cuda", 0)
    a, b = (int(v) for v in args.train_seeds.split(":"))
    xs, ts, its = [], [], []
    t0 = time.perf_counter()
    for seed in range(a, b):
        solver = make_solver(training_config(seed), dev)
        x, t, it = make_pairs(solver)
        xs.append(x)
        ts.append(t)
        its.append(it)
        del solver
        log(f"[data] seed {seed}: {x.shape[0]} pairs, plain PDHG {it} iterations "
            f"({time.perf_counter() - t0:.0f}s)")
    base = relative_mse_of_identity(xs, ts)
    n_pairs = sum(x.shape[0] for x in xs)
    log(f"[data] {n_pairs} pairs; loss of the identity map {base:.5f}")
    net = WarmStartNet(dev, seed=args.init_seed)
    hist = train(net, xs, ts, TRAIN_SHAPE, args.steps, args.lr, log=log)
    del xs, ts
    torch.cuda.empty_cache()
    ea, eb = (int(v) for v in args.eval_seeds.split(":"))
    rows = evaluate(net, range(ea, eb), dev, log=log)
    net.save(args.out, train_shape=list(TRAIN_SHAPE), train_seeds=[a, b], steps=args.steps, lr=args.lr,
             identity_loss=base, final_loss=hist[-1][1] if hist else float("nan"))
    report = {"train_seeds": [a, b], "pairs": n_pairs, "steps": args.steps,
              "identity_loss": base, "loss_history": hist, "plain_iterations_train": its, "eval": rows}
    if args.report:
        with open(args.report, "w") as f:
            json.dump(report, f, indent=1)
    log(f"[train] saved {args.out}")
    return 0


if __name__ == "__main__":
    sys.exit(main())
