"""Parity at the benchmark shapes (80x192x192, 32 coils, 6 echoes, 6 bins, 9000x512 cone samples).

* TVME data term on a full bin (0.72 M samples x 192 channels) with the autotuner's launch modes
  (grouped interpolation, XCD chunks, support pruning on): the forward NUFFT and the
  density-weighted adjoint of single (coil, echo) channels against the float64 oracle, one channel
  per echo so the oracle runs in seconds.
* TVMW at the same shape: the db6 zero-mode analysis / synthesis of one (bin, echo) image at the
  automatic level (2: 80 -> 45 -> 28) against the oracle, the full-bin normal equation, and two
  PDHG iterations that are bitwise identical between two runs.
* The learned warm start on a held-out generator seed (2; the benchmark uses seed 1, training
  used 64x96x96 problems): fewer than half the plain iterations to tol 1e-3 with the final
  objective within +-0.1 %.
"""

import numpy as np
import pytest

from oracle import nufft as on
from oracle import wavelet as ow
from problems import rel

pytestmark = pytest.mark.gpu

TOL = 1e-5


@pytest.fixture(scope="module")
def torch_mod():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


@pytest.fixture(scope="module")
def tvme_full(torch_mod):
    """TVME solver at the benchmark size on generator seed 2, k-space simulated on the GPU,
    launch parameters from the autotuner (bins on one stream)."""
    from toporecon_hip import synth
    from toporecon_hip.algorithms import TvmeReconstructor
    cfg = synth.config_full_tvme()
    cfg.seed = 2
    print("tvme 80x192x192 c=32 e=6 b=6 seed 2: building", flush=True)
    acq = synth.make_acquisition(cfg)
    s = TvmeReconstructor(None, acq.coord, acq.dcf, acq.mps, acq.resp, None, cfg.bins, cfg.l2_coupling,
                          lamda_m=cfg.lambda_motion, lamda_e=cfg.lambda_echo, sigma=1 / 6, tau=1 / 6,
                          max_iter=cfg.max_iter, tol=cfg.tol, margin=0, device=0, bin_edges=acq.bin_edges,
                          show_pbar=False, echoes=cfg.echoes)
    s.simulate_kspace(acq.truth, cfg.snr_db, seed=cfg.seed + 104729)
    tuned = s.autotune(bin_streams=False)
    print("autotuned", {k: tuned.get(k) for k in ("interp_xcd_chunk", "interp_kgroups", "spread_xcd_chunk")},
          flush=True)
    yield cfg, acq, s, tuned
    s.plan.close()


def test_tvme_full_bin_channels_match_oracle(torch_mod, tvme_full):
    torch = torch_mod
    from toporecon_hip import _lib
    cfg, acq, s, tuned = tvme_full
    assert tuned["interp_kgroups"] in (0, 1, 2, 3)
    b = 0
    E, C, K = s.E, s.C, s.E * s.C
    M = s.plan.samples(b)
    coord = np.asarray(acq.coord[acq.bin_index(b)]).reshape(-1, 3)
    dcf = np.asarray(acq.dcf[acq.bin_index(b)]).reshape(-1)
    assert coord.shape[0] == M
    rng = np.random.default_rng(8)
    shape = tuple(s.img_shape)
    x = (rng.standard_normal((E,) + shape) + 1j * rng.standard_normal((E,) + shape)).astype(np.complex64)
    xt = torch.from_numpy(x).cuda()
    maps = s.maps.cpu().numpy()
    st = s.stream
    # forward over all 192 channels, compared on one channel per echo
    a_sorted = torch.empty((M, K), dtype=torch.complex64, device="cuda")
    _lib.call("tr_data_simulate", s.plan.handle, b, xt.data_ptr(), s.maps.data_ptr(), E, C, a_sorted.data_ptr(), st)
    a = torch.empty((K, M), dtype=torch.complex64, device="cuda")
    _lib.call("tr_plan_scatter_samples", s.plan.handle, b, a_sorted.data_ptr(), K, a.data_ptr(), st)
    a = a.cpu().numpy()
    chans = [(e, (11 * e + 5) % C) for e in range(E)]
    for e, c in chans[:3]:
        ref = on.nufft(x[e].astype(np.complex128) * maps[c], coord)
        err = rel(a[e * C + c], ref)
        print(f"forward channel (e={e}, c={c}): rel-L2 {err:.2e}", flush=True)
        assert err < TOL
    # density-weighted adjoint: q nonzero on one coil per echo, so g[e] = conj(S_c) A^H(dcf q_ec)
    q = np.zeros((K, M), dtype=np.complex64)
    for e, c in chans:
        q[e * C + c] = (rng.standard_normal(M) + 1j * rng.standard_normal(M)).astype(np.complex64)
    qt = torch.from_numpy(q).cuda()
    q_sorted = torch.empty((M, K), dtype=torch.complex64, device="cuda")
    _lib.call("tr_plan_gather_samples", s.plan.handle, b, qt.data_ptr(), K, q_sorted.data_ptr(), st)
    g = torch.empty((E,) + shape, dtype=torch.complex64, device="cuda")
    _lib.call("tr_data_adjoint", s.plan.handle, b, q_sorted.data_ptr(), s.dcf[b].data_ptr(), s.maps.data_ptr(),
              E, C, g.data_ptr(), 0, st)
    g = g.cpu().numpy()
    for e, c in chans[:3]:
        ref = np.conj(maps[c].astype(np.complex128)) * on.nufft_adjoint(q[e * C + c] * dcf, coord, shape)
        err = rel(g[e], ref)
        print(f"adjoint channel (e={e}, c={c}): rel-L2 {err:.2e}", flush=True)
        assert err < TOL


@pytest.mark.timeout(600)
def test_learned_warm_start_full_size_held_out(torch_mod, tvme_full):
    from toporecon_hip import objective
    from toporecon_hip.learned import WarmStartNet, default_weights_path, warm_start
    cfg, acq, s, tuned = tvme_full
    print("plain solve", flush=True)
    u_plain = s.run(progress_every=50).clone()
    it_plain = s.iterations
    phi_plain = objective.objective(s, u_plain)
    del u_plain
    net = WarmStartNet.load(default_weights_path(), 0)
    print("learned solve", flush=True)
    u_warm = s.run(init=warm_start(s, net), dual_init="residual")
    it_warm = s.iterations
    phi_warm = objective.objective(s, u_warm)
    d = (phi_warm - phi_plain) / phi_plain
    print(f"plain {it_plain} it, learned {it_warm} it, objective difference {d:+.2e}", flush=True)
    assert it_warm < 0.5 * it_plain, (it_warm, it_plain)
    assert abs(d) <= 1e-3


@pytest.mark.timeout(600)
def test_tvmw_full_size(torch_mod, tvme_full):
    torch = torch_mod
    from toporecon_hip import _lib, synth
    from toporecon_hip.algorithms import TvmwReconstructor
    _, acq, s_tvme, _ = tvme_full
    s_tvme.plan.release_workspace()
    cfg = synth.config_full_tvme()
    cfg.seed = 2
    print("tvmw 80x192x192 c=32 e=6 b=6", flush=True)
    s = TvmwReconstructor(None, acq.coord, acq.dcf, acq.mps, acq.resp, None, cfg.bins, True, lambda1=1e-5,
                          lambda2=1e-5, lambda3=1e-5, sigma=1 / 6, tau=1 / 6, max_iter=2, tol=0.0, margin=0,
                          device=0, bin_edges=acq.bin_edges, show_pbar=False, echoes=cfg.echoes)
    assert s.level_w2 == ow.max_level(s.img_shape) == 2
    s.simulate_kspace(acq.truth, cfg.snr_db, seed=cfg.seed + 104729)
    st = s.stream
    shape = tuple(s.img_shape)
    lib = _lib.load()
    rng = np.random.default_rng(9)
    # db6 analysis / synthesis of one (bin, echo) image at level 2
    x = (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(np.complex64)
    xt = torch.from_numpy(x).cuda()
    ncoef = int(lib.tr_w2_coeff_count(*shape, 2))
    work = torch.empty((int(lib.tr_w2_work_count(*shape, 2)),), dtype=torch.complex64, device="cuda")
    ct = torch.empty((1, ncoef), dtype=torch.complex64, device="cuda")
    _lib.call("tr_w2_forward", xt.data_ptr(), 1, *shape, 2, ct.data_ptr(), work.data_ptr(), st)
    ref = np.concatenate([v.reshape(-1) for v in ow.tree_leaves(ow.wavedec3(x.astype(np.complex128), 2))])
    assert ref.size == ncoef
    e_fwd = rel(ct.cpu().numpy()[0], ref)
    d = (rng.standard_normal(ncoef) + 1j * rng.standard_normal(ncoef)).astype(np.complex64)
    dt = torch.from_numpy(d[None]).cuda()
    out = torch.empty((1,) + shape, dtype=torch.complex64, device="cuda")
    _lib.call("tr_w2_adjoint", dt.data_ptr(), 1, *shape, 2, out.data_ptr(), work.data_ptr(), st)
    tree, off = [], 0
    shapes = ow.level_shapes(shape, 2)
    n = int(np.prod(shapes[-1]))
    tree.append(d[off:off + n].reshape(shapes[-1]).astype(np.complex128))
    off += n
    for lv in (2, 1):
        n = int(np.prod(shapes[lv]))
        lev = {}
        for k in ow._KEYS:
            lev[k] = d[off:off + n].reshape(shapes[lv]).astype(np.complex128)
            off += n
        tree.append(lev)
    e_adj = rel(out.cpu().numpy()[0], ow.waverec3(tree, shape))
    print(f"db6 level 2 analysis {e_fwd:.2e}, synthesis {e_adj:.2e}", flush=True)
    assert e_fwd < 1e-6 and e_adj < 1e-6
    del work, ct, dt, out
    # normal equation of a full bin
    E, C = s.E, s.C
    b = 0
    xe = torch.complex(torch.randn((E,) + shape, device="cuda"), torch.randn((E,) + shape, device="cuda"))
    ax = torch.empty_like(s.y[b])
    _lib.call("tr_data_simulate", s.plan.handle, b, xe.data_ptr(), s.maps.data_ptr(), E, C, ax.data_ptr(), st)
    g = torch.empty_like(xe)
    _lib.call("tr_data_adjoint", s.plan.handle, b, ax.data_ptr(), s.dcf[b].data_ptr(), s.maps.data_ptr(), E, C,
              g.data_ptr(), 0, st)
    lhs = torch.sum(xe.conj().to(torch.complex128) * g.to(torch.complex128))
    rhs = torch.sum(s.dcf[b].to(torch.float64)[:, None] * torch.abs(ax.to(torch.complex128)) ** 2)
    print("normal equation", float(lhs.real), float(rhs), flush=True)
    assert abs(float(lhs.real) - float(rhs)) / float(rhs) < 1e-5
    assert abs(float(lhs.imag)) / float(rhs) < 1e-5
    del xe, ax, g
    # two TVMW iterations, twice: bitwise identical
    u1 = s.run().clone()
    h1 = list(s.history)
    u2 = s.run()
    assert s.iterations == 2 and h1 == s.history
    assert torch.equal(torch.view_as_real(u1), torch.view_as_real(u2))
    assert bool(torch.isfinite(torch.view_as_real(u1)).all())
    s.plan.close()
