"""HIP NUFFT parity against the float64 oracle (rel-L2 <= 1e-5 per operator application)."""

import numpy as np
import pytest

from oracle import nufft as on

pytestmark = pytest.mark.gpu

TOL = 1e-5


def _coords(rng, shape, M, special=True):
    half = np.array(shape) / 2.0
    c = (rng.uniform(-1, 1, (M, 3)) * half).astype(np.float32)
    if special:
        c[0] = 0.0                                   # k = 0: integral scaled position, 5 taps
        c[1] = -half                                 # corner of k-space, wraps
        c[2] = half * 0.999
        c[3] = np.round(c[3])                        # on-grid sample
        c[4] = [0.0, -half[1], half[2]]
    return c


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def _cplx(rng, shape):
    return (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(np.complex64)


@pytest.fixture(scope="module")
def torch_mod():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


@pytest.mark.parametrize("shape", [(12, 16, 16), (9, 11, 7), (8, 20, 6), (16, 16, 24)])
def test_forward_single_image(torch_mod, shape):
    from toporecon_hip.nufft_backend import create_nufft_backend
    print("forward single", shape, flush=True)
    rng = np.random.default_rng(1)
    coords = [_coords(rng, shape, 1500), _coords(rng, shape, 700)]
    be = create_nufft_backend("hip", img_shape=shape, coordinates=coords, device=0)
    img = _cplx(rng, shape)
    for b in range(2):
        out = be.forward(img, b).cpu().numpy()
        ref = on.nufft(img.astype(np.complex128), coords[b])
        assert out.shape == ref.shape
        e = np.abs(out - ref) / np.abs(ref).max()
        bad = np.nonzero(e > 1e-5)[0]
        print("bin", b, "bad samples", bad[:20].tolist(), "coords", coords[b][bad[:5]].tolist(), flush=True)
        assert _rel(out, ref) < TOL, (b, _rel(out, ref))


@pytest.mark.parametrize("shape", [(12, 16, 16), (9, 11, 7), (8, 20, 6)])
def test_adjoint_single_image(torch_mod, shape):
    from toporecon_hip.nufft_backend import create_nufft_backend
    print("adjoint single", shape, flush=True)
    rng = np.random.default_rng(2)
    coords = [_coords(rng, shape, 1200).reshape(30, 40, 3)]
    be = create_nufft_backend("hip", img_shape=shape, coordinates=coords, device=0)
    k = _cplx(rng, (30, 40))
    out = be.adjoint(k, 0).cpu().numpy()
    ref = on.nufft_adjoint(k.astype(np.complex128), coords[0], shape)
    assert out.shape == tuple(shape)
    assert _rel(out, ref) < TOL, _rel(out, ref)


@pytest.mark.parametrize("E,C", [(2, 3), (6, 32), (1, 8), (3, 12)])
def test_batched_coils_echoes(torch_mod, E, C):
    from toporecon_hip.nufft_backend import HipNufftBackend
    print("batched", E, C, flush=True)
    shape = (8, 12, 10)
    rng = np.random.default_rng(E * 100 + C)
    coords = [_coords(rng, shape, 900)]
    be = HipNufftBackend(shape, coords, 0)
    img = _cplx(rng, (E,) + shape)
    maps = _cplx(rng, (C,) + shape)
    out = be.forward_batch(img, maps, 0).cpu().numpy()
    ref = on.nufft(img[:, None].astype(np.complex128) * maps[None], coords[0])
    assert _rel(out, ref) < TOL
    k = _cplx(rng, (E, C, 900))
    dcf = rng.uniform(0.1, 1.0, 900).astype(np.float32)
    adj = be.adjoint_batch(k, maps, 0, dcf=dcf).cpu().numpy()
    x = on.nufft_adjoint(k.astype(np.complex128) * dcf, coords[0], shape)
    ref = np.sum(np.conj(maps)[None] * x, axis=1)
    assert _rel(adj, ref) < TOL


@pytest.mark.parametrize("shape,E,C", [((8, 16, 192), 2, 16), ((8, 192, 8), 2, 16), ((80, 8, 8), 2, 16),
                                       ((32, 48, 64), 2, 16), ((16, 96, 48), 2, 16),
                                       ((8, 16, 192), 3, 8), ((80, 8, 8), 3, 8), ((16, 96, 48), 3, 8)])
def test_two_level_fft_lengths(torch_mod, shape, E, C):
    """Batches on grids whose lengths (240, 100, 80, 60, 40, 120) take the two-level register FFT
    in the x, y and z passes, forward and density-weighted adjoint: 32 channels (whole 16-channel
    chunks) and 24 channels (a half-empty second chunk, whose dead lanes the y / z passes mask
    with out-of-range buffer offsets)."""
    from toporecon_hip.nufft_backend import HipNufftBackend
    print("two-level", shape, E, C, flush=True)
    M = 3000
    rng = np.random.default_rng(sum(shape))
    coords = [_coords(rng, shape, M)]
    be = HipNufftBackend(shape, coords, 0)
    img = _cplx(rng, (E,) + shape)
    maps = _cplx(rng, (C,) + shape)
    out = be.forward_batch(img, maps, 0).cpu().numpy()
    ref = on.nufft(img[:, None].astype(np.complex128) * maps[None], coords[0])
    assert _rel(out, ref) < TOL, _rel(out, ref)
    k = _cplx(rng, (E, C, M))
    dcf = rng.uniform(0.1, 1.0, M).astype(np.float32)
    adj = be.adjoint_batch(k, maps, 0, dcf=dcf).cpu().numpy()
    x = on.nufft_adjoint(k.astype(np.complex128) * dcf, coords[0], shape)
    ref = np.sum(np.conj(maps)[None] * x, axis=1)
    assert _rel(adj, ref) < TOL, _rel(adj, ref)


def test_touched_points_match_oracle_taps(torch_mod):
    """The plan's per-bin count of distinct grid points under the taps (the interpolation's
    algorithmic grid bytes) equals the support of the oracle's gridding of all-ones samples
    (every Kaiser-Bessel tap weight, including |x| = 1, is positive)."""
    from toporecon_hip.nufft_backend import HipNufftBackend
    shape = (12, 16, 14)
    rng = np.random.default_rng(17)
    coords = [_coords(rng, shape, 3000), _coords(rng, shape, 40)]
    be = HipNufftBackend(shape, coords, 0)
    gshape = on.oversamp_shape(shape)
    beta32 = float(np.float32(on.kb_beta()))
    for b, c in enumerate(coords):
        g = on.gridding(np.ones(len(c), np.complex128), on.scale_coord(c, shape), gshape, beta32)
        assert be.plan.touched_points(b) == int(np.count_nonzero(g)), b
        assert be.plan.touched_points(b) <= be.plan.info(b)[11] <= int(np.prod(gshape))


def test_adjointness_and_determinism(torch_mod):
    from toporecon_hip.nufft_backend import HipNufftBackend
    shape = (12, 16, 14)
    rng = np.random.default_rng(7)
    coords = [_coords(rng, shape, 4000)]
    be = HipNufftBackend(shape, coords, 0)
    x = _cplx(rng, shape)
    y = _cplx(rng, (4000,))
    Ax = be.forward(x, 0).cpu().numpy().astype(np.complex128)
    Ahy1 = be.adjoint(y, 0).cpu().numpy()
    Ahy2 = be.adjoint(y, 0).cpu().numpy()
    assert np.array_equal(Ahy1.view(np.uint64), Ahy2.view(np.uint64))  # bitwise rerun
    lhs = np.vdot(Ax, y)
    rhs = np.vdot(x.astype(np.complex128), Ahy1.astype(np.complex128))
    assert abs(lhs - rhs) / abs(lhs) < 1e-5


def test_empty_and_ragged_bins(torch_mod):
    from toporecon_hip.nufft_backend import HipNufftBackend
    shape = (8, 8, 8)
    rng = np.random.default_rng(8)
    coords = [_coords(rng, shape, 5), np.zeros((0, 3), np.float32), _coords(rng, shape, 333, False)]
    be = HipNufftBackend(shape, coords, 0)
    img = _cplx(rng, shape)
    assert be.forward(img, 1).numel() == 0
    adj = be.adjoint(np.zeros((0,), np.complex64), 1).cpu().numpy()
    assert not np.any(adj)
    for b in (0, 2):
        ref = on.nufft(img.astype(np.complex128), coords[b])
        assert _rel(be.forward(img, b).cpu().numpy(), ref) < TOL


def test_errors(torch_mod):
    from toporecon_hip import _lib
    from toporecon_hip.nufft_backend import HipNufftBackend
    rng = np.random.default_rng(9)
    be = HipNufftBackend((8, 8, 8), [_coords(rng, (8, 8, 8), 10)], 0)
    with pytest.raises(IndexError, match="outside"):
        be.forward(np.zeros((8, 8, 8), np.complex64), 1)
    with pytest.raises(ValueError, match="must end in dimension 3"):
        HipNufftBackend((8, 8, 8), [np.zeros((10, 2), np.float32)], 0)
    with pytest.raises(ValueError, match="At least one trajectory bin"):
        HipNufftBackend((8, 8, 8), [], 0)
    with pytest.raises(_lib.HipLibraryError, match="not supported"):
        HipNufftBackend((1700, 4, 4), [_coords(rng, (1700, 4, 4), 10)], 0)  # grid 2125: 17 | 2125 > 1024


@pytest.mark.parametrize("shape,frac", [((12, 16, 20), 0.25), ((10, 14, 9), 0.4), ((16, 16, 16), 0.05)])
def test_partial_support_forward_adjoint(torch_mod, shape, frac):
    """Samples confined to the centre of k-space leave whole grid columns, and whole x
    planes of them, outside the support.  The pruned passes skip those points; forward and
    adjoint still match the oracle and equal the unpruned passes bit for bit."""
    from toporecon_hip.nufft_backend import create_nufft_backend
    print("partial support", shape, frac, flush=True)
    rng = np.random.default_rng(9)
    half = np.array(shape) / 2.0
    coords = [(rng.uniform(-1, 1, (900, 3)) * half * frac).astype(np.float32),
              (rng.uniform(-1, 1, (500, 3)) * half * frac * 0.5).astype(np.float32)]
    be = create_nufft_backend("hip", img_shape=shape, coordinates=coords, device=0)
    plan = be.plan
    G = int(np.prod(plan.grid_shape))
    assert plan.get_param("support_points") < G
    assert plan.get_param("live_columns") < plan.grid_shape[1] * plan.grid_shape[2]
    img = _cplx(rng, shape)
    k = _cplx(rng, (900,))
    fwd = be.forward(img, 0).cpu().numpy()
    adj = be.adjoint(k, 0).cpu().numpy()
    assert _rel(fwd, on.nufft(img.astype(np.complex128), coords[0])) < TOL
    assert _rel(adj, on.nufft_adjoint(k.astype(np.complex128), coords[0], shape)) < TOL
    plan.set_param("prune", 0)
    fwd0 = be.forward(img, 0).cpu().numpy()
    adj0 = be.adjoint(k, 0).cpu().numpy()
    plan.set_param("prune", 1)
    assert np.array_equal(fwd.view(np.uint64), fwd0.view(np.uint64))
    assert np.array_equal(adj.view(np.uint64), adj0.view(np.uint64))


def test_grouped_interp_never_reads_unwritten_workspace(torch_mod):
    """With support pruning the forward passes leave grid points outside the k-space support
    unwritten; the grouped interpolation's 5^3 boxes must stay inside the support, or weight-0
    terms would read stale workspace (0 * NaN = NaN).  The workspace is filled with NaN first."""
    import torch

    from toporecon_hip import _lib, synth
    from toporecon_hip.nufft_backend import NufftPlan
    print("grouped interpolation vs stale workspace", flush=True)
    torch = torch_mod
    shape = (16, 24, 24)
    coord = synth.cone_trajectory(300, 64, shape).reshape(-1, 3)
    dev = torch.device("cuda", 0)
    plan = NufftPlan(shape, [torch.from_numpy(coord).to(dev)], dev)
    G = int(np.prod(plan.grid_shape))
    assert plan.get_param("support_points") < G
    E, C = 3, 32
    K = E * C
    rng = np.random.default_rng(7)
    img = torch.from_numpy(_cplx(rng, (E, *shape))).to(dev)
    maps = torch.from_numpy(_cplx(rng, (C, *shape))).to(dev)
    nan_grid = torch.full((2 * G * K,), float("nan"), dtype=torch.float32, device=dev).view(torch.complex64)
    out_img = torch.empty((E, *shape), dtype=torch.complex64, device=dev)
    s = _lib.stream_handle(dev)
    M = plan.samples(0)
    plan.set_param("interp_kgroups", 3)
    ys = []
    for prune in (1, 0):
        plan.set_param("prune", prune)
        _lib.call("tr_nufft_adjoint_grid", plan.handle, nan_grid.data_ptr(), maps.data_ptr(), E, C,
                  out_img.data_ptr(), s)
        y = torch.empty((M, K), dtype=torch.complex64, device=dev)
        _lib.call("tr_data_simulate", plan.handle, 0, img.data_ptr(), maps.data_ptr(), E, C, y.data_ptr(), s)
        ys.append(y.cpu().numpy())
    plan.close()
    assert np.all(np.isfinite(ys[0]))
    assert np.array_equal(ys[0], ys[1])


def test_spreading_orders_and_heavy_segments(torch_mod):
    """The spreading's cluster visiting orders (longest first, bricks, bricks longest first, index
    order over every cluster) and support pruning give identical bits, also for clusters whose pair lists are long
    enough to be summed in segments (every cone interleave starts at k = 0, so the centre cluster
    collects thousands of pairs), and the result matches the oracle."""
    torch = torch_mod
    from toporecon_hip import _lib, synth
    from toporecon_hip.nufft_backend import NufftPlan
    print("spreading orders / heavy segments", flush=True)
    shape = (12, 16, 16)
    coord = synth.cone_trajectory(3000, 24, shape).reshape(-1, 3)
    dev = torch.device("cuda", 0)
    plan = NufftPlan(shape, [torch.from_numpy(coord).to(dev)], dev)
    assert plan.get_param("spread_order") == 0xffff
    E, C = 2, 3
    K = E * C
    M = plan.samples(0)
    rng = np.random.default_rng(21)
    maps = torch.from_numpy(_cplx(rng, (C, *shape))).to(dev)
    q = _cplx(rng, (K, M))
    dcf = rng.uniform(0.2, 1.0, M).astype(np.float32)
    qt = torch.from_numpy(q).to(dev)
    dt = torch.from_numpy(dcf).to(dev)
    s = _lib.stream_handle(dev)
    outs = []
    for order, prune in ((0xffff, 1), (0x111, 1), (0x822, 1), (0x10422, 1), (0, 1), (0xffff, 0)):
        plan.set_param("spread_order", order)
        plan.set_param("prune", prune)
        g = torch.empty((E, *shape), dtype=torch.complex64, device=dev)
        _lib.call("tr_nufft_adjoint", plan.handle, 0, qt.data_ptr(), dt.data_ptr(), maps.data_ptr(), E, C,
                  g.data_ptr(), 0, s)
        outs.append(g.cpu().numpy())
    for o in outs[1:]:
        assert np.array_equal(o.view(np.uint64), outs[0].view(np.uint64))
    mp = maps.cpu().numpy().astype(np.complex128)
    for e in range(E):
        ref = sum(np.conj(mp[c]) * on.nufft_adjoint(q[e * C + c] * dcf, coord, shape) for c in range(C))
        assert _rel(outs[0][e], ref) < TOL
    with pytest.raises(Exception):
        plan.set_param("spread_order", 0x20000)
    with pytest.raises(Exception):
        plan.set_param("spread_order", 0x1ffff)
    plan.close()


@pytest.mark.parametrize("tile,seg", [(4, 1024), (16, 1024), (32, 64), (8, 64), (2, 65536)])
def test_sample_geometry_forward_adjoint(torch_mod, tile, seg):
    """Plans built with other sort tiles and heavy-cluster segment lengths (64: the centre
    clusters of a cone trajectory are cut into dozens of segments; 65536: never cut) match the
    oracle forward and density-weighted adjoint, rerun bit for bit, and report their geometry."""
    torch = torch_mod
    from toporecon_hip import _lib, synth
    from toporecon_hip.nufft_backend import NufftPlan
    print("geometry", tile, seg, flush=True)
    shape = (12, 16, 16)
    coord = synth.cone_trajectory(2000, 24, shape).reshape(-1, 3)
    dev = torch.device("cuda", 0)
    plan = NufftPlan(shape, [torch.from_numpy(coord).to(dev)], dev, sort_tile=tile, spread_segment=seg)
    assert (plan.get_param("sort_tile"), plan.get_param("spread_segment")) == (tile, seg)
    with pytest.raises(Exception):
        plan.set_param("sort_tile", 8)
    E, C = 2, 3
    K = E * C
    M = plan.samples(0)
    rng = np.random.default_rng(tile * 7 + seg)
    img = _cplx(rng, (E, *shape))
    maps = _cplx(rng, (C, *shape))
    q = _cplx(rng, (K, M))
    dcf = rng.uniform(0.2, 1.0, M).astype(np.float32)
    it, mt, qt, dt = (torch.from_numpy(v).to(dev) for v in (img, maps, q, dcf))
    st = _lib.stream_handle(dev)
    fwd = torch.empty((K, M), dtype=torch.complex64, device=dev)
    _lib.call("tr_nufft_forward", plan.handle, 0, it.data_ptr(), mt.data_ptr(), E, C, fwd.data_ptr(), st)
    outs = []
    for _ in range(2):
        g = torch.empty((E, *shape), dtype=torch.complex64, device=dev)
        _lib.call("tr_nufft_adjoint", plan.handle, 0, qt.data_ptr(), dt.data_ptr(), mt.data_ptr(), E, C,
                  g.data_ptr(), 0, st)
        outs.append(g.cpu().numpy())
    assert np.array_equal(outs[0].view(np.uint64), outs[1].view(np.uint64))
    f = fwd.cpu().numpy()
    mp = maps.astype(np.complex128)
    for e in range(E):
        for c in range(C):
            ref = on.nufft(img[e].astype(np.complex128) * mp[c], coord)
            assert _rel(f[e * C + c], ref) < TOL
        ref = sum(np.conj(mp[c]) * on.nufft_adjoint(q[e * C + c] * dcf, coord, shape) for c in range(C))
        assert _rel(outs[0][e], ref) < TOL
    plan.close()


def test_sample_geometry_limits(torch_mod):
    torch = torch_mod
    from toporecon_hip.nufft_backend import NufftPlan
    dev = torch.device("cuda", 0)
    c = [torch.zeros((10, 3), dtype=torch.float32, device=dev)]
    for tile, seg in ((1, 1024), (33, 1024), (8, 63), (8, 65537)):
        with pytest.raises(Exception):
            NufftPlan((8, 8, 8), c, dev, sort_tile=tile, spread_segment=seg)


@pytest.mark.parametrize("shape", [(13, 9, 7), (136, 10, 8), (8, 150, 9), (9, 8, 220), (6, 256, 7)])
def test_general_grid_lengths(torch_mod, shape):
    """Grid lengths the compile-time FFT plans do not cover (17, 170, 188, 275 with prime factors
    17, 47, 11 x 25; 320 above 256 points) against the oracle, forward and density-weighted adjoint
    over 2 echoes x 3 coils."""
    torch = torch_mod
    from toporecon_hip import _lib
    from toporecon_hip.nufft_backend import HipNufftBackend
    print("general lengths", shape, flush=True)
    rng = np.random.default_rng(sum(shape))
    coords = [_coords(rng, shape, 1200)]
    be = HipNufftBackend(shape, coords, 0)
    E, C = 2, 3
    K = E * C
    M = coords[0].shape[0]
    img = _cplx(rng, (E, *shape))
    maps = _cplx(rng, (C, *shape))
    ti, tm = torch.from_numpy(img).cuda(), torch.from_numpy(maps).cuda()
    out = torch.empty((K, M), dtype=torch.complex64, device="cuda")
    s = _lib.stream_handle()
    _lib.call("tr_nufft_forward", be.plan.handle, 0, ti.data_ptr(), tm.data_ptr(), E, C, out.data_ptr(), s)
    got = out.cpu().numpy()
    ref = np.stack([on.nufft(img[e].astype(np.complex128) * maps[c], coords[0]) for e in range(E) for c in range(C)])
    assert _rel(got, ref) < TOL, _rel(got, ref)
    q = _cplx(rng, (K, M))
    dcf = rng.uniform(0.2, 1.0, M).astype(np.float32)
    qt, dt = torch.from_numpy(q).cuda(), torch.from_numpy(dcf).cuda()  # held across the launch
    g = torch.empty((E, *shape), dtype=torch.complex64, device="cuda")
    _lib.call("tr_nufft_adjoint", be.plan.handle, 0, qt.data_ptr(), dt.data_ptr(), tm.data_ptr(), E, C,
              g.data_ptr(), 0, s)
    refa = np.stack([sum(np.conj(maps[c].astype(np.complex128)) * on.nufft_adjoint(q[e * C + c] * dcf, coords[0], shape)
                         for c in range(C)) for e in range(E)])
    assert _rel(g.cpu().numpy(), refa) < TOL, _rel(g.cpu().numpy(), refa)


def test_many_echoes_and_long_rows(torch_mod):
    """12 echoes: the inverse x pass runs in groups of 8 echoes and, for 600-voxel rows, the forward
    x pass stages the echoes' image rows in groups that fit its LDS; both match the oracle."""
    torch = torch_mod
    from toporecon_hip import _lib
    from toporecon_hip.nufft_backend import HipNufftBackend
    print("many echoes", flush=True)
    for shape in ((6, 8, 10), (4, 5, 600)):
        rng = np.random.default_rng(31)
        coords = [_coords(rng, shape, 700)]
        be = HipNufftBackend(shape, coords, 0)
        E, C = 12, 1
        M = coords[0].shape[0]
        img = _cplx(rng, (E, *shape))
        out = torch.empty((E, M), dtype=torch.complex64, device="cuda")
        s = _lib.stream_handle()
        it = torch.from_numpy(img).cuda()  # held across the launch
        _lib.call("tr_nufft_forward", be.plan.handle, 0, it.data_ptr(), None, E, C, out.data_ptr(), s)
        ref = np.stack([on.nufft(img[e].astype(np.complex128), coords[0]) for e in range(E)])
        assert _rel(out.cpu().numpy(), ref) < TOL
        q = _cplx(rng, (E, M))
        qt = torch.from_numpy(q).cuda()
        g = torch.empty((E, *shape), dtype=torch.complex64, device="cuda")
        _lib.call("tr_nufft_adjoint", be.plan.handle, 0, qt.data_ptr(), None, None, E, C, g.data_ptr(), 0, s)
        refa = np.stack([on.nufft_adjoint(q[e].astype(np.complex128), coords[0], shape) for e in range(E)])
        assert _rel(g.cpu().numpy(), refa) < TOL


@pytest.mark.parametrize("E,C", [(2, 65), (3, 50), (6, 32), (4, 75), (2, 193)])
def test_interp_channel_layouts_bitwise(torch_mod, E, C):
    """Channel counts of 130, 150, 192, 300 and 386 (two to four channel groups of the 16-byte-load
    kernels, the last one ragged: lanes past the group's channels repeat its first pair and store
    nothing) through every channel-group mode of the interpolation and every epilogue: the forward
    NUFFT (sample order restored) against the oracle, and the forward NUFFT, the k-space simulation,
    the residual and the data-dual update on sorted samples bit for bit against the per-sample
    kernel (mode 0; past its 256-channel limit, the 64-channel-group kernel, mode 1)."""
    import torch

    from toporecon_hip import _lib
    from toporecon_hip.nufft_backend import HipNufftBackend
    K = E * C
    print("interp channel layouts", E, C, K, flush=True)
    shape = (8, 12, 10)
    M = 900
    rng = np.random.default_rng(K)
    coords = [_coords(rng, shape, M)]
    be = HipNufftBackend(shape, coords, 0)
    dev = torch.device("cuda", 0)
    img = _cplx(rng, (E,) + shape)
    maps = _cplx(rng, (C,) + shape)
    img_t = torch.from_numpy(img).to(dev)
    maps_t = torch.from_numpy(maps).to(dev)
    y = torch.from_numpy(_cplx(rng, (M, K))).to(dev)
    q0 = torch.from_numpy(_cplx(rng, (M, K))).to(dev)
    s = _lib.stream_handle(dev)
    h = be.plan.handle
    modes = (0, 1, 2, 3) if K <= 256 else (1, 2, 3)
    runs = {}
    for kg in modes:
        be.plan.set_param("interp_kgroups", kg)
        fwd = be.forward_batch(img_t, maps_t, 0).cpu().numpy()
        sim = torch.empty((M, K), dtype=torch.complex64, device=dev)
        _lib.call("tr_data_simulate", h, 0, img_t.data_ptr(), maps_t.data_ptr(), E, C, sim.data_ptr(), s)
        res = torch.empty((M, K), dtype=torch.complex64, device=dev)
        _lib.call("tr_data_residual", h, 0, img_t.data_ptr(), maps_t.data_ptr(), E, C, y.data_ptr(),
                  res.data_ptr(), s)
        q = q0.clone()
        _lib.call("tr_data_dual_update", h, 0, img_t.data_ptr(), maps_t.data_ptr(), E, C, y.data_ptr(),
                  q.data_ptr(), 0.25, 1.25, s)
        torch.cuda.synchronize()
        runs[kg] = [fwd, sim.cpu().numpy(), res.cpu().numpy(), q.cpu().numpy()]
    ref = on.nufft(img[:, None].astype(np.complex128) * maps[None], coords[0]).reshape(K, M)
    assert _rel(runs[3][0].reshape(K, M), ref) < TOL
    base = runs[modes[0]]
    assert np.array_equal(base[2], base[1] - y.cpu().numpy())
    for kg in modes[1:]:
        for name, a, b in zip(("forward", "simulate", "residual", "dual"), runs[kg], base):
            assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), (kg, name)
