"""JSENSE on the GPU against the float64 oracle (oracle/jsense.py).

* the three 'valid' convolution kernels at the default map-kernel width 16 on a non-cubic
  calibration grid, and at small widths with coil counts that take every coils-per-thread path;
  bitwise reproducible;
* the axis-matrix kernel and the root-sum-of-squares normalisation against NumPy;
* a whole estimation on a small problem (map-kernel width 5, calibration width 10, 3 x 4 CG steps)
  against the oracle step by step (objective after every half-step) and in the final maps;
* the default settings (width 16, calibration 24, 10 x 10 steps) on a 32 x 28 x 30 image: the
  first round against the oracle, then the full run's properties (objective never rising, unit
  root sum of squares, bitwise reproducible) and its time;
* the dataset entry point (raw CFL files in, ``mps.cfl`` out) against the array entry point.
"""

import time

import numpy as np
import pytest

from oracle import jsense as oj
from oracle import nufft as on
from problems import rel

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_mod():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def crandn(rng, *shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


def dev(torch, a):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.complex64)).cuda()


def run_convs(torch, img, ker):
    """(conv, adj_img(z), adj_ker(z)) on the GPU for img (I^3), ker (C, W^3), z random."""
    from toporecon_hip import _lib
    lib = _lib.load()
    C, W = ker.shape[0], ker.shape[1]
    I = img.shape
    G = tuple(i - W + 1 for i in I)
    rng = np.random.default_rng(11)
    z = crandn(rng, C, *G)
    it, kt, zt = dev(torch, img), dev(torch, ker), dev(torch, z)
    work = torch.empty(int(lib.tr_js_work_count(*I, C, W)), dtype=torch.complex64, device="cuda")
    out = torch.empty((C,) + G, dtype=torch.complex64, device="cuda")
    ai = torch.empty(I, dtype=torch.complex64, device="cuda")
    ak = torch.empty((C, W, W, W), dtype=torch.complex64, device="cuda")
    s = _lib.stream_handle()
    _lib.call("tr_js_conv", it.data_ptr(), *I, kt.data_ptr(), C, W, out.data_ptr(), work.data_ptr(), s)
    _lib.call("tr_js_conv_adj_img", zt.data_ptr(), *I, kt.data_ptr(), C, W, ai.data_ptr(), work.data_ptr(), s)
    _lib.call("tr_js_conv_adj_ker", zt.data_ptr(), it.data_ptr(), *I, C, W, ak.data_ptr(), work.data_ptr(), s)
    torch.cuda.synchronize()
    return z, out.cpu().numpy(), ai.cpu().numpy(), ak.cpu().numpy()


@pytest.mark.parametrize("C, W, grd", [(8, 16, (23, 21, 22)), (3, 5, (9, 7, 8)), (6, 4, (5, 6, 7)), (1, 1, (3, 4, 5))])
def test_conv_kernels_match_oracle(torch_mod, C, W, grd):
    torch = torch_mod
    rng = np.random.default_rng(10 + C)
    I = tuple(g + W - 1 for g in grd)
    img = crandn(rng, *I).astype(np.complex64)
    ker = crandn(rng, C, W, W, W).astype(np.complex64)
    print(f"conv C={C} W={W} grid {grd}", flush=True)
    z, out, ai, ak = run_convs(torch, img, ker)
    img64, ker64 = img.astype(np.complex128), ker.astype(np.complex128)
    e1 = rel(out, oj.conv_valid(img64[None], ker64))
    e2 = rel(ai, oj.conv_valid_adj_img(z, ker64, I))
    e3 = rel(ak, oj.conv_valid_adj_ker(z, img64, ker64.shape))
    print(f"  forward {e1:.2e}  adjoint image {e2:.2e}  adjoint maps {e3:.2e}", flush=True)
    assert max(e1, e2, e3) < 2e-6
    _, out2, ai2, ak2 = run_convs(torch, img, ker)
    assert np.array_equal(out, out2) and np.array_equal(ai, ai2) and np.array_equal(ak, ak2)


def test_axis_apply_and_rss(torch_mod):
    torch = torch_mod
    from toporecon_hip import _lib
    rng = np.random.default_rng(12)
    x = crandn(rng, 3, 7, 9, 11).astype(np.complex64)
    xt = dev(torch, x)
    s = _lib.stream_handle()
    for axis, n_out in ((0, 13), (1, 1), (2, 6), (0, 4), (1, 9)):
        n_in = x.shape[axis + 1]
        m = crandn(rng, n_out, n_in).astype(np.complex64)
        mt = dev(torch, m)
        shp = list(x.shape)
        shp[axis + 1] = n_out
        out = torch.empty(shp, dtype=torch.complex64, device="cuda")
        outer = int(np.prod(x.shape[:axis + 1]))
        inner = int(np.prod(x.shape[axis + 2:]))
        _lib.call("tr_js_axis_apply", xt.data_ptr(), outer, n_in, inner, mt.data_ptr(), n_out, out.data_ptr(), s)
        want = np.moveaxis(np.tensordot(m.astype(np.complex128), x.astype(np.complex128), axes=([1], [axis + 1])),
                           0, axis + 1)
        e = rel(out.cpu().numpy(), want)
        print(f"axis {axis} {n_in}->{n_out}: {e:.2e}", flush=True)
        assert e < 1e-6
    maps = crandn(rng, 5, 4, 6, 7).astype(np.complex64)
    mt = dev(torch, maps)
    _lib.call("tr_js_rss_normalize", mt.data_ptr(), 5, 4 * 6 * 7, s)
    want = maps / np.sqrt(np.sum(np.abs(maps.astype(np.complex128)) ** 2, axis=0))
    assert rel(mt.cpu().numpy(), want) < 1e-6


def kooshball(spokes, readout, kmax, seed):
    rng = np.random.default_rng(seed)
    d = rng.standard_normal((spokes, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    r = np.linspace(0, 1, readout)
    return (r[None, :, None] * d[:, None, :] * np.asarray(kmax, float)).astype(np.float32)


def phantom_kspace(shape, coils, spokes, readout, seed):
    """k-space (1, C, spokes, readout) of a smooth object under smooth coil maps (float64 NUFFT)."""
    rng = np.random.default_rng(seed)
    coord = kooshball(spokes, readout, np.asarray(shape) / 2, seed)
    g = np.meshgrid(*[(np.arange(n) - n // 2) / n for n in shape], indexing="ij")
    r2 = sum(v ** 2 for v in g)
    obj = (r2 < 0.16).astype(float) * (1.0 + 0.5 * g[0]) + 0.3 * (r2 < 0.04)
    maps = []
    for c in range(coils):
        ctr = rng.uniform(-0.5, 0.5, 3)
        maps.append(np.exp(-sum((v - cc) ** 2 for v, cc in zip(g, ctr)) / 0.3) * np.exp(1j * rng.uniform(0, 6.28)))
    maps = np.stack(maps)
    ksp = on.nufft(maps * obj, coord).astype(np.complex64)
    dens = (0.05 + np.sum(coord.astype(np.float64) ** 2, axis=-1)).astype(np.float32)
    return ksp[None], coord, dens


def test_small_estimation_matches_oracle(torch_mod):
    from toporecon_hip.jsense import JsenseHip
    shape = (16, 16, 16)
    ksp, coord, dens = phantom_kspace(shape, 4, 600, 24, 13)
    W, calib, rounds, inner = 5, 10, 3, 4
    print("small JSENSE: 4 coils, 600 x 24 samples, W 5, calibration 10", flush=True)
    prob = oj.calibration(ksp[0].astype(np.complex128), coord, dens, mps_ker_width=W, ksp_calib_width=calib)
    h_ref = []
    _, mps_ref = oj.jsense(prob, max_iter=rounds, max_inner_iter=inner, history=h_ref)
    maps_ref = oj.to_shape(oj.maps_from_kernel(mps_ref, prob.img_shape), shape)
    s = JsenseHip(ksp[0], coord, dens, 0, mps_ker_width=W, ksp_calib_width=calib, max_iter=rounds,
                  max_inner_iter=inner)
    try:
        assert s.grd_shape == prob.grd_shape and s.img_shape == prob.img_shape and s.M == prob.coord.shape[0]
        h = []
        s.run(history=h)
        maps = s.maps(shape).cpu().numpy()
    finally:
        s.close()
    dh = max(abs(a - b) / b for a, b in zip(h, h_ref))
    e = rel(maps, maps_ref)
    print(f"objective per half-step {['%.4g' % v for v in h]}; max rel diff {dh:.2e}; maps {e:.2e}", flush=True)
    assert dh < 1e-3
    assert e < 1e-3


@pytest.mark.timeout(600)
def test_default_settings(torch_mod):
    torch = torch_mod
    from toporecon_hip.jsense import JsenseHip, estimate_sensitivity_maps
    shape = (32, 28, 30)
    ksp, coord, dens = phantom_kspace(shape, 8, 2000, 48, 14)
    print("default JSENSE: 8 coils, 2000 x 48 samples, image 32 x 28 x 30", flush=True)
    # first round (2 CG steps per half) against the oracle
    prob = oj.calibration(ksp[0].astype(np.complex128), coord, dens)
    h_ref = []
    _, mps_ref = oj.jsense(prob, max_iter=1, max_inner_iter=2, history=h_ref)
    s = JsenseHip(ksp[0], coord, dens, 0, max_iter=1, max_inner_iter=2)
    try:
        print("calibration grid", s.grd_shape, "samples", s.M, flush=True)
        assert s.grd_shape == prob.grd_shape
        h = []
        _, mps = s.run(history=h)
        e = rel(mps.cpu().numpy(), mps_ref)
    finally:
        s.close()
    dh = max(abs(a - b) / b for a, b in zip(h, h_ref))
    print(f"first alternation: objective {dh:.2e}, map kernels {e:.2e}", flush=True)
    assert dh < 1e-4 and e < 1e-4
    # the full default run
    h = []
    s = JsenseHip(ksp[0], coord, dens, 0)
    try:
        s.run(history=h)
    finally:
        s.close()
    assert all(b <= a * (1 + 1e-4) for a, b in zip(h, h[1:])), h
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    m1 = estimate_sensitivity_maps(ksp, coord, dens, image_shape_zyx=shape)
    dt = time.perf_counter() - t0
    m2 = estimate_sensitivity_maps(ksp, coord, dens, image_shape_zyx=shape)
    print(f"default run {dt:.3f} s; objective {h[0]:.4g} -> {h[-1]:.4g}", flush=True)
    assert m1.shape == (8,) + shape and m1.dtype == np.complex64
    assert np.array_equal(m1, m2)
    # estimate_shape of the trajectory is one short of the image here, so the maps pass through the
    # k-space resize; their root sum of squares is 1 up to that band limit, exactly 1 before it
    s = JsenseHip(ksp[0], coord, dens, 0, max_iter=1, max_inner_iter=1)
    try:
        s.run()
        raw = s.maps().cpu().numpy()
    finally:
        s.close()
    np.testing.assert_allclose(np.sum(np.abs(raw) ** 2, axis=0), 1.0, rtol=1e-5)


def test_dataset_entry_point(torch_mod, tmp_path):
    from toporecon_hip.data import cfl
    from toporecon_hip.data.trajectory import prepare_trajectory
    from toporecon_hip.jsense import JsenseConfig, estimate_sensitivity_from_dataset, estimate_sensitivity_maps
    shape = (16, 18, 20)  # z, y, x
    ksp, coord, dens = phantom_kspace(shape, 3, 500, 20, 15)
    raw_xyz = (coord[..., ::-1] / 40.0).astype(np.float32)  # any scale: legacy_max rescales per axis
    cfl.write_cfl(tmp_path / "ksp", ksp.reshape(1, 1, 1, 1, 1, 3, 500, 20, 1))
    cfl.write_cfl(tmp_path / "ktraj", raw_xyz.reshape(1, 1, 1, 1, 1, 1, 500, 20, 3))
    cfl.write_cfl(tmp_path / "dens", dens.reshape(1, 1, 1, 1, 1, 1, 500, 20, 1))
    (tmp_path / "imageDim.txt").write_text("20:18:16\n")
    (tmp_path / "voxelSize.txt").write_text("1.0:1.0:1.0\n")
    cfg = JsenseConfig(mps_kernel_width=4, kspace_calibration_width=8, max_iter=2, max_inner_iter=3)
    out = estimate_sensitivity_from_dataset(tmp_path, tmp_path / "work", config=cfg)
    got = cfl.read_cfl(out)
    assert got.shape == (1, 1, 1, 1, 1, 3) + shape
    t = prepare_trajectory(raw_xyz, dens, image_shape_zyx=shape, voxel_size_zyx=(1, 1, 1), readout_fraction=0.95,
                           scaling="legacy_max")
    want = estimate_sensitivity_maps(ksp[..., :t.readout], t.coordinates_zyx, t.density,
                                     image_shape_zyx=t.image_shape_zyx, config=cfg)
    assert np.array_equal(got[0, 0, 0, 0, 0], want)
