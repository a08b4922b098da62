"""Input conditioning of the three algorithms on hand-built raw datasets with known answers.

Each dataset is written with values whose conditioned result is fixed by hand:

* respiratory trace r_i = i (i = 0..99): the 2 % / 98 % percentiles are 1.98 and 97.02, so TVM and
  TVME keep r in 2..97, then the first int(100 / acc) = 50 of those (acc = 2): readouts 2..51
  (tvm.py:755-763, tvme.py:745-757).  TVMW keeps every readout (tvmw.py:941-951);
* bin edges are the percentiles of the kept trace at 0, 20, .., 100 %: for r = 2..51 that is
  2 + 49 k / 5, and membership edge_b <= r < edge_{b+1} leaves the maximum (51) in no bin
  (tvm.py:857-877); for r = 0..99 (TVMW) the maximum 99 is left out;
* readout fraction 0.98 of 10 readouts = 9, and a corrupt value of 5000 in readouts 7 and 8 of a
  kept interleave shrinks it to 7 (the `while coord.max() > 1000` loop, tvm.py:769-772);
* coordinates: the stored [x, y, z] trajectory is reversed to [z, y, x] and scaled by voxel size
  times matrix size per axis (image 8:6:4 X:Y:Z, voxel 1:2:0.5 -> factors 2, 12, 8);
* density: 10 on the interleave that carries the maximum kept respiratory value, 4 on one
  interleave inside the bins, 1 elsewhere.  TVM divides by the maximum over the kept readouts (10)
  and again after bin selection (0.4), so the inside-bin values end as 1 and 0.25 (tvm.py:803-805,
  954-959); TVME divides once (0.4 and 0.1, tvme.py:802-804); TVMW divides once after selection
  (tvmw.py:1129-1133);
* k-space: its largest magnitude over the selected data is |8 + 6j| = 10 (a larger value sits on
  a readout the conditioning drops), so the sample becomes 0.8 + 0.6j;
* coil maps: at one voxel both coils are 0 (stays 0), at one the coils are 3 and 4j (0.6, 0.8j),
  and at one the root-sum-of-squares is 1e-13: TVM / TVMW divide by it (1.0), TVME floors it at
  1e-12 (0.1) (tvm.py:945-950, tvme.py:943-948, tvmw.py:1122-1127).
"""

import numpy as np
import pytest

from toporecon_hip.data import cfl
from toporecon_hip.mains import build_parser, load_conditioned

NTR, NRO = 100, 10
SHAPE = (4, 6, 8)  # Z, Y, X


def _write(root, dcf_hot):
    rng = np.random.default_rng(3)
    raw = rng.uniform(-0.45, 0.45, (NTR, NRO, 3)).astype(np.float32)  # [x, y, z]
    raw[20, 7:9, :] = 5000.0  # corrupt tail of a kept interleave
    dens = np.ones((NTR, NRO), np.float32)
    for i, v in dcf_hot.items():
        dens[i] = v
    ksp = np.ones((1, 2, NTR, NRO), np.complex64)
    ksp[0, 1, 30, 3] = 8 + 6j
    ksp[0, 0, 99, 2] = 100.0         # the maximum respiratory value: dropped by every algorithm
    ksp[0, 0, 30, 8] = 50.0          # readout index 8 of 10: beyond the shrunk readout count
    mps = np.ones((2,) + SHAPE, np.complex64)
    mps[:, 0, 0, 0] = 0
    mps[0, 0, 0, 1], mps[1, 0, 0, 1] = 3, 4j
    mps[0, 0, 0, 2], mps[1, 0, 0, 2] = 1e-13, 0
    cfl.write_cfl(root / "ksp", ksp.reshape(1, 1, 1, 1, 1, 2, NTR, NRO, 1))
    cfl.write_cfl(root / "ktraj", raw.reshape(1, 1, 1, 1, 1, 1, NTR, NRO, 3))
    cfl.write_cfl(root / "dens", dens.reshape(1, 1, 1, 1, 1, 1, NTR, NRO, 1))
    cfl.write_cfl(root / "mps", mps.reshape((1, 1, 1, 1, 1, 2) + SHAPE))
    cfl.write_cfl(root / "resp", np.arange(NTR, dtype=np.float32).reshape(-1, 1))
    (root / "imageDim.txt").write_text("8:6:4\n")
    (root / "voxelSize.txt").write_text("1.0:2.0:0.5\n")
    (root / "tr.txt").write_text("5.0\n")
    return raw


def _run(tmp_path, alg, dcf_hot, acc=True):
    raw = _write(tmp_path, dcf_hot)
    argv = [str(tmp_path), "recon", "--num-bins", "5"]
    if acc and alg != "tvmw":
        argv += ["--acceleration", "2"]
    return load_conditioned(build_parser(alg).parse_args(argv), alg), raw


@pytest.mark.parametrize("alg", ["tvm", "tvme"])
def test_tvm_tvme_trace_bins_readouts_and_coordinates(tmp_path, alg):
    d, raw = _run(tmp_path, alg, {51: 10.0, 30: 4.0})
    # kept 2..51, bins of 2..51 with edges 2, 11.8, 21.6, 31.4, 41.2, 51: readouts 2..50 selected
    np.testing.assert_allclose(d["bin_edges"], [2.0, 11.8, 21.6, 31.4, 41.2, 51.0], rtol=0, atol=1e-12)
    assert np.array_equal(d["resp"], np.arange(2, 51, dtype=np.float32))
    assert d["coord"].shape == (49, 7, 3) and d["ksp"].shape == (1, 2, 49, 7)
    want = raw[2:51, :7, ::-1] * np.array([2.0, 12.0, 8.0], np.float32)
    np.testing.assert_array_equal(d["coord"], want.astype(np.float32))
    assert d["img_shape"] == [4, 6, 8]


@pytest.mark.parametrize("alg, inside_hot, inside", [("tvm", 1.0, 0.25), ("tvme", 0.4, 0.1)])
def test_density_normalisation_tvm_twice_tvme_once(tmp_path, alg, inside_hot, inside):
    d, _ = _run(tmp_path, alg, {51: 10.0, 30: 4.0})
    dcf = d["dcf"]
    row = list(range(2, 51)).index(30)
    np.testing.assert_allclose(dcf[row], inside_hot, rtol=1e-7)
    others = np.delete(dcf, row, axis=0)
    np.testing.assert_allclose(others, inside, rtol=1e-7)


def test_tvmw_keeps_every_readout_and_normalises_density_after_selection(tmp_path):
    d, raw = _run(tmp_path, "tvmw", {99: 10.0, 30: 4.0})
    np.testing.assert_allclose(d["bin_edges"], [0.0, 19.8, 39.6, 59.4, 79.2, 99.0], rtol=0, atol=1e-12)
    assert np.array_equal(d["resp"], np.arange(0, 99, dtype=np.float32))
    assert d["coord"].shape == (99, 7, 3)
    np.testing.assert_array_equal(d["coord"], (raw[:99, :7, ::-1] * np.array([2.0, 12.0, 8.0], np.float32)))
    np.testing.assert_allclose(d["dcf"][30], 1.0, rtol=1e-7)          # 4 / 4: the 10 is not selected
    np.testing.assert_allclose(np.delete(d["dcf"], 30, axis=0), 0.25, rtol=1e-7)


@pytest.mark.parametrize("alg", ["tvm", "tvme", "tvmw"])
def test_kspace_maximum_over_selected_data(tmp_path, alg):
    d, _ = _run(tmp_path, alg, {})
    ksp = d["ksp"]
    assert float(np.max(np.abs(ksp))) == pytest.approx(1.0, rel=1e-7)
    row = list(d["resp"]).index(30.0)
    np.testing.assert_allclose(ksp[0, 1, row, 3], 0.8 + 0.6j, rtol=1e-6)
    np.testing.assert_allclose(ksp[0, 0, row, 0], 0.1, rtol=1e-6)


@pytest.mark.parametrize("alg, tiny", [("tvm", 1.0), ("tvme", 0.1), ("tvmw", 1.0)])
def test_coil_map_rss_zero_and_floor(tmp_path, alg, tiny):
    d, _ = _run(tmp_path, alg, {})
    m = d["mps"]
    assert m[0, 0, 0, 0] == 0 and m[1, 0, 0, 0] == 0
    np.testing.assert_allclose([m[0, 0, 0, 1], m[1, 0, 0, 1]], [0.6, 0.8j], rtol=1e-6)
    np.testing.assert_allclose(m[0, 0, 0, 2], tiny, rtol=1e-5)
    np.testing.assert_allclose(np.abs(m[:, 1, 1, 1]), 1 / np.sqrt(2), rtol=1e-6)
