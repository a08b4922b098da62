"""Shard formats, input conditioning and the reconstruct entry point (CPU)."""

import json
import os

import numpy as np
import pytest

from toporecon_hip import conditioning as cond
from toporecon_hip import synth
from toporecon_hip.data import cfl
from toporecon_hip.mains import build_parser, load_conditioned, _manifest


def test_cfl_round_trip_and_reversed_header(tmp_path):
    a = (np.arange(60, dtype=np.float32).reshape(3, 4, 5) - 7j).astype(np.complex64)
    cfl.write_cfl(tmp_path / "arr", a)
    assert (tmp_path / "arr.hdr").read_text() == "# Dimensions\n5 4 3\n"
    assert cfl.read_cfl_header(tmp_path / "arr.cfl") == (3, 4, 5)
    assert np.array_equal(cfl.read_cfl(tmp_path / "arr"), a)
    assert np.array_equal(cfl.read_cfl(tmp_path / "arr", mmap=True), a)
    assert (tmp_path / "arr.cfl").stat().st_size == 60 * 8
    (tmp_path / "arr.cfl").write_bytes(b"\0" * 8)
    with pytest.raises(ValueError, match="payload size mismatch"):
        cfl.read_cfl(tmp_path / "arr")


def test_kspace_shard_and_maps_readers(tmp_path):
    E, C, T, R = 3, 4, 6, 5
    canon = (np.arange(E * C * T * R) + 1j).astype(np.complex64).reshape(E, C, T, R)
    cfl.write_cfl(tmp_path / "ksp", canon.reshape(1, 1, 1, E, 1, C, T, R, 1))
    assert cfl.kspace_layout(cfl.read_cfl_header(tmp_path / "ksp")) == (E, C, T, R)
    got = cfl.read_kspace_shard(tmp_path / "ksp", echo_start=1, echo_end=3, coil_indices=[3, 0],
                                trajectory_indices=[5, 2, 4], readout=2)
    assert np.array_equal(got, canon[1:3][:, [3, 0]][:, :, [5, 2, 4]][..., :2])
    with pytest.raises(IndexError):
        cfl.read_kspace_shard(tmp_path / "ksp", echo_start=0, echo_end=1, coil_indices=[4])
    with pytest.raises(ValueError):
        cfl.read_kspace_shard(tmp_path / "ksp", echo_start=2, echo_end=2, coil_indices=[0])
    with pytest.raises(ValueError, match="non-singleton"):
        cfl.kspace_layout((2, 1, 1, E, 1, C, T, R, 1))
    maps = (np.arange(C * 2 * 3 * 4) * 1j).astype(np.complex64).reshape(C, 2, 3, 4)
    cfl.write_cfl(tmp_path / "mps", maps.reshape(1, 1, 1, 1, 1, C, 2, 3, 4))
    assert np.array_equal(cfl.read_mps_coils(tmp_path / "mps", coil_indices=[2, 1]), maps[[2, 1]])


def _small_cfg(algorithm="tvme"):
    cfg = synth.config_scaled(synth.config_small_tvm(), (10, 12, 14), coils=3, echoes=2, bins=3,
                              interleaves=80, readouts=24, seed=4)
    cfg.algorithm = algorithm
    return cfg


def _write_dataset(root, algorithm="tvme", voxel=(2.0, 1.5, 1.0)):
    from oracle import nufft as on
    cfg = _small_cfg(algorithm)
    acq, ksp, coord, dens, resp = synth.make_raw(cfg, lambda x, c: on.nufft(x, c))
    synth.write_raw_dataset(acq, ksp, coord, dens, resp, root, voxel_size=voxel)
    return cfg, acq, ksp, coord, dens, resp


@pytest.mark.parametrize("algorithm", ["tvm", "tvme", "tvmw"])
def test_conditioning_matches_generator(tmp_path, algorithm):
    cfg, acq, ksp, coord, dens, resp = _write_dataset(tmp_path / "raw", algorithm)
    args = build_parser(algorithm).parse_args([str(tmp_path / "raw"), "img", "--num-bins", "3"])
    d = load_conditioned(args, algorithm)
    keep = (cond.drop_respiratory_outliers(resp) if algorithm != "tvmw"
            else np.arange(resp.size))
    nro = int(0.98 * cfg.readouts)
    edges = cond.bin_edges(resp[keep], 3)
    assert np.allclose(d["bin_edges"], edges)
    mask = np.zeros(keep.size, bool)
    for b in range(3):
        mask |= cond.bin_members(resp[keep], edges, b)
    tr = keep[mask]
    # coordinates survive the x,y,z / voxel*n round trip through the file to float32 rounding
    assert np.allclose(d["coord"], coord[tr, :nro], rtol=0, atol=2e-5)
    assert d["ksp"].shape == (2, 3, tr.size, nro)
    assert np.abs(d["ksp"]).max() == pytest.approx(1.0)
    assert np.allclose(np.sum(np.abs(d["mps"]) ** 2, axis=0), 1.0, atol=1e-5)
    assert d["dcf"].max() == pytest.approx(1.0)
    if algorithm == "tvme":
        assert np.allclose(d["coord"], acq.coord[mask], atol=2e-5)
        assert np.allclose(d["dcf"], acq.dcf[mask])


def test_coordinate_scaling_known_answer(tmp_path):
    """Raw [x, y, z] = [1, 2, 3] with voxel (z,y,x) = (0.5, 0.25, 0.1), shape (10, 20, 30) -> (15, 10, 3)."""
    root = tmp_path / "raw"
    T, R = 4, 8
    cfl.write_cfl(root / "ktraj", np.tile(np.array([1.0, 2.0, 3.0], np.complex64), (T, R, 1)))
    cfl.write_cfl(root / "dens", np.ones((T, R), np.complex64))
    cfl.write_cfl(root / "resp", np.linspace(0, 1, T).astype(np.complex64)[:, None])
    cfl.write_cfl(root / "ksp", np.ones((1, 1, 1, 1, 1, 1, T, R, 1), np.complex64))
    cfl.write_cfl(root / "mps", np.ones((1, 1, 1, 1, 1, 1, 10, 20, 30), np.complex64))
    (root / "imageDim.txt").write_text("30:20:10\n")
    (root / "voxelSize.txt").write_text("0.1:0.25:0.5\n")
    args = build_parser("tvmw").parse_args([str(root), "img", "--num-bins", "1", "--frac", "1.0"])
    d = load_conditioned(args, "tvmw")
    assert np.allclose(d["coord"][0, 0], (15.0, 10.0, 3.0))
    assert d["coord"].shape[0] == T - 1  # the maximum respiratory value falls in no bin


def test_manifest_keys():
    args = build_parser("tvme").parse_args(["in", "stem", "--l2-coupling"])
    m = _manifest(args, "tvme", 6, "in")
    assert m["algorithm"] == "tvme" and m["nufft_backend"] == "hip"
    assert m["output_pattern"] == "stem_e{e0}-{e1}_m{b0}-{b1}"
    assert m["parameters"]["lambda_echo"] == 1e-5 and m["parameters"]["l2_coupling"] is True
    assert json.loads(json.dumps(m)) == m
