"""Shard assembly (stitch): the reference's known answers (tests/unit/test_postprocessing.py:64-114)
on a motion x echo grid of uneven shards, plus the validation paths of
src/toporecon/postprocessing/stitch.py and the single shard the reconstruct entry point writes."""

import json

import numpy as np
import pytest

from toporecon_hip.data import cfl
from toporecon_hip.mains import build_parser, cli_main, _manifest
from toporecon_hip.postprocessing import stitch


def _split(total, group, groups):
    return total * group // groups, total * (group + 1) // groups


def _write_run(d, *, stem="reconstruction", bins=5, echoes=7, mg=2, eg=3, drop_last=False, manifest_extra=None):
    """Shards of a (bins x echoes) grid cut into mg x eg uneven tiles; voxel value 100 b + e."""
    full = np.empty((bins, 1, 1, echoes, 1, 1, 2, 3, 4), np.complex64)
    for b in range(bins):
        for e in range(echoes):
            full[b, :, :, e] = 100 * b + e
    tiles = [(_split(bins, i, mg), _split(echoes, j, eg)) for i in range(mg) for j in range(eg)]
    if drop_last:
        tiles.pop()
    for (b0, b1), (e0, e1) in tiles:
        cfl.write_cfl(d / f"{stem}_e{e0}-{e1 - 1}_m{b0}-{b1 - 1}", full[b0:b1, :, :, e0:e1])
    man = {"format_version": 1, "output_stem": stem, "parameters": {"num_bins": bins},
           "distributed": {"motion_groups": mg, "echo_groups": eg}}
    if manifest_extra:
        man["parameters"].update(manifest_extra)
    (d / "run_manifest.json").write_text(json.dumps(man), encoding="utf-8")
    return full


def test_uneven_grid_custom_prefix_and_output_name(tmp_path):
    full = _write_run(tmp_path, stem="custom_prefix")
    out = stitch(tmp_path, output_name="final_image")
    assert out == (tmp_path / "final_image").resolve()
    assert np.array_equal(cfl.read_cfl(out), full)
    assert (tmp_path / "final_image.hdr").read_text() == "# Dimensions\n4 3 2 1 1 7 1 1 5\n"


def test_cli_default_output_name(tmp_path, capsys):
    full = _write_run(tmp_path, bins=2, echoes=2, mg=1, eg=1)
    assert cli_main(["stitch", str(tmp_path)]) == 0
    assert np.array_equal(cfl.read_cfl(tmp_path / "imout"), full)
    assert "imout.hdr" in capsys.readouterr().out


def test_missing_tile_rejected_without_output(tmp_path):
    _write_run(tmp_path, drop_last=True)
    with pytest.raises(ValueError, match="Expected 6 shards"):
        stitch(tmp_path)
    assert not (tmp_path / "imout.hdr").exists() and not (tmp_path / "imout.cfl").exists()


def test_output_cannot_replace_an_input_shard(tmp_path):
    _write_run(tmp_path, bins=2, echoes=2, mg=1, eg=1)
    with pytest.raises(ValueError, match="cannot overwrite"):
        stitch(tmp_path, output_name="reconstruction_e0-1_m0-1")


def test_existing_output_needs_overwrite_and_output_path(tmp_path):
    full = _write_run(tmp_path, bins=3, echoes=2, mg=3, eg=1)
    stitch(tmp_path)
    with pytest.raises(FileExistsError):
        stitch(tmp_path)
    assert np.array_equal(cfl.read_cfl(stitch(tmp_path, overwrite=True)), full)
    target = tmp_path / "elsewhere" / "img.cfl"
    assert cli_main(["stitch", str(tmp_path), "--output", str(target)]) == 0
    assert np.array_equal(cfl.read_cfl(tmp_path / "elsewhere" / "img"), full)
    with pytest.raises(ValueError, match="must be a filename"):
        stitch(tmp_path, output_name="sub/img")
    with pytest.raises(ValueError, match="end in .hdr/.cfl"):
        stitch(tmp_path, output=tmp_path / "img.npy")


def test_grid_validation(tmp_path):
    # the manifest's echo count must be covered exactly
    (tmp_path / "a").mkdir()
    _write_run(tmp_path / "a", bins=2, echoes=4, mg=1, eg=2, manifest_extra={"num_echoes": 5})
    with pytest.raises(ValueError, match="echo shard ranges do not cover 0:5"):
        stitch(tmp_path / "a")
    # motion ranges with a gap (bins 0-1 and 3-4 of 5): the count matches, the partition does not
    (tmp_path / "b").mkdir()
    x = np.zeros((2, 1, 1, 1, 1, 1, 2, 2, 2), np.complex64)
    cfl.write_cfl(tmp_path / "b" / "r_e0-0_m0-1", x)
    cfl.write_cfl(tmp_path / "b" / "r_e0-0_m3-4", x)
    (tmp_path / "b" / "run_manifest.json").write_text(json.dumps(
        {"output_stem": "r", "parameters": {"num_bins": 5}, "distributed": {"motion_groups": 2, "echo_groups": 1}}))
    with pytest.raises(ValueError, match="motion shard ranges have a gap or overlap"):
        stitch(tmp_path / "b")
    # a shard whose file name disagrees with its motion extent
    (tmp_path / "c").mkdir()
    cfl.write_cfl(tmp_path / "c" / "r_e0-0_m0-2", x)
    (tmp_path / "c" / "run_manifest.json").write_text(json.dumps(
        {"output_stem": "r", "parameters": {"num_bins": 3}, "distributed": {"motion_groups": 1, "echo_groups": 1}}))
    with pytest.raises(ValueError, match="Motion range does not match"):
        stitch(tmp_path / "c")
    # the prefix may not leave the run directory; no manifest / no shards
    with pytest.raises(ValueError, match="stay inside INPUT_DIR"):
        stitch(tmp_path / "c", stem="../b/r")
    (tmp_path / "d").mkdir()
    with pytest.raises(FileNotFoundError, match="Missing reconstruction manifest"):
        stitch(tmp_path / "d")
    with pytest.raises(FileNotFoundError, match="No shards found"):
        stitch(tmp_path / "c", stem="nothing")


def test_reconstruct_shard_and_manifest_stitch(tmp_path):
    """The single shard and manifest the reconstruct entry point writes assemble to the same image."""
    B, E = 3, 2
    img = (np.arange(B * E * 4 * 5 * 6) * (1 + 0.5j)).astype(np.complex64).reshape(B, E, 4, 5, 6)
    out = np.zeros((B, 1, 1, E, 1, 1, 4, 5, 6), np.complex64)
    out[:, 0, 0, :, 0, 0] = img
    args = build_parser("tvme").parse_args([str(tmp_path), "recon", "--num-bins", str(B)])
    cfl.write_cfl(tmp_path / f"recon_e0-{E - 1}_m0-{B - 1}", out)
    (tmp_path / "run_manifest.json").write_text(json.dumps(_manifest(args, "tvme", E, str(tmp_path))))
    got = cfl.read_cfl(stitch(tmp_path))
    assert np.array_equal(got, out)
