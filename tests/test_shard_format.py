"""The output shard the reconstruct entry point writes is the reference's shard format.

The reference's ``toporecon stitch`` (src/toporecon/postprocessing/stitch.py:56-161) assembles the
shards unchanged, so this build has no stitch of its own; what it must guarantee is what stitch
checks: the file name ``{stem}_e{e0}-{e1}_m{b0}-{b1}`` with inclusive ranges (stitch.py:56-86), a
9-D ``(B, 1, 1, E, 1, 1, Z, Y, X)`` complex64 CFL whose singleton axes 1, 2, 4, 5 are 1 and whose
motion / echo extents match the name (stitch.py:147-159), the BART header with reversed
dimensions, and a ``run_manifest.json`` whose ``parameters.num_bins``, ``parameters.num_echoes``,
``distributed.motion_groups`` and ``distributed.echo_groups`` describe a grid the shards cover
exactly (stitch.py:100-146).  The checks below restate those rules independently and run them on
what ``write_output`` produced.
"""

import json
import re

import numpy as np
import pytest

from toporecon_hip.data import cfl
from toporecon_hip.mains import _manifest, build_parser, write_output

SHARD = re.compile(r"^(?P<stem>.+)_e(?P<e0>\d+)-(?P<e1>\d+)_m(?P<m0>\d+)-(?P<m1>\d+)\.hdr$")


def _check_run(d):
    man = json.loads((d / "run_manifest.json").read_text())
    stem = man["output_stem"]
    bins = int(man["parameters"]["num_bins"])
    mg, eg = int(man["distributed"]["motion_groups"]), int(man["distributed"]["echo_groups"])
    shards = []
    for p in sorted(d.glob(f"{stem}_e*_m*.hdr")):
        m = SHARD.match(p.name)
        assert m and m["stem"] == stem
        e0, e1, m0, m1 = (int(m[k]) for k in ("e0", "e1", "m0", "m1"))
        base = p.with_suffix("")
        shape = cfl.read_cfl_header(base)
        cfl.validate_cfl_storage(base, shape)
        assert len(shape) == 9 and all(shape[a] == 1 for a in (1, 2, 4, 5))
        assert shape[0] == m1 - m0 + 1 and shape[3] == e1 - e0 + 1
        shards.append((m0, m1 + 1, e0, e1 + 1, shape))
    assert len(shards) == mg * eg
    echoes = int(man["parameters"].get("num_echoes") or max(s[3] for s in shards))
    # the motion and echo ranges partition [0, bins) and [0, echoes)
    for lo_i, hi_i, total in ((0, 1, bins), (2, 3, echoes)):
        cursor = 0
        for lo, hi in sorted({(s[lo_i], s[hi_i]) for s in shards}):
            assert lo == cursor and hi > lo
            cursor = hi
        assert cursor == total
    return man, shards


@pytest.mark.parametrize("alg", ["tvm", "tvme", "tvmw"])
def test_written_shard_is_what_stitch_accepts(tmp_path, alg):
    B, E = 3, 2
    rng = np.random.default_rng(5)
    img = (rng.standard_normal((B, E, 4, 5, 6)) + 1j * rng.standard_normal((B, E, 4, 5, 6))).astype(np.complex64)
    args = build_parser(alg).parse_args([str(tmp_path), "recon", "--num-bins", str(B), "--nufft-backend", "hip"])
    base = write_output(img, tmp_path, "recon", _manifest(args, alg, E, str(tmp_path)))
    assert base.endswith(f"recon_e0-{E - 1}_m0-{B - 1}")
    man, shards = _check_run(tmp_path)
    assert man["algorithm"] == alg and man["nufft_backend"] == "hip"
    assert shards == [(0, B, 0, E, (B, 1, 1, E, 1, 1, 4, 5, 6))]
    assert (tmp_path / f"recon_e0-{E - 1}_m0-{B - 1}.hdr").read_text().splitlines()[1] == "6 5 4 1 1 2 1 1 3"
    got = cfl.read_cfl(base)
    assert got.dtype == np.complex64 and np.array_equal(got[:, 0, 0, :, 0, 0], img)


def test_crop_to_original_keeps_the_format(tmp_path):
    img = np.ones((2, 1, 6, 8, 10), np.complex64)
    args = build_parser("tvme").parse_args([str(tmp_path), "recon", "--num-bins", "2"])
    base = write_output(img, tmp_path, "recon", _manifest(args, "tvme", 1, str(tmp_path)), crop_zyx=(4, 6, 8))
    _, shards = _check_run(tmp_path)
    assert shards[0][4] == (2, 1, 1, 1, 1, 1, 4, 6, 8)
    assert cfl.read_cfl(base).shape == (2, 1, 1, 1, 1, 1, 4, 6, 8)


def test_no_stitch_subcommand():
    from toporecon_hip.mains import cli_main
    with pytest.raises(SystemExit):
        cli_main(["stitch", "somewhere"])
