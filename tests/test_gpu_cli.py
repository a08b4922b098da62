"""`reconstruct` end to end on the GPU: raw dataset files in, reference-format shard + manifest out."""

import json
import os
import subprocess
import sys

import numpy as np
import pytest

from oracle import pdhg as op
from problems import rel
from test_formats import _write_dataset
from toporecon_hip.data import cfl
from toporecon_hip.mains import build_parser, load_conditioned, run_main

pytestmark = pytest.mark.gpu

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def torch_mod():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


@pytest.mark.parametrize("algorithm", ["tvm", "tvme", "tvmw"])
def test_reconstruct_writes_reference_shard(torch_mod, tmp_path, algorithm):
    print("cli", algorithm, flush=True)
    _write_dataset(tmp_path / "raw", algorithm)
    out = tmp_path / "out"
    lam = {"tvm": ["--lamda", "2e-3"], "tvme": ["--lamda_m", "2e-3", "--lamda_e", "2e-3"],
           "tvmw": ["--lambda1", "2e-3", "--lambda2", "2e-3", "--lambda3", "2e-3"]}[algorithm]
    argv = [str(tmp_path / "raw"), "img", "--num-bins", "3", "--max-iter", "4", "--tol", "0",
            "--l2-coupling", "--output-dir", str(out)] + lam
    assert run_main(algorithm, argv) == 0
    shard = out / "img_e0-1_m0-2"
    img = cfl.read_cfl(shard)
    assert img.shape == (3, 1, 1, 2, 1, 1, 10, 12, 14)
    man = json.loads((out / "run_manifest.json").read_text())
    assert man["algorithm"] == algorithm and man["nufft_backend"] == "hip"
    # the same conditioned inputs through the float64 oracle
    args = build_parser(algorithm).parse_args(argv)
    d = load_conditioned(args, algorithm)
    prob = op.bin_problem(d["ksp"], d["coord"], d["dcf"], d["mps"], d["resp"], d["bin_edges"])
    kw = dict(lam_m=2e-3, lam_e=2e-3, lam1=2e-3, lam2=2e-3, lam3=2e-3, l2_coupling=True, max_iter=4, tol=0.0)
    o_ubar, _, _ = op.run(prob, algorithm, **kw)
    assert rel(img[:, 0, 0, :, 0, 0], o_ubar) < 3e-5


def test_module_cli_forwards_arguments(torch_mod, tmp_path):
    _write_dataset(tmp_path / "raw", "tvm")
    out = tmp_path / "out"
    env = dict(os.environ)
    env["PYTHONPATH"] = os.pathsep.join([os.path.join(REPO, "multinodemri-miccai26_amd"),
                                         env.get("PYTHONPATH", "")])
    r = subprocess.run([sys.executable, "-m", "toporecon_hip", "reconstruct", "--algorithm", "tvm", "--",
                        str(tmp_path / "raw"), "img", "--num_bins", "3", "--max_iter", "2",
                        "--output-dir", str(out)], capture_output=True, text=True, env=env, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    assert (out / "img_e0-1_m0-2.cfl").exists()
    with pytest.raises(RuntimeError, match="SigPy is required"):
        run_main("tvm", [str(tmp_path / "raw"), "img", "--nufft-backend", "sigpy"])


def test_tvme_cli_learned_warm_start(torch_mod, tmp_path):
    """--warm-start learned (TVME): same shard format, the manifest says so, and the image
    agrees with the plain solve to the stopping tolerance's accuracy."""
    _write_dataset(tmp_path / "raw", "tvme")
    base = [str(tmp_path / "raw"), "img", "--num-bins", "3", "--max-iter", "200", "--tol", "1e-3",
            "--l2-coupling", "--lamda_m", "2e-3", "--lamda_e", "2e-3"]
    assert run_main("tvme", base + ["--output-dir", str(tmp_path / "plain")]) == 0
    assert run_main("tvme", base + ["--output-dir", str(tmp_path / "warm"), "--warm-start", "learned"]) == 0
    plain = cfl.read_cfl(tmp_path / "plain" / "img_e0-1_m0-2")
    warm = cfl.read_cfl(tmp_path / "warm" / "img_e0-1_m0-2")
    assert warm.shape == plain.shape
    # both stop on the relative-change rule; compare what they minimise (float64 oracle objective)
    args = build_parser("tvme").parse_args(base)
    d = load_conditioned(args, "tvme")
    prob = op.bin_problem(d["ksp"], d["coord"], d["dcf"], d["mps"], d["resp"], d["bin_edges"])
    kw = dict(lam_m=2e-3, lam_e=2e-3, l2_coupling=True)
    phi_p = op.objective(prob, plain[:, 0, 0, :, 0, 0].astype(np.complex128), "tvme", **kw)
    phi_w = op.objective(prob, warm[:, 0, 0, :, 0, 0].astype(np.complex128), "tvme", **kw)
    print("objective plain", phi_p, "warm", phi_w, "image rel diff", rel(warm, plain), flush=True)
    # On this 10x12x14 fixture the relative-change rule stops both runs early, so only a coarse
    # agreement is asserted here; test_learned_warm_start_gains_over_plain checks the 0.1 %
    # objective bound on a realistic held-out problem.
    assert abs(phi_w - phi_p) / phi_p < 3e-2
    man = json.loads((tmp_path / "warm" / "run_manifest.json").read_text())
    assert man["parameters"]["warm_start"] == "learned"
    assert man["parameters"]["warm_start_cg_steps"] == 4
    assert "warm_start" not in json.loads((tmp_path / "plain" / "run_manifest.json").read_text())["parameters"]
