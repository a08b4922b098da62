"""`reconstruct` entry points with the reference's arguments, conditioning and output shards.

Single-process restatement of ``main`` in src/toporecon/algorithms/tvm.py:494-1088,
tvme.py:578-1053 and tvmw.py:789-1240 with the HIP reconstructors: the same
flags (kebab-case and legacy spellings), the same raw-dataset files (``ksp``,
``ktraj``, ``dens``, ``imageDim.txt``, ``voxelSize.txt``; ``mps`` and ``resp``
from ``--prepared-dir``), the same input conditioning (a17), and the same
output shard ``{stem}_e{e0}-{e1}_m{b0}-{b1}`` (9-D ``(B,1,1,E,1,1,Z,Y,X)`` CFL)
plus ``run_manifest.json`` (a18).  ``--nufft-backend`` accepts ``hip`` (the
default here) beside ``sigpy``.
"""

from __future__ import annotations

import argparse
import json
import logging
import os
import time
from pathlib import Path

import numpy as np

from . import conditioning as cond
from .data import cfl

ALGORITHMS = ("tvm", "tvme", "tvmw")


def build_parser(algorithm: str) -> argparse.ArgumentParser:
    p = argparse.ArgumentParser(description=f"{algorithm.upper()} PDHG reconstruction (HIP hot path).")
    p.add_argument("--readout-fraction", "--frac", dest="frac", type=float, default=0.98)
    p.add_argument("--num-bins", "--num_bins", dest="num_bins", type=int, default=6)
    if algorithm == "tvm":
        p.add_argument("--lambda-motion", "--lamda", dest="lamda", type=float, default=1e-5)
    elif algorithm == "tvme":
        p.add_argument("--lambda-motion", "--lamda_m", dest="lamda_m", type=float, default=1e-5)
        p.add_argument("--lambda-echo", "--lamda_e", dest="lamda_e", type=float, default=1e-5)
    else:
        p.add_argument("--lambda-motion", "--lambda1", dest="lambda1", type=float, default=1e-5)
        p.add_argument("--lambda-echo-wavelet", "--lambda2", dest="lambda2", type=float, default=1e-5)
        p.add_argument("--lambda-spatial-wavelet", "--lambda3", dest="lambda3", type=float, default=1e-5)
    p.add_argument("--max-iter", "--max_iter", dest="max_iter", type=int, default=300)
    p.add_argument("--tol", type=float, default=1e-3)
    if algorithm in ("tvm", "tvme"):
        p.add_argument("--acceleration", "--acc", dest="acc", type=int, default=1)
    p.add_argument("--fov-scale", "--fov_scale", dest="fov_scale", type=float, nargs=3,
                   default=[1.0, 1.0, 1.0])
    p.add_argument("--crop-to-original", "--crop_to_orig", dest="crop_to_orig", action="store_true")
    p.add_argument("--echo-groups", "--echo_groups", dest="echo_groups", type=int, default=1)
    p.add_argument("--motion-groups", "--motion_groups", dest="motion_groups", type=int, default=1)
    p.add_argument("--l2-coupling", "--l2_coupling", dest="l2_coupling", action="store_true")
    p.add_argument("--show-progress", "--show_pbar", dest="show_pbar", action="store_true")
    p.add_argument("--device", type=int, default=0)
    p.add_argument("--multi-gpu", "--multi_gpu", dest="multi_gpu", action="store_true")
    if algorithm == "tvm":
        p.add_argument("--reduce-per-bin", "--reduce_per_bin", dest="reduce_per_bin", action="store_true")
    p.add_argument("--fp16-communication", "--use_fp16_comm", dest="use_fp16_comm", action="store_true")
    p.add_argument("--nufft-backend", choices=("sigpy", "hip"), default="hip")
    if algorithm == "tvme":
        # not a reference flag: start PDHG from the trained CNN's estimate (same tolerance)
        p.add_argument("--warm-start", choices=("none", "learned"), default="none")
        p.add_argument("--warm-start-weights", type=str, default=None)
        p.add_argument("--warm-start-cg-steps", type=int, default=None,
                       help="conjugate-gradient data-consistency steps after the CNN (default 4)")
    p.add_argument("--prepared-dir", type=str, default=None)
    p.add_argument("--output-dir", type=str, default=".")
    p.add_argument("--verbose", action="store_true")
    p.add_argument("input_dir", type=str)
    p.add_argument("img_file", type=str)
    return p


def center_crop_zyx(arr, out_shape_zyx):
    """Centre crop of the last three axes (tvme.py:45-54)."""
    Nz0, Ny0, Nx0 = out_shape_zyx
    Nz, Ny, Nx = arr.shape[-3:]
    if Nz0 > Nz or Ny0 > Ny or Nx0 > Nx:
        raise ValueError(f"Requested crop {out_shape_zyx} larger than input {(Nz, Ny, Nx)}")
    z0, y0, x0 = (Nz - Nz0) // 2, (Ny - Ny0) // 2, (Nx - Nx0) // 2
    return arr[..., z0:z0 + Nz0, y0:y0 + Ny0, x0:x0 + Nx0]


def load_conditioned(args, algorithm):
    """Read the raw + prepared files and condition them like the reference main."""
    prepared = args.prepared_dir or args.input_dir
    img_shape = [int(v) for v in cfl.read_zyx_txt(os.path.join(args.input_dir, "imageDim.txt"))]
    orig_img_shape = list(img_shape)
    voxel = cfl.read_zyx_txt(os.path.join(args.input_dir, "voxelSize.txt"))
    coord_full = np.squeeze(cfl.read_cfl(os.path.join(args.input_dir, "ktraj"))).real[..., ::-1]
    dcf_full = np.squeeze(cfl.read_cfl(os.path.join(args.input_dir, "dens"))).real
    resp_full = np.squeeze(cfl.read_cfl(os.path.join(prepared, "resp"))).real.reshape(-1)
    if algorithm in ("tvm", "tvme"):
        tr_keep = cond.drop_respiratory_outliers(resp_full, args.acc)
    else:
        tr_keep = np.arange(resp_full.size, dtype=np.int64)
    num_ro = cond.readout_count(args.frac, coord_full.shape[-2])
    coord = np.array(coord_full[tr_keep, :num_ro])
    dcf = np.array(dcf_full[tr_keep, :num_ro])
    resp = resp_full[tr_keep]
    while coord.max() > 1000:
        num_ro -= 1
        coord = coord[:, :num_ro]
        dcf = dcf[:, :num_ro]
    for d in range(coord.shape[-1]):
        coord[..., d] *= voxel[d] * img_shape[d]
    fov = np.asarray(args.fov_scale, dtype=float)
    if not np.allclose(fov, 1.0):
        if np.any(fov < 1.0):
            raise ValueError("FoV scale factors must be >= 1.0 for all axes (z y x).")
        coord = coord.copy()
        for d in range(3):
            coord[..., d] *= fov[d]
        img_shape = [int(img_shape[d] * fov[d]) for d in range(3)]
    if algorithm in ("tvm", "tvme"):
        dcf = cond.normalize_dcf(dcf)
    edges = cond.bin_edges(resp, args.num_bins)
    mask = np.zeros(resp.shape, dtype=bool)
    for b in range(args.num_bins):
        mask |= cond.bin_members(resp, edges, b)
    idx = np.nonzero(mask)[0].astype(np.int64)
    tr_idx = tr_keep[idx]
    coord, dcf, resp = coord[idx], dcf[idx], resp[idx]
    if algorithm in ("tvm", "tvmw"):
        dcf = cond.normalize_dcf(dcf)
    kfile = os.path.join(args.input_dir, "ksp")
    E_total = cfl.kspace_layout(cfl.read_cfl_header(kfile))[0]
    C_total = int(cfl.read_cfl_header(os.path.join(prepared, "mps"))[-4])
    ksp = cfl.read_kspace_shard(kfile, echo_start=0, echo_end=E_total, coil_indices=range(C_total),
                                readout=num_ro, trajectory_indices=tr_idx)
    mps = cfl.read_mps_coils(os.path.join(prepared, "mps"), coil_indices=range(C_total))
    if tuple(mps.shape[-3:]) != tuple(img_shape):
        raise ValueError(f"MPS grid {tuple(mps.shape[-3:])} does not match the requested "
                         f"reconstruction grid {tuple(img_shape)}. Prepare sensitivity maps with the "
                         f"same FoV scale.")
    ksp = cond.normalize_kspace(ksp)
    mps = cond.normalize_maps(mps, algorithm)
    return dict(ksp=ksp, coord=coord.astype(np.float32), dcf=dcf.astype(np.float32), resp=resp, mps=mps,
                bin_edges=edges, E_total=E_total, img_shape=img_shape, orig_img_shape=orig_img_shape,
                fov=fov, prepared=prepared)


def _cg_steps(args):
    if args.warm_start_cg_steps is not None:
        if args.warm_start_cg_steps < 0:
            raise ValueError("--warm-start-cg-steps must be >= 0")
        return args.warm_start_cg_steps
    from .learned.warm import DEFAULT_CG_STEPS
    return DEFAULT_CG_STEPS


def _manifest(args, algorithm, E_total, prepared):
    params = {"readout_fraction": args.frac, "num_bins": args.num_bins, "num_echoes": E_total,
              "max_iter": args.max_iter, "tol": args.tol, "fov_scale_zyx": list(args.fov_scale),
              "crop_to_original": args.crop_to_orig, "l2_coupling": args.l2_coupling}
    if algorithm == "tvm":
        params.update(lambda_motion=args.lamda, acceleration=args.acc, reduce_per_bin=args.reduce_per_bin)
        density = "maximum_then_global_maximum_after_motion_sharding"
    elif algorithm == "tvme":
        params.update(lambda_motion=args.lamda_m, lambda_echo=args.lamda_e, acceleration=args.acc)
        if getattr(args, "warm_start", "none") != "none":  # absent in the reference's manifest
            params.update(warm_start=args.warm_start, warm_start_cg_steps=_cg_steps(args))
        density = "maximum"
    else:
        params.update(lambda_motion=args.lambda1, lambda_echo_wavelet=args.lambda2,
                      lambda_spatial_wavelet=args.lambda3)
        density = "global_maximum_after_motion_sharding"
    return {
        "format_version": 1,
        "algorithm": algorithm,
        "nufft_backend": args.nufft_backend,
        "input_directory": os.path.abspath(args.input_dir),
        "prepared_directory": os.path.abspath(prepared),
        "output_stem": args.img_file,
        "output_pattern": f"{args.img_file}_e{{e0}}-{{e1}}_m{{b0}}-{{b1}}",
        "parameters": params,
        "distributed": {"world_size": 1, "num_nodes": 1, "motion_groups": args.motion_groups,
                        "echo_groups": args.echo_groups, "fp16_communication": args.use_fp16_comm},
        "normalization": {"kspace": "global_max_magnitude", "mps": "coil_root_sum_of_squares",
                          "density": density},
    }


def run_main(algorithm: str, argv=None) -> int:
    t_main = time.perf_counter()
    args = build_parser(algorithm).parse_args(argv)
    if args.verbose:
        logging.basicConfig(level=logging.DEBUG)
    if args.nufft_backend != "hip":
        raise RuntimeError("SigPy is required for SigPyNufftBackend. "
                           "Install the toporecon package dependencies first.")
    if args.motion_groups != 1 or args.echo_groups != 1:
        raise ValueError("one process owns the whole motion x echo grid: use --motion-groups 1 "
                         "--echo-groups 1")
    data = load_conditioned(args, algorithm)
    from .algorithms import TvmeReconstructor, TvmReconstructor, TvmwReconstructor
    B = args.num_bins
    common = dict(sigma=1 / 6, tau=1 / 6, max_iter=args.max_iter, tol=args.tol, margin=0,
                  device=args.device, B_total=B, b0=0, b1=B, bin_edges=data["bin_edges"],
                  show_pbar=args.show_pbar, nufft_backend="hip")
    ksp, coord, dcf, mps, resp = data["ksp"], data["coord"], data["dcf"], data["mps"], data["resp"]
    t0 = time.perf_counter()
    if algorithm == "tvm":
        rec = TvmReconstructor(ksp, coord, dcf, mps, resp, None, B, args.l2_coupling, lamda=args.lamda,
                               reduce_per_bin=args.reduce_per_bin, **common)
    elif algorithm == "tvme":
        rec = TvmeReconstructor(ksp, coord, dcf, mps, resp, None, B, args.l2_coupling, lamda_m=args.lamda_m,
                                lamda_e=args.lamda_e, E_total=data["E_total"], e0=0, e1=data["E_total"],
                                **common)
    else:
        rec = TvmwReconstructor(ksp, coord, dcf, mps, resp, None, B, args.l2_coupling, lambda1=args.lambda1,
                                lambda2=args.lambda2, lambda3=args.lambda3, E_total=data["E_total"], e0=0,
                                e1=data["E_total"], **common)
    if getattr(args, "warm_start", "none") == "learned":
        from .learned import WarmStartNet, default_weights_path, warm_start
        path = args.warm_start_weights or default_weights_path()
        if not os.path.exists(path):
            raise FileNotFoundError(f"--warm-start learned: no weights at {path}")
        net = WarmStartNet.load(path, args.device)
        u0, q0 = warm_start(rec, net, cg_steps=_cg_steps(args), with_dual=True)
        img = rec.run(init=u0, dual_init=q0).cpu().numpy()
    else:
        img = rec.run().cpu().numpy()
    logging.info("Reconstruction done in %.2f seconds (%d iterations).", time.perf_counter() - t0,
                 rec.iterations)
    crop = data["orig_img_shape"] if args.crop_to_orig and not np.allclose(data["fov"], 1.0) else None
    write_output(img, args.output_dir, args.img_file, _manifest(args, algorithm, data["E_total"], data["prepared"]),
                 crop_zyx=crop)
    logging.info("Main done in %.2f seconds.", time.perf_counter() - t_main)
    return 0


def write_output(img, output_dir, stem, manifest, crop_zyx=None):
    """The reconstruction shard of one process that owns every bin and echo: ubar (B, E, Z, Y, X)
    as the 9-D (B, 1, 1, E, 1, 1, Z, Y, X) complex64 CFL ``{stem}_e0-{E-1}_m0-{B-1}`` plus
    ``run_manifest.json`` (tvm.py:1020-1083, tvme.py:984-1048, tvmw.py:1170-1234); the reference's
    ``toporecon stitch`` assembles it unchanged.  Returns the shard's base path."""
    img = np.asarray(img)
    Bn, E = img.shape[:2]
    Z, Y, X = img.shape[-3:]
    out = np.zeros((Bn, 1, 1, E, 1, 1, Z, Y, X), dtype=np.complex64)
    out[:, 0, 0, :, 0, 0] = img
    if crop_zyx is not None:
        out = center_crop_zyx(out, crop_zyx)
    os.makedirs(output_dir, exist_ok=True)
    base = os.path.join(output_dir, f"{stem}_e0-{E - 1}_m0-{Bn - 1}")
    cfl.write_cfl(base, out)
    with open(os.path.join(output_dir, "run_manifest.json"), "w", encoding="utf-8") as f:
        json.dump(manifest, f, indent=2, sort_keys=True)
        f.write("\n")
    return base


def tvm_main(argv=None) -> int:
    return run_main("tvm", argv)


def tvme_main(argv=None) -> int:
    return run_main("tvme", argv)


def tvmw_main(argv=None) -> int:
    return run_main("tvmw", argv)


def cli_main(argv=None) -> int:
    """``reconstruct --algorithm {tvm,tvme,tvmw} [--] ARGS``, ``resp INPUT --work-dir DIR [--device]`` and ``jsense
    INPUT --work-dir DIR [--device --readout-fraction --coordinate-scaling --fov-scale --show-progress]``
    as in src/toporecon/cli.py:23-57, :75-88, :135-196.  The shards are the reference's format, so its own ``toporecon
    stitch`` assembles them."""
    p = argparse.ArgumentParser(prog="toporecon_hip")
    p.add_argument("--verbose", action="store_true")
    sub = p.add_subparsers(dest="command", required=True)
    r = sub.add_parser("reconstruct")
    r.add_argument("--algorithm", choices=ALGORITHMS, required=True)
    r.add_argument("algorithm_arguments", nargs=argparse.REMAINDER)
    rp = sub.add_parser("resp", help="Estimate resp.hdr/.cfl in a separate work directory.")
    rp.add_argument("input", type=Path)
    rp.add_argument("--work-dir", type=Path, required=True)
    rp.add_argument("--device", type=int, default=None,
                    help="GPU index for the band-pass of the centre traces (default: SciPy on the host).")
    jp = sub.add_parser("jsense", help="Estimate mps.hdr/.cfl in a separate work directory.")
    jp.add_argument("input", type=Path)
    jp.add_argument("--work-dir", type=Path, required=True)
    jp.add_argument("--device", type=int, default=0, help="GPU device index (the HIP path has no CPU mode).")
    jp.add_argument("--readout-fraction", type=float, default=0.95,
                    help="Fraction of readout samples used by JSENSE.")
    jp.add_argument("--coordinate-scaling", choices=("legacy_max", "physical"), default="legacy_max",
                    help="Coordinate scaling for JSENSE (legacy_max: the historical per-axis maximum).")
    jp.add_argument("--fov-scale", type=float, nargs=3, default=(1.0, 1.0, 1.0), metavar=("Z", "Y", "X"),
                    help="Reconstruction FOV scale in z, y, x order.")
    jp.add_argument("--show-progress", action="store_true", help="Show a progress bar over the JSENSE rounds.")
    args = p.parse_args(argv)
    logging.basicConfig(level=logging.DEBUG if args.verbose else logging.INFO,
                        format="%(levelname)s: %(message)s")
    if args.command == "resp":
        from .preprocessing import estimate_respiratory_from_dataset
        args.work_dir.mkdir(parents=True, exist_ok=True)
        out = estimate_respiratory_from_dataset(args.input, args.work_dir, device=args.device)
        print(f"Wrote {out}.hdr and {out}.cfl")
        return 0
    if args.command == "jsense":
        from .jsense import JsenseConfig, estimate_sensitivity_from_dataset
        args.work_dir.mkdir(parents=True, exist_ok=True)
        cfg = JsenseConfig(readout_fraction=args.readout_fraction, coordinate_scaling=args.coordinate_scaling)
        out = estimate_sensitivity_from_dataset(args.input, args.work_dir, device=args.device,
                                                fov_scale_zyx=args.fov_scale, show_progress=args.show_progress,
                                                config=cfg)
        print(f"Wrote {out}.hdr and {out}.cfl")
        return 0
    fwd = list(args.algorithm_arguments)
    if fwd[:1] == ["--"]:
        fwd = fwd[1:]
    return run_main(args.algorithm, fwd)
