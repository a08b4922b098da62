"""Seeded synthetic multi-echo, motion-resolved 3-D cone acquisitions.

Generates the inputs the reconstruction path consumes, in the reference's
conventions (trajectory in [z, y, x] SigPy units, k in [-n/2, n/2] per axis;
k-space (E, C, Ntr, Nro); maps (C, Z, Y, X); one respiratory value per
interleave):

* cone-like trajectory: interleave axes from a Fibonacci sphere, radius rising
  linearly from 0 to n/2 per axis, a fixed cone opening and a 4-turn twist;
* density clip(|k|^2, 1e-3) normalised to mean 1;
* phantom of seeded ellipsoids with magnitude U(0.2, 1), R2* U(20, 80)/s and
  off-resonance U(-50, 50) Hz, signal m exp(-TE R2*) exp(i 2 pi df TE), shifted
  along z by 2 sin(2 pi b / B) voxels in motion bin b;
* smooth complex Gaussian-blob coil sensitivities normalised to RSS = 1;
* respiratory trace: 4 s period sinusoid sampled at TR = 5 ms.

K-space simulation (NUFFT of coil-weighted images plus complex Gaussian noise
at a given SNR) is left to the caller's forward operator, so the same
generator feeds the HIP path and CPU checks.
"""

from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import List, Tuple

import numpy as np

from . import conditioning as cond


@dataclass
class AcquisitionConfig:
    img_shape: Tuple[int, int, int]
    coils: int
    echo_times_ms: List[float]
    bins: int
    interleaves: int
    readouts: int
    algorithm: str = "tvm"
    lambda_motion: float = 1e-5
    lambda_echo: float = 1e-5
    l2_coupling: bool = False
    readout_fraction: float = 0.98
    tol: float = 1e-3
    max_iter: int = 100
    snr_db: float = 30.0
    seed: int = 0
    tr_s: float = 0.005
    resp_period_s: float = 4.0
    cone_angle: float = 0.3
    turns: float = 4.0
    ellipsoids: int = 6
    extra: dict = field(default_factory=dict)

    @property
    def echoes(self) -> int:
        return len(self.echo_times_ms)


def config_small_tvm() -> AcquisitionConfig:
    """TVM at 48x64x64, 8 coils, 2 echoes (1.2, 2.6 ms), 4 bins, 1536x256 samples."""
    return AcquisitionConfig((48, 64, 64), 8, [1.2, 2.6], 4, 1536, 256, algorithm="tvm",
                             max_iter=100, seed=0)


def config_full_tvme() -> AcquisitionConfig:
    """TVME at 80x192x192, 32 coils, 6 echoes (1.0 + 1.3 e ms), 6 bins, 9000x512 samples."""
    return AcquisitionConfig((80, 192, 192), 32, [1.0 + 1.3 * e for e in range(6)], 6, 9000, 512,
                             algorithm="tvme", l2_coupling=True, max_iter=300, seed=1)


def config_scaled(base: AcquisitionConfig, img_shape, coils=None, echoes=None, bins=None,
                  interleaves=None, readouts=None, seed=None) -> AcquisitionConfig:
    """Same generators at another size (used for tests and sub-problems)."""
    tes = base.echo_times_ms if echoes is None else [1.0 + 1.3 * e for e in range(echoes)]
    return AcquisitionConfig(tuple(img_shape), base.coils if coils is None else coils, tes,
                             base.bins if bins is None else bins,
                             base.interleaves if interleaves is None else interleaves,
                             base.readouts if readouts is None else readouts,
                             algorithm=base.algorithm, lambda_motion=base.lambda_motion,
                             lambda_echo=base.lambda_echo, l2_coupling=base.l2_coupling,
                             readout_fraction=base.readout_fraction, tol=base.tol,
                             max_iter=base.max_iter, snr_db=base.snr_db,
                             seed=base.seed if seed is None else seed)


def batched_instance_configs(seeds=range(100, 112)):
    """Independent TVM problems of uneven size from the small-TVM generator, sizes drawn per
    seed: matrices from 32^3 (first seed) to 96x128x128 (last seed), 8-32 coils, 1-6 echoes,
    4-8 bins, 0.2M-3M samples per echo (about 1.5-4 per voxel), tol 1e-3, at most 100
    iterations."""
    seeds = list(seeds)
    out = []
    for k, seed in enumerate(seeds):
        rng = np.random.default_rng(seed)
        if k == 0:
            shape = (32, 32, 32)
        elif k == len(seeds) - 1:
            shape = (96, 128, 128)
        else:
            yx = int(rng.choice([32, 48, 64, 96, 128]))
            shape = (int(rng.choice([32, 48, 64, 80, 96])), yx, yx)
        coils = int(rng.choice([8, 12, 16, 24, 32]))
        echoes = int(rng.integers(1, 7))
        bins = int(rng.integers(4, 9))
        nvox = shape[0] * shape[1] * shape[2]
        samples = float(np.clip(nvox * rng.uniform(1.5, 4.0), 0.2e6, 3.0e6))
        readouts = int(max(64, round(3 * max(shape) / 32) * 32))
        interleaves = max(20 * bins, int(round(samples / readouts)))
        cfg = config_scaled(config_small_tvm(), shape, coils=coils, echoes=echoes, bins=bins,
                            interleaves=interleaves, readouts=readouts, seed=seed)
        cfg.max_iter = 100
        out.append(cfg)
    return out


def fibonacci_axes(n: int) -> np.ndarray:
    """n unit vectors (z, y, x) on a Fibonacci sphere."""
    i = np.arange(n, dtype=np.float64)
    z = 1.0 - 2.0 * (i + 0.5) / n
    r = np.sqrt(np.maximum(1.0 - z * z, 0.0))
    phi = i * math.pi * (3.0 - math.sqrt(5.0))
    return np.stack([z, r * np.cos(phi), r * np.sin(phi)], axis=-1)


def cone_trajectory(n_il: int, n_ro: int, img_shape, turns: float = 4.0,
                    cone_angle: float = 0.3) -> np.ndarray:
    """(n_il, n_ro, 3) float32 trajectory in [z, y, x] SigPy units."""
    a = fibonacci_axes(n_il)
    ref = np.where(np.abs(a[:, :1]) < 0.9, np.array([[1.0, 0.0, 0.0]]), np.array([[0.0, 1.0, 0.0]]))
    u = np.cross(a, ref)
    u /= np.linalg.norm(u, axis=-1, keepdims=True)
    v = np.cross(a, u)
    t = np.linspace(0.0, 1.0, n_ro)
    phi = 2.0 * math.pi * turns * t
    ca, sa = math.cos(cone_angle), math.sin(cone_angle)
    d = (ca * a[:, None, :]
         + sa * (np.cos(phi)[None, :, None] * u[:, None, :]
                 + np.sin(phi)[None, :, None] * v[:, None, :]))
    half = np.array([n / 2.0 for n in img_shape])
    k = t[None, :, None] * d * half[None, None, :]
    return k.astype(np.float32)


def density(coord: np.ndarray) -> np.ndarray:
    d = np.maximum(np.sum(coord.astype(np.float64) ** 2, axis=-1), 1e-3)
    return (d / d.mean()).astype(np.float32)


def respiratory_trace(n_il: int, tr_s: float = 0.005, period_s: float = 4.0) -> np.ndarray:
    t = np.arange(n_il) * tr_s
    return np.sin(2.0 * math.pi * t / period_s).astype(np.float32)


def _grid(img_shape):
    axes = [(np.arange(n, dtype=np.float32) - n / 2.0) / n for n in img_shape]
    return np.meshgrid(*axes, indexing="ij")


def phantom(img_shape, echo_times_ms, num_bins, seed, n_ellipsoids=6) -> np.ndarray:
    """(B, E, Z, Y, X) complex64 seeded ellipsoid phantom."""
    rng = np.random.RandomState(seed)
    Z, Y, X = img_shape
    zz, yy, xx = _grid(img_shape)
    params = []
    for _ in range(n_ellipsoids):
        params.append(dict(c=rng.uniform(-0.22, 0.22, 3), r=rng.uniform(0.08, 0.3, 3),
                           m=rng.uniform(0.2, 1.0), r2=rng.uniform(20.0, 80.0),
                           df=rng.uniform(-50.0, 50.0)))
    tes = np.asarray(echo_times_ms, dtype=np.float64) * 1e-3
    out = np.zeros((num_bins, len(tes), Z, Y, X), dtype=np.complex64)
    for b in range(num_bins):
        dz = 2.0 * math.sin(2.0 * math.pi * b / num_bins) / Z
        for p in params:
            inside = (((zz - p["c"][0] - dz) / p["r"][0]) ** 2 + ((yy - p["c"][1]) / p["r"][1]) ** 2
                      + ((xx - p["c"][2]) / p["r"][2]) ** 2) <= 1.0
            for e, te in enumerate(tes):
                val = p["m"] * math.exp(-te * p["r2"]) * complex(math.cos(2 * math.pi * p["df"] * te),
                                                                  math.sin(2 * math.pi * p["df"] * te))
                out[b, e][inside] += np.complex64(val)
    return out


def coil_maps(coils, img_shape, seed) -> np.ndarray:
    """(C, Z, Y, X) complex64 smooth Gaussian-blob sensitivities, RSS-normalised."""
    rng = np.random.RandomState(seed + 7919)
    zz, yy, xx = _grid(img_shape)
    maps = np.empty((coils,) + tuple(img_shape), dtype=np.complex64)
    for c in range(coils):
        th = 2.0 * math.pi * c / coils
        cz = 0.25 * (1 if c % 2 == 0 else -1)
        cy, cx = 0.6 * math.cos(th), 0.6 * math.sin(th)
        w = 0.45
        mag = np.exp(-((zz - cz) ** 2 + (yy - cy) ** 2 + (xx - cx) ** 2) / (2 * w * w))
        ph = rng.uniform(-math.pi, math.pi) + 2.0 * (rng.uniform(-1, 1) * yy + rng.uniform(-1, 1) * xx)
        maps[c] = (mag * np.exp(1j * ph)).astype(np.complex64)
    rss = np.sqrt(np.sum(np.abs(maps) ** 2, axis=0))
    return (maps / np.maximum(rss, 1e-12)[None]).astype(np.complex64)


@dataclass
class Acquisition:
    """Conditioned trajectory side of a synthetic acquisition (no k-space yet)."""

    cfg: AcquisitionConfig
    coord: np.ndarray        # (Ntr_kept, Nro_kept, 3) float32
    dcf: np.ndarray          # (Ntr_kept, Nro_kept) float32, normalised by max
    resp: np.ndarray         # (Ntr_kept,)
    bin_edges: np.ndarray    # (B + 1,)
    sim_bin: np.ndarray      # (Ntr_kept,) motion state used to simulate each interleave
    mps: np.ndarray          # (C, Z, Y, X) complex64, RSS-normalised
    truth: np.ndarray        # (B, E, Z, Y, X) complex64

    def bin_index(self, b: int) -> np.ndarray:
        return np.nonzero(cond.bin_members(self.resp, self.bin_edges, b))[0]


def make_acquisition(cfg: AcquisitionConfig) -> Acquisition:
    """Trajectory, density, respiratory binning, maps and phantom, conditioned like TVME main."""
    coord_full = cone_trajectory(cfg.interleaves, cfg.readouts, cfg.img_shape, cfg.turns,
                                 cfg.cone_angle)
    dens_full = density(coord_full)
    resp_full = respiratory_trace(cfg.interleaves, cfg.tr_s, cfg.resp_period_s)
    keep = cond.drop_respiratory_outliers(resp_full)
    nro = cond.readout_count(cfg.readout_fraction, cfg.readouts)
    coord = np.ascontiguousarray(coord_full[keep, :nro])
    dcf = cond.normalize_dcf(np.ascontiguousarray(dens_full[keep, :nro]))
    resp = resp_full[keep]
    edges = cond.bin_edges(resp, cfg.bins)
    sim_bin = np.clip(np.searchsorted(edges, resp, side="right") - 1, 0, cfg.bins - 1)
    mps = coil_maps(cfg.coils, cfg.img_shape, cfg.seed)
    truth = phantom(cfg.img_shape, cfg.echo_times_ms, cfg.bins, cfg.seed, cfg.ellipsoids)
    return Acquisition(cfg, coord, dcf, resp, edges, sim_bin, mps, truth)


def noise_sigma(signal_rms: float, snr_db: float) -> float:
    return float(signal_rms) * 10.0 ** (-snr_db / 20.0)


def simulate_kspace_numpy(acq: Acquisition, forward) -> np.ndarray:
    """(E, C, Ntr, Nro) complex64 k-space, normalised by max|ksp| (tvme.py:938-940).

    ``forward(images (E, C, Z, Y, X), coord (M, 3)) -> (E, C, M)`` is the NUFFT used.
    """
    cfg = acq.cfg
    E, C = cfg.echoes, cfg.coils
    ntr, nro = acq.coord.shape[:2]
    ksp = np.zeros((E, C, ntr, nro), dtype=np.complex64)
    for b in range(cfg.bins):
        idx = np.nonzero(acq.sim_bin == b)[0]
        if idx.size == 0:
            continue
        img = acq.truth[b][:, None] * acq.mps[None]
        y = forward(img, acq.coord[idx].reshape(-1, 3))
        ksp[:, :, idx] = np.asarray(y).reshape(E, C, idx.size, nro)
    rms = float(np.sqrt(np.mean(np.abs(ksp) ** 2)))
    sig = noise_sigma(rms, cfg.snr_db) / math.sqrt(2.0)
    rng = np.random.RandomState(cfg.seed + 104729)
    ksp += (rng.standard_normal(ksp.shape) * sig + 1j * rng.standard_normal(ksp.shape) * sig).astype(np.complex64)
    return cond.normalize_kspace(ksp)


def write_raw_dataset(acq: "Acquisition", ksp_full, coord_full, dens_full, resp_full, root,
                      voxel_size=(1.0, 1.0, 1.0)):
    """Write a raw MICCAI-style dataset (+ prepared mps/resp) for the reference CLI layout.

    ksp_full (E, C, Ntr, Nro) complex; coord_full (Ntr, Nro, 3) in [z, y, x] SigPy units;
    written as ktraj in [x, y, z] order divided by voxel * n (the main multiplies back,
    tvme.py:778-783); imageDim.txt / voxelSize.txt as "X:Y:Z".
    """
    import os

    from .data import cfl
    os.makedirs(root, exist_ok=True)
    Z, Y, X = acq.cfg.img_shape
    E, C, T, R = ksp_full.shape
    cfl.write_cfl(os.path.join(root, "ksp"), np.asarray(ksp_full).reshape(1, 1, 1, E, 1, C, T, R, 1))
    vz = np.asarray(voxel_size, dtype=np.float64)
    scale = vz * np.array([Z, Y, X], dtype=np.float64)
    raw = (np.asarray(coord_full, dtype=np.float64) / scale)[..., ::-1].astype(np.float32)
    cfl.write_cfl(os.path.join(root, "ktraj"), raw.reshape((1, 1, 1, 1, 1, 1, T, R, 3)))
    cfl.write_cfl(os.path.join(root, "dens"), np.asarray(dens_full).reshape(1, 1, 1, 1, 1, 1, T, R, 1))
    with open(os.path.join(root, "imageDim.txt"), "w") as f:
        f.write(f"{X}:{Y}:{Z}\n")
    with open(os.path.join(root, "voxelSize.txt"), "w") as f:
        f.write(":".join(str(v) for v in vz[::-1]) + "\n")
    with open(os.path.join(root, "tr.txt"), "w") as f:
        f.write(f"{acq.cfg.tr_s * 1000.0}\n")
    cfl.write_cfl(os.path.join(root, "mps"), acq.mps.reshape((1, 1, 1, 1, 1, C, Z, Y, X)))
    cfl.write_cfl(os.path.join(root, "resp"), np.asarray(resp_full).reshape(-1, 1))


def make_raw(cfg: AcquisitionConfig, forward, noise_seed=None):
    """Unconditioned raw arrays of a synthetic acquisition, as stored in a raw dataset.

    Returns (acq, ksp_full (E, C, Ntr, Nro), coord_full, dens_full, resp_full); every
    interleave is simulated with the motion state of its respiratory value.
    ``forward(images (E, C, Z, Y, X), coord (M, 3)) -> (E, C, M)``.
    """
    acq = make_acquisition(cfg)
    coord_full = cone_trajectory(cfg.interleaves, cfg.readouts, cfg.img_shape, cfg.turns, cfg.cone_angle)
    dens_full = density(coord_full)
    resp_full = respiratory_trace(cfg.interleaves, cfg.tr_s, cfg.resp_period_s)
    sim = np.clip(np.searchsorted(acq.bin_edges, resp_full, side="right") - 1, 0, cfg.bins - 1)
    E, C = cfg.echoes, cfg.coils
    ksp = np.zeros((E, C, cfg.interleaves, cfg.readouts), dtype=np.complex64)
    for b in range(cfg.bins):
        idx = np.nonzero(sim == b)[0]
        if idx.size:
            y = forward(acq.truth[b][:, None] * acq.mps[None], coord_full[idx].reshape(-1, 3))
            ksp[:, :, idx] = np.asarray(y).reshape(E, C, idx.size, cfg.readouts)
    rms = float(np.sqrt(np.mean(np.abs(ksp) ** 2)))
    sig = noise_sigma(rms, cfg.snr_db) / math.sqrt(2.0)
    rng = np.random.RandomState(cfg.seed + 104729 if noise_seed is None else noise_seed)
    ksp += (rng.standard_normal(ksp.shape) * sig + 1j * rng.standard_normal(ksp.shape) * sig).astype(np.complex64)
    return acq, ksp, coord_full, dens_full, resp_full
