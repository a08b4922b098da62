"""Trajectory preparation for JSENSE: axis order, readout truncation, coordinate scaling.

Restates ``toporecon.data.trajectory.prepare_trajectory`` (src/toporecon/data/trajectory.py:32-107):
the raw ``[trajectory, readout, 3]`` coordinates are ``[x, y, z]`` and are reversed once; the
first ``max(1, int(fraction * readout))`` readout samples are kept and shrunk further while any
coordinate exceeds the instability threshold; ``physical`` scaling multiplies each axis by
voxel size x matrix size, ``legacy_max`` by ``matrix / (2 max)`` (the historical JSENSE script);
the FOV scale multiplies coordinates and the reconstruction shape.  Host code: one pass over the
trajectory, once per dataset.
"""

from __future__ import annotations

from dataclasses import dataclass
from typing import Literal, Sequence

import numpy as np

CoordinateScaling = Literal["physical", "legacy_max"]


@dataclass(frozen=True)
class PreparedTrajectory:
    """Trajectory and density prepared for a NUFFT (trajectory.py:14-23)."""

    coordinates_zyx: np.ndarray
    density: np.ndarray
    image_shape_zyx: tuple
    original_image_shape_zyx: tuple
    readout: int
    scaling: str


def _three(values, name: str) -> tuple:
    if len(values) != 3:
        raise ValueError(f"{name} must contain exactly three values.")
    return tuple(values)


def prepare_trajectory(trajectory_xyz, density, *, image_shape_zyx: Sequence[int],
                       voxel_size_zyx: Sequence[float], readout_fraction: float,
                       fov_scale_zyx: Sequence[float] = (1.0, 1.0, 1.0), scaling: str = "physical",
                       unstable_coordinate_threshold: float | None = 1000.0) -> PreparedTrajectory:
    """(trajectory, readout, 3) [x, y, z] float32 coordinates -> [z, y, x] NUFFT coordinates."""
    coords = np.asarray(trajectory_xyz, dtype=np.float32)
    dens = np.asarray(density, dtype=np.float32)
    if coords.ndim != 3 or coords.shape[-1] != 3:
        raise ValueError(f"trajectory must have shape [trajectory, readout, 3], got {coords.shape}.")
    if dens.shape != coords.shape[:2]:
        raise ValueError(f"density shape {dens.shape} does not match trajectory {coords.shape[:2]}.")
    if not 0 < readout_fraction <= 1:
        raise ValueError("readout_fraction must be in (0, 1].")
    shape = tuple(int(v) for v in _three(image_shape_zyx, "image_shape"))
    voxel = tuple(float(v) for v in _three(voxel_size_zyx, "voxel_size"))
    fov = tuple(float(v) for v in _three(fov_scale_zyx, "fov_scale"))
    if any(v < 1 for v in fov):
        raise ValueError("fov_scale values must be >= 1.")

    readout = max(1, int(readout_fraction * coords.shape[-2]))
    coords = coords[:, :readout, ::-1].copy()
    dens = dens[:, :readout].copy()
    if unstable_coordinate_threshold is not None:
        while readout > 1 and float(coords.max()) > unstable_coordinate_threshold:
            readout -= 1
            coords = coords[:, :readout]
            dens = dens[:, :readout]

    if scaling == "physical":
        for a in range(3):
            coords[..., a] *= voxel[a] * shape[a]
    elif scaling == "legacy_max":
        for a in range(3):
            top = float(np.max(coords[..., a]))
            if top == 0:
                raise ValueError(f"trajectory axis {a} has zero maximum.")
            coords[..., a] *= shape[a] / (2.0 * top)
    else:
        raise ValueError(f"Unknown coordinate scaling mode: {scaling!r}")

    recon_shape = tuple(int(shape[a] * fov[a]) for a in range(3))
    for a in range(3):
        coords[..., a] *= fov[a]
    return PreparedTrajectory(np.ascontiguousarray(coords), np.ascontiguousarray(dens), recon_shape, shape,
                              readout, scaling)
