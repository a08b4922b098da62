"""BART-style CFL/HDR arrays in the reference's convention (src/toporecon/data/cfl.py).

The header lists dimensions in reverse NumPy order after a ``# Dimensions``
line; the payload is the complex64 array in C order.  Shard readers return the
canonical layouts the solvers consume: k-space ``[echo, coil, trajectory,
readout]`` from the 9-D ``(1,1,1,E,1,C,Ntr,Nro,1)`` file (cfl.py:101-187) and
maps ``[coil, z, y, x]`` (cfl.py:190-233).
"""

from __future__ import annotations

from pathlib import Path
from typing import Sequence

import numpy as np

DTYPE = np.dtype(np.complex64)


def base_path(name) -> Path:
    p = Path(name)
    return p.with_suffix("") if p.suffix in (".cfl", ".hdr") else p


def read_cfl_header(name) -> tuple:
    """Dimensions in NumPy order (the file stores them reversed)."""
    hdr = base_path(name).with_suffix(".hdr")
    dims = None
    for line in hdr.read_text(encoding="utf-8").splitlines():
        line = line.strip()
        if line and not line.startswith("#"):
            dims = [int(v) for v in line.split()]
            break
    if not dims:
        raise ValueError(f"No dimensions found in CFL header: {hdr}")
    if min(dims) < 1:
        raise ValueError(f"CFL dimensions must be positive: {dims}")
    return tuple(dims[::-1])


def expected_cfl_bytes(shape: Sequence[int]) -> int:
    return int(np.prod(tuple(shape), dtype=np.int64)) * DTYPE.itemsize


def validate_cfl_storage(name, shape=None) -> None:
    root = base_path(name)
    shape = tuple(shape) if shape is not None else read_cfl_header(root)
    want = expected_cfl_bytes(shape)
    got = root.with_suffix(".cfl").stat().st_size
    if got != want:
        raise ValueError(f"CFL payload size mismatch for {root}: expected {want} bytes, found {got} bytes.")


def read_cfl(name, *, mmap: bool = False) -> np.ndarray:
    root = base_path(name)
    shape = read_cfl_header(root)
    validate_cfl_storage(root, shape)
    if mmap:
        return np.memmap(root.with_suffix(".cfl"), mode="r", dtype=DTYPE, shape=shape)
    return np.fromfile(root.with_suffix(".cfl"), dtype=DTYPE,
                       count=int(np.prod(shape, dtype=np.int64))).reshape(shape)


def write_cfl(name, array) -> None:
    root = base_path(name)
    root.parent.mkdir(parents=True, exist_ok=True)
    a = np.ascontiguousarray(array, dtype=DTYPE)
    root.with_suffix(".hdr").write_text("# Dimensions\n" + " ".join(str(d) for d in a.shape[::-1]) + "\n",
                                        encoding="utf-8")
    a.tofile(root.with_suffix(".cfl"))


def kspace_layout(shape) -> tuple:
    """(echoes, coils, trajectories, readout) of a legacy >= 8-D k-space shape."""
    d = tuple(int(v) for v in shape)
    if len(d) < 8:
        raise ValueError(f"Unsupported k-space CFL shape: {d}.")
    out = (d[-6], d[-4], d[-3], d[-2])
    if int(np.prod(out, dtype=np.int64)) != int(np.prod(d, dtype=np.int64)):
        raise ValueError(f"Unsupported non-singleton dimensions in k-space CFL shape {d}.")
    return out


def read_kspace_shard(name, *, echo_start: int, echo_end: int, coil_indices, readout=None,
                      trajectory_indices=None) -> np.ndarray:
    """[echo, coil, trajectory, readout] block; requested index orders are preserved."""
    root = base_path(name)
    stored = read_cfl_header(root)
    validate_cfl_storage(root, stored)
    E, C, T, R = kspace_layout(stored)
    if not 0 <= echo_start < echo_end <= E:
        raise ValueError(f"Invalid echo range [{echo_start}, {echo_end}) for {E} echoes.")
    coils = np.asarray(coil_indices, dtype=np.int64).reshape(-1)
    if coils.size == 0:
        raise ValueError("At least one coil index is required.")
    if coils.min() < 0 or coils.max() >= C:
        raise IndexError(f"Coil indices [{coils.min()}, {coils.max()}] are outside [0, {C}).")
    ro = R if readout is None else int(readout)
    if not 1 <= ro <= R:
        raise ValueError(f"readout must be in [1, {R}], got {ro}.")
    mm = np.memmap(root.with_suffix(".cfl"), mode="r", dtype=DTYPE, shape=(E, C, T, R))
    block = np.take(mm[echo_start:echo_end], coils, axis=1)
    if trajectory_indices is not None:
        tr = np.asarray(trajectory_indices, dtype=np.int64).reshape(-1)
        if tr.size and (tr.min() < 0 or tr.max() >= T):
            raise IndexError(f"Trajectory indices are outside [0, {T}).")
        block = np.take(block, tr, axis=2)
    return np.array(block[..., :ro], dtype=DTYPE, order="C")


def read_mps_coils(name, *, coil_indices) -> np.ndarray:
    """Selected sensitivity maps as [coil, z, y, x]."""
    root = base_path(name)
    shape = read_cfl_header(root)
    validate_cfl_storage(root, shape)
    if len(shape) < 4:
        raise ValueError(f"Unsupported MPS CFL shape: {shape}.")
    C = int(shape[-4])
    sp = tuple(int(v) for v in shape[-3:])
    if C * int(np.prod(sp, dtype=np.int64)) != int(np.prod(shape, dtype=np.int64)):
        raise ValueError(f"Unsupported non-singleton dimensions in MPS CFL shape {shape}.")
    coils = np.asarray(coil_indices, dtype=np.int64).reshape(-1)
    if coils.size == 0:
        raise ValueError("At least one coil index is required.")
    if coils.min() < 0 or coils.max() >= C:
        raise IndexError(f"Coil indices [{coils.min()}, {coils.max()}] are outside [0, {C}).")
    mm = np.memmap(root.with_suffix(".cfl"), mode="r", dtype=DTYPE, shape=shape)
    sel = np.moveaxis(mm, -4, 0)[coils]
    return np.array(sel.reshape((coils.size,) + sp), dtype=DTYPE, order="C")


def read_zyx_txt(path) -> list:
    """'X:Y:Z' text (imageDim.txt / voxelSize.txt) -> [z, y, x]."""
    line = Path(path).read_text(encoding="utf-8").splitlines()[0]
    return [float(v) for v in line.split(":")][::-1]
