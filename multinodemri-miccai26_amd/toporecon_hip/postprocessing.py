"""Assembly of the per-(motion, echo) reconstruction shards into one CFL image.

Mirrors ``toporecon.postprocessing.stitch`` (src/toporecon/postprocessing/stitch.py:192-256):
a run writes shards ``{stem}_e{e0}-{e1}_m{b0}-{b1}`` (inclusive ranges, 9-D
``(B_g,1,1,E_g,1,1,Z,Y,X)`` complex64) plus ``run_manifest.json``; ``stitch``
checks that the shards tile the motion x echo grid the manifest records and
writes the ``(B,1,1,E,1,1,Z,Y,X)`` image.  Same arguments, defaults and error
types / messages as the reference; the output is written under a temporary
name and renamed into place, so a failed run leaves no partial image.
"""

from __future__ import annotations

import json
import os
import re
import uuid
from dataclasses import dataclass
from pathlib import Path

import numpy as np

from .data import cfl

_SINGLETON_AXES = (1, 2, 4, 5)
_MOTION_AXIS, _ECHO_AXIS = 0, 3


@dataclass(frozen=True)
class Shard:
    base: Path
    echoes: tuple   # [start, stop)
    motion: tuple   # [start, stop)
    shape: tuple


def _manifest(directory: Path) -> dict:
    path = directory / "run_manifest.json"
    if not path.is_file():
        raise FileNotFoundError(f"Missing reconstruction manifest: {path}")
    try:
        data = json.loads(path.read_text(encoding="utf-8"))
    except json.JSONDecodeError as error:
        raise ValueError(f"Invalid reconstruction manifest: {path}") from error
    if not isinstance(data, dict):
        raise ValueError(f"Invalid reconstruction manifest: {path}")
    return data


def _shard_dir(directory: Path, stem: str):
    """(directory holding the shards, file-name prefix); the stem may name a subdirectory
    of ``directory`` but not leave it (stitch.py:41-53)."""
    rel = Path(stem)
    if rel.is_absolute() or not rel.name:
        raise ValueError("The input prefix must be relative to INPUT_DIR.")
    where = (directory / rel.parent).resolve()
    if where != directory and directory not in where.parents:
        raise ValueError("The input prefix must stay inside INPUT_DIR.")
    if not where.is_dir():
        raise FileNotFoundError(f"Shard directory does not exist: {where}")
    return where, rel.name


def find_shards(directory: Path, stem: str) -> list:
    """Every ``{stem}_e{e0}-{e1}_m{b0}-{b1}.hdr`` with its payload checked (stitch.py:56-86)."""
    where, prefix = _shard_dir(directory, stem)
    name_re = re.compile(re.escape(prefix) + r"_e(\d+)-(\d+)_m(\d+)-(\d+)\.hdr")
    found = []
    for hdr in sorted(where.iterdir()):
        m = name_re.fullmatch(hdr.name)
        if m is None:
            continue
        base = hdr.with_suffix("")
        if not base.with_suffix(".cfl").is_file():
            raise FileNotFoundError(f"Missing CFL payload for shard: {base}")
        shape = cfl.read_cfl_header(base)
        cfl.validate_cfl_storage(base, shape)
        e0, e1, b0, b1 = (int(g) for g in m.groups())
        found.append(Shard(base=base, echoes=(e0, e1 + 1), motion=(b0, b1 + 1), shape=tuple(shape)))
    if not found:
        raise FileNotFoundError(f"No shards found for input prefix {stem!r}.")
    return found


def _covers(ranges, total: int, name: str) -> None:
    """The distinct [start, stop) ranges partition [0, total) (stitch.py:89-97)."""
    end = 0
    for start, stop in sorted(ranges):
        if start != end or stop <= start:
            raise ValueError(f"The {name} shard ranges have a gap or overlap.")
        end = stop
    if end != total:
        raise ValueError(f"The {name} shard ranges do not cover 0:{total}.")


def _grid_config(manifest: dict):
    try:
        params, dist = manifest["parameters"], manifest["distributed"]
        nbins, mg, eg = int(params["num_bins"]), int(dist["motion_groups"]), int(dist["echo_groups"])
    except (KeyError, TypeError, ValueError) as error:
        raise ValueError("run_manifest.json is missing its grid configuration.") from error
    if min(nbins, mg, eg) < 1:
        raise ValueError("The grid configuration must contain positive integers.")
    return params, nbins, mg, eg


def output_shape(shards: list, manifest: dict) -> tuple:
    """Validate the shard grid against the manifest; the assembled 9-D shape (stitch.py:100-168)."""
    params, nbins, mg, eg = _grid_config(manifest)
    if len(shards) != mg * eg:
        raise ValueError(f"Expected {mg * eg} shards, found {len(shards)}.")
    recorded = params.get("num_echoes")
    try:
        nechoes = max(s.echoes[1] for s in shards) if recorded is None else int(recorded)
    except (TypeError, ValueError) as error:
        raise ValueError("run_manifest.json contains an invalid echo count.") from error
    if nechoes < 1:
        raise ValueError("run_manifest.json contains an invalid echo count.")
    motion = {s.motion for s in shards}
    echoes = {s.echoes for s in shards}
    if len(motion) != mg or len(echoes) != eg:
        raise ValueError("Shard ranges do not match the manifest grid configuration.")
    if {(s.motion, s.echoes) for s in shards} != {(m, e) for m in motion for e in echoes}:
        raise ValueError("The motion/echo grid contains a missing or duplicate tile.")
    _covers(motion, nbins, "motion")
    _covers(echoes, nechoes, "echo")
    first = list(shards[0].shape)
    if len(first) != 9:
        raise ValueError(f"Expected a 9D reconstruction shard, got {first}.")
    if any(first[a] != 1 for a in _SINGLETON_AXES):
        raise ValueError("Reconstruction shard singleton axes are invalid.")
    for s in shards:
        if len(s.shape) != 9:
            raise ValueError(f"Unexpected shard shape: {s.base}")
        if s.shape[_MOTION_AXIS] != s.motion[1] - s.motion[0]:
            raise ValueError(f"Motion range does not match {s.base.name}.")
        if s.shape[_ECHO_AXIS] != s.echoes[1] - s.echoes[0]:
            raise ValueError(f"Echo range does not match {s.base.name}.")
        if any(s.shape[a] != first[a] for a in range(9) if a not in (_MOTION_AXIS, _ECHO_AXIS)):
            raise ValueError(f"Incompatible shard shape: {s.base.name}.")
    first[_MOTION_AXIS], first[_ECHO_AXIS] = nbins, nechoes
    return tuple(first)


def _output_base(directory: Path, output_name: str, output) -> Path:
    """``output`` (a path) wins; else ``output_name``, a bare file name inside the run
    directory (stitch.py:171-189)."""
    if output is None:
        name = Path(output_name)
        if name.is_absolute() or len(name.parts) != 1 or name.name != output_name or output_name in (".", ".."):
            raise ValueError("--output-name must be a filename, not a path.")
        target = directory / name
    else:
        target = Path(output).expanduser()
    if target.suffix not in ("", ".hdr", ".cfl"):
        raise ValueError("The output must be a base path or end in .hdr/.cfl.")
    return cfl.base_path(target).resolve()


def stitch(directory, *, stem=None, output_name: str = "imout", output=None, overwrite: bool = False) -> Path:
    """Assemble all shards of one reconstruction run; returns the output's base path
    (``.hdr`` / ``.cfl`` beside it).  stitch.py:192-256."""
    directory = Path(directory).expanduser().resolve()
    if not directory.is_dir():
        raise FileNotFoundError(f"Reconstruction directory does not exist: {directory}")
    manifest = _manifest(directory)
    stem = stem or manifest.get("output_stem")
    if not isinstance(stem, str) or not stem:
        raise ValueError("run_manifest.json is missing output_stem.")
    shards = find_shards(directory, stem)
    shape = output_shape(shards, manifest)
    base = _output_base(directory, output_name, output)
    hdr, payload = base.with_suffix(".hdr"), base.with_suffix(".cfl")
    inputs = {p.resolve() for s in shards for p in (s.base.with_suffix(".hdr"), s.base.with_suffix(".cfl"))}
    if hdr in inputs or payload in inputs:
        raise ValueError("The assembled output cannot overwrite an input shard.")
    if not overwrite and (hdr.exists() or payload.exists()):
        raise FileExistsError(f"Output already exists: {base}")
    base.parent.mkdir(parents=True, exist_ok=True)
    tmp = base.parent / f".stitch_{uuid.uuid4().hex}"
    try:
        # the payload is streamed shard by shard through memory maps, so no image-sized host buffer
        dst = np.memmap(tmp.with_suffix(".cfl"), mode="w+", dtype=cfl.DTYPE, shape=shape)
        for s in shards:
            src = cfl.read_cfl(s.base, mmap=True)
            dst[s.motion[0]:s.motion[1], :, :, s.echoes[0]:s.echoes[1]] = src
            del src
        dst.flush()
        del dst
        tmp.with_suffix(".hdr").write_text("# Dimensions\n" + " ".join(str(d) for d in shape[::-1]) + "\n",
                                           encoding="utf-8")
        cfl.validate_cfl_storage(tmp, shape)
        os.replace(tmp.with_suffix(".cfl"), payload)
        os.replace(tmp.with_suffix(".hdr"), hdr)
    finally:
        for p in (tmp.with_suffix(".cfl"), tmp.with_suffix(".hdr")):
            if p.exists():
                p.unlink()
    return base
