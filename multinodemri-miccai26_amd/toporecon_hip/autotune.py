"""Launch-parameter autotuner for the gridding kernels of a NUFFT plan.

The interpolation (forward gather at the samples) and the deterministic
spreading (adjoint gather over grid clusters) read the oversampled grid or the
sample rows through each XCD's own L2.  Which block order keeps the most of
that traffic inside one L2 depends on the grid size, the channel count and the
sample density, so it is measured, not guessed: for every candidate XCD chunk
the tuner times the two kernels on the plan's own bins with HIP events (the
library's per-kernel profiling counters) and keeps the fastest.  Choices are
cached per problem signature in memory and, optionally, in a JSON file.

Only the order in which workgroups visit samples or clusters changes; every
result bit is the same for every choice.
"""

from __future__ import annotations

import json
import os

from . import _lib

INTERP_CANDIDATES = (0, 16, 64, 256, 1024)
# sample geometry of a plan (fixed when it is built): sort tile edges, then heavy-cluster segment
# lengths, tried one axis at a time from the defaults
TILE_CANDIDATES = (4, 8, 16)
SEGMENT_CANDIDATES = (256, 1024, 4096)
# all channels per block / 64-channel groups / 16-byte loads over <=128-channel groups / the same with
# one gather per group of samples with identical taps
INTERP_KGROUPS = (0, 1, 2, 3)
SPREAD_CANDIDATES = (0, 64)

_cache: dict = {}


def signature(plan, K: int) -> str:
    m = [plan.samples(b) for b in range(plan.num_bins)]
    return f"n={plan.img_shape} g={plan.grid_shape} K={K} M={max(m) if m else 0}"


def _load(path):
    if path and os.path.exists(path):
        with open(path) as f:
            return json.load(f)
    return {}


def _time_kernel(name, fn, reps):
    _lib.profile_enable(True)
    for _ in range(reps):
        fn()
    prof = _lib.profile_read()
    _lib.profile_enable(False)
    ms, n, _ = prof.get(name, (0.0, 0, 0.0))
    return ms / n if n else float("inf")


def tune(plan, *, image, maps, E: int, C: int, ksp, dcf, stream, bins=(0,), reps: int = 2,
         cache_path: str | None = None, force: bool = False) -> dict:
    """Pick the interpolation and spreading XCD chunks for ``plan`` and set them.

    ``image`` (E, Z, Y, X) complex64, ``maps`` (C, Z, Y, X) complex64, ``ksp`` a
    (M_b, E*C) complex64 scratch for the largest timed bin and ``dcf`` (M_b,)
    float32, all on the plan's device; ``image`` is overwritten by the adjoint
    timing.  Returns {"interp_xcd_chunk", "spread_xcd_chunk", "timings_ms"}.
    """
    K = E * C
    key = signature(plan, K)
    disk = _load(cache_path)
    if not force:
        hit = _cache.get(key) or disk.get(key)
        if hit:
            plan.set_param("interp_xcd_chunk", hit["interp_xcd_chunk"])
            plan.set_param("interp_kgroups", hit.get("interp_kgroups", 0))
            plan.set_param("spread_xcd_chunk", hit["spread_xcd_chunk"])
            return dict(hit, cached=True)
    lib_call = _lib.call
    timings = {"interp": {}, "spread": {}}

    def fwd():
        for b in bins:
            lib_call("tr_data_simulate", plan.handle, b, image.data_ptr(), maps.data_ptr(), E, C,
                     ksp.data_ptr(), stream)

    def adj():
        for b in bins:
            lib_call("tr_data_adjoint", plan.handle, b, ksp.data_ptr(), dcf.data_ptr(), maps.data_ptr(),
                     E, C, image.data_ptr(), 0, stream)

    fwd()  # warm the workspace at its full size
    for kg in INTERP_KGROUPS:
        plan.set_param("interp_kgroups", kg)
        for c in INTERP_CANDIDATES:
            plan.set_param("interp_xcd_chunk", c)
            timings["interp"][(kg, c)] = _time_kernel("interp", fwd, reps)
    best_kg, best_i = min(timings["interp"], key=timings["interp"].get)
    plan.set_param("interp_kgroups", best_kg)
    plan.set_param("interp_xcd_chunk", best_i)
    adj()  # warm the adjoint path, so the first candidate is not charged for it
    for c in SPREAD_CANDIDATES:
        plan.set_param("spread_xcd_chunk", c)
        timings["spread"][c] = _time_kernel("spread", adj, reps)
    best_s = min(timings["spread"], key=timings["spread"].get)
    plan.set_param("spread_xcd_chunk", best_s)
    out = {"interp_xcd_chunk": best_i, "interp_kgroups": best_kg, "spread_xcd_chunk": best_s,
           "timings_ms": {k: {str(c): round(t, 4) for c, t in v.items()} for k, v in timings.items()}}
    _cache[key] = out
    if cache_path:
        disk[key] = out
        with open(cache_path, "w") as f:
            json.dump(disk, f, indent=1, sort_keys=True)
    return dict(out, cached=False)


def tune_geometry(img_shape, coords, *, image, maps, E: int, C: int, dcf, device, stream, reps: int = 2,
                  launch: dict | None = None) -> dict:
    """Pick the sample geometry (sort tile edge, then heavy-cluster segment length) for a plan of
    ``coords`` (one bin's (M, 3) float32 device coordinates, the largest bin of the problem).

    Every candidate is a plan of its own built from ``coords`` with the launch parameters in
    ``launch`` (the tuned ones of the problem's plan); the interpolation and the spreading of the
    data term are timed on it with HIP events and the geometry with the smallest sum is kept.
    ``image`` (E, Z, Y, X) and ``dcf`` (M,) are scratch of the right sizes.  Unlike the launch
    parameters, the geometry changes the order in which float32 products are added, so the
    results of two geometries agree to rounding, not bit for bit.
    Returns {"sort_tile", "spread_segment", "timings_ms"}.
    """
    import torch
    from .nufft_backend import NufftPlan
    K = E * C
    M = int(coords.shape[0])
    ksp = torch.zeros((M, K), dtype=torch.complex64, device=image.device)

    def cost(tile, seg):
        plan = NufftPlan(img_shape, [coords], device, sort_tile=tile, spread_segment=seg)
        try:
            for k, v in (launch or {}).items():
                if k in NufftPlan.LAUNCH_PARAMS:
                    plan.set_param(k, v)

            def fwd():
                _lib.call("tr_data_simulate", plan.handle, 0, image.data_ptr(), maps.data_ptr(), E, C,
                          ksp.data_ptr(), stream)

            def adj():
                _lib.call("tr_data_adjoint", plan.handle, 0, ksp.data_ptr(), dcf.data_ptr(), maps.data_ptr(),
                          E, C, image.data_ptr(), 0, stream)

            fwd()
            adj()  # warm both paths (workspace, partial sums) before timing
            return _time_kernel("interp", fwd, reps) + _time_kernel("spread", adj, reps)
        finally:
            plan.close()

    timings = {}
    best_t, best_s = 8, 1024
    for t in TILE_CANDIDATES:
        timings[(t, best_s)] = cost(t, best_s)
    best_t = min((k for k in timings), key=timings.get)[0]
    for sg in SEGMENT_CANDIDATES:
        if (best_t, sg) not in timings:
            timings[(best_t, sg)] = cost(best_t, sg)
    best_t, best_s = min(timings, key=timings.get)
    del ksp
    return {"sort_tile": best_t, "spread_segment": best_s,
            "timings_ms": {f"{t}x{sg}": round(v, 4) for (t, sg), v in timings.items()}}
