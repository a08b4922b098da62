"""NUFFT plugin boundary with the MI355X HIP backend (``--nufft-backend hip``).

Mirrors src/toporecon/nufft_backend.py: the ``NufftBackend`` contract (:11-52),
its error behaviour, and the ``create_nufft_backend`` registry (:119-135), with
a new registered name ``"hip"`` beside ``"sigpy"``.  ``HipNufftBackend`` keeps
the per-image, per-bin ``forward`` / ``adjoint`` calls of the reference and adds
batched entry points (all echoes x coils of a bin in one launch) used by the
HIP solvers.  Arrays are torch tensors on the GPU (the backend's ``xp`` is
``torch``); NumPy inputs are copied to the device.
"""

from __future__ import annotations

import ctypes
from abc import ABC, abstractmethod
from typing import Any, Iterable, Sequence

import numpy as np

from . import _lib


class NufftBackend(ABC):
    """Minimal contract required by the reconstruction algorithms (nufft_backend.py:11-52).

    Coordinates are stored per local motion bin, already in the internal
    ``[z, y, x]`` SigPy grid-coordinate convention.
    """

    def __init__(self, img_shape: Sequence[int]) -> None:
        self.img_shape = tuple(int(value) for value in img_shape)
        if len(self.img_shape) != 3 or any(value < 1 for value in self.img_shape):
            raise ValueError(f"img_shape must be three positive values: {img_shape}.")

    @property
    @abstractmethod
    def xp(self) -> Any:
        """Array module used by this backend."""

    @property
    @abstractmethod
    def num_bins(self) -> int:
        """Number of local trajectory bins."""

    @abstractmethod
    def to_device(self, array: Any) -> Any:
        """Move or convert an array to the backend device."""

    @abstractmethod
    def forward(self, image: Any, bin_index: int) -> Any:
        """Apply the forward NUFFT for one image and one motion bin."""

    @abstractmethod
    def adjoint(self, kspace: Any, bin_index: int) -> Any:
        """Apply the adjoint NUFFT for one k-space array and one motion bin."""

    def _validate_bin(self, bin_index: int) -> int:
        resolved = int(bin_index)
        if not 0 <= resolved < self.num_bins:
            raise IndexError(f"bin_index {resolved} is outside [0, {self.num_bins}).")
        return resolved


def resolve_device(device: Any):
    """torch.device for an int / str / torch.device / object with an ``id``."""
    import torch
    if isinstance(device, torch.device):
        dev = device
    elif isinstance(device, int):
        if device < 0:
            raise ValueError("the hip NUFFT backend runs on a GPU; device -1 (CPU) is not supported")
        dev = torch.device("cuda", device)
    elif isinstance(device, str):
        dev = torch.device(device)
    elif hasattr(device, "id"):
        return resolve_device(int(device.id))
    else:
        raise TypeError(f"unsupported device {device!r}")
    if dev.type != "cuda":
        raise ValueError("the hip NUFFT backend runs on a GPU device")
    if dev.index is None:
        dev = torch.device("cuda", torch.cuda.current_device())
    return dev


class NufftPlan:
    """Owner of a native ``tr_plan`` (coordinates, sort permutation, CSR, workspace)."""

    def __init__(self, img_shape, coords_dev, device, sort_tile=8, spread_segment=1024):
        import torch
        self.device = device
        self.img_shape = tuple(int(v) for v in img_shape)
        self._coords = [c.contiguous() for c in coords_dev]
        n = len(self._coords)
        arr_p = (ctypes.c_void_p * n)(*[c.data_ptr() for c in self._coords])
        arr_m = (ctypes.c_int64 * n)(*[int(c.shape[0]) for c in self._coords])
        handle = ctypes.c_void_p()
        lib = _lib.load()
        with torch.cuda.device(device):
            stream = _lib.stream_handle(device)
            rc = lib.tr_plan_create_geom(ctypes.byref(handle), *self.img_shape, n, arr_p, arr_m, int(sort_tile),
                                         int(spread_segment), stream)
        _lib.check(rc, "tr_plan_create_geom")
        self.handle = handle
        self.num_bins = n
        info = self.info(0)
        self.grid_shape = tuple(int(v) for v in info[3:6])

    def info(self, bin_index: int):
        out = (ctypes.c_int64 * 12)()
        _lib.call("tr_plan_info", self.handle, int(bin_index), out)
        return list(out)

    def samples(self, bin_index: int) -> int:
        return int(self.info(bin_index)[7])

    def nnz(self, bin_index: int) -> int:
        return int(self.info(bin_index)[8])

    def touched_points(self, bin_index: int) -> int:
        """Distinct grid points under the taps of bin ``bin_index`` (what its interpolation reads)."""
        return int(self.info(bin_index)[10])

    def release_workspace(self):
        _lib.call("tr_plan_release_workspace", self.handle)

    PARAMS = {"interp_xcd_chunk": 1, "spread_xcd_chunk": 2, "prune": 3, "support_points": 4,
              "live_columns": 5, "interp_kgroups": 6, "workspace_slots": 7, "lines_xcd_lines": 8,
              "spread_order": 10, "sort_tile": 11, "spread_segment": 12}
    LAUNCH_PARAMS = ("interp_xcd_chunk", "spread_xcd_chunk", "prune", "interp_kgroups", "workspace_slots",
                     "lines_xcd_lines", "spread_order")

    def set_param(self, name: str, value: int):
        """Set a launch parameter (block order only; results are unchanged).  sort_tile and
        spread_segment are fixed when the plan is built (constructor arguments)."""
        _lib.call("tr_plan_set_param", self.handle, self.PARAMS[name], int(value))

    def get_param(self, name: str) -> int:
        out = ctypes.c_int64()
        _lib.call("tr_plan_get_param", self.handle, self.PARAMS[name], ctypes.byref(out))
        return int(out.value)

    def close(self):
        if getattr(self, "handle", None):
            _lib.load().tr_plan_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class HipNufftBackend(NufftBackend):
    """Hand-written HIP NUFFT for MI355X with SigPy 0.1.27 semantics (width 4, 1.25x)."""

    def __init__(self, img_shape: Sequence[int], coordinates: Iterable[Any], device: Any) -> None:
        super().__init__(img_shape)
        import torch
        _lib.load()  # fail loudly before touching the inputs
        self.device = resolve_device(device)
        prepared = []
        for index, coordinate in enumerate(coordinates):
            array = coordinate if hasattr(coordinate, "shape") else np.asarray(coordinate)
            if array.ndim < 2 or array.shape[-1] != len(self.img_shape):
                raise ValueError(
                    f"coordinates[{index}] must end in dimension {len(self.img_shape)}, got "
                    f"{tuple(array.shape)}.")
            prepared.append(array)
        if not prepared:
            raise ValueError("At least one trajectory bin is required.")
        self.coord_shapes = [tuple(a.shape) for a in prepared]
        coords_dev = [self._as_tensor(a, torch.float32).reshape(-1, 3) for a in prepared]
        self.plan = NufftPlan(self.img_shape, coords_dev, self.device)

    # -- contract -----------------------------------------------------------
    @property
    def xp(self) -> Any:
        import torch
        return torch

    @property
    def num_bins(self) -> int:
        return self.plan.num_bins

    def to_device(self, array: Any) -> Any:
        import torch
        if isinstance(array, torch.Tensor):
            return array.to(self.device)
        a = np.asarray(array)
        return torch.from_numpy(np.ascontiguousarray(a)).to(self.device)

    def forward(self, image: Any, bin_index: int) -> Any:
        b = self._validate_bin(bin_index)
        import torch
        img = self._as_tensor(image, torch.complex64)
        if tuple(img.shape) != self.img_shape:
            raise ValueError(f"image shape {tuple(img.shape)} != img_shape {self.img_shape}")
        out = self.forward_batch(img[None], None, b)
        return out.reshape(self.coord_shapes[b][:-1])

    def adjoint(self, kspace: Any, bin_index: int) -> Any:
        b = self._validate_bin(bin_index)
        import torch
        k = self._as_tensor(kspace, torch.complex64)
        return self.adjoint_batch(k.reshape(1, 1, -1), None, b)[0]

    # -- batched entry points -------------------------------------------------
    def forward_batch(self, images, maps, bin_index: int):
        """(E, Z, Y, X) images x (C, Z, Y, X) maps -> (E, C, M_b); maps None means C = 1."""
        import torch
        b = self._validate_bin(bin_index)
        images = self._as_tensor(images, torch.complex64)
        E = int(images.shape[0])
        C = 1 if maps is None else int(maps.shape[0])
        maps_t = None if maps is None else self._as_tensor(maps, torch.complex64)
        M = self.plan.samples(b)
        out = torch.empty((E, C, M), dtype=torch.complex64, device=self.device)
        with torch.cuda.device(self.device):
            _lib.call("tr_nufft_forward", self.plan.handle, b, images.data_ptr(),
                      _lib.ptr(maps_t), E, C, out.data_ptr(), _lib.stream_handle(self.device))
        return out

    def adjoint_batch(self, kspace, maps, bin_index: int, dcf=None, out=None, accumulate=False):
        """(E, C, M_b) -> (E, Z, Y, X) = sum_c conj(S_c) NUFFT^H(dcf * k[e, c])."""
        import torch
        b = self._validate_bin(bin_index)
        k = self._as_tensor(kspace, torch.complex64)
        E, C = int(k.shape[0]), int(k.shape[1])
        if maps is None and C != 1:
            raise ValueError("maps are required for more than one coil")
        maps_t = None if maps is None else self._as_tensor(maps, torch.complex64)
        dcf_t = None if dcf is None else self._as_tensor(dcf, torch.float32).reshape(-1)
        if out is None:
            out = torch.empty((E,) + self.img_shape, dtype=torch.complex64, device=self.device)
            accumulate = False
        with torch.cuda.device(self.device):
            _lib.call("tr_nufft_adjoint", self.plan.handle, b, k.data_ptr(), _lib.ptr(dcf_t),
                      _lib.ptr(maps_t), E, C, out.data_ptr(), int(bool(accumulate)),
                      _lib.stream_handle(self.device))
        return out

    def _as_tensor(self, a, dtype):
        import torch
        if isinstance(a, torch.Tensor):
            t = a.to(device=self.device, dtype=dtype)
        else:
            t = torch.from_numpy(np.ascontiguousarray(np.asarray(a), dtype=_np_dtype(dtype))).to(self.device)
        return t.contiguous()


def _np_dtype(dtype):
    import torch
    return {torch.complex64: np.complex64, torch.float32: np.float32}[dtype]


AVAILABLE_BACKENDS = ("sigpy", "hip")


def create_nufft_backend(name: str, *, img_shape: Sequence[int], coordinates: Iterable[Any],
                         device: Any) -> NufftBackend:
    """Construct a registered backend (nufft_backend.py:119-135) with ``hip`` added."""
    normalized = name.strip().lower()
    if normalized == "hip":
        return HipNufftBackend(img_shape=img_shape, coordinates=coordinates, device=device)
    if normalized == "sigpy":
        raise RuntimeError(
            "SigPy is required for SigPyNufftBackend. "
            "Install the toporecon package dependencies first.")
    raise ValueError(f"Unknown NUFFT backend {name!r}. Available backends: sigpy, hip.")
