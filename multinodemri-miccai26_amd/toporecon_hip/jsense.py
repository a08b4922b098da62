"""JSENSE coil-sensitivity estimation on the HIP NUFFT.

Mirrors ``toporecon.preprocessing.jsense`` (src/toporecon/preprocessing/jsense.py:16-140): the same
``JsenseConfig``, ``estimate_sensitivity_maps`` and ``estimate_sensitivity_from_dataset`` with the
same arguments, shapes and errors.  The reference hands echo 0 to SigPy 0.1.27's
``sigpy.mri.app.JsenseRecon`` (jsense.py:91) with SigPy's defaults unless the config overrides them;
this module runs the same algorithm with the library's kernels:

* calibration data (JsenseRecon._get_data): the samples with max_axis |k| < calib_width / 2, their
  k-space divided by its largest magnitude and weighted by sqrt(density); ``img_shape`` and the
  calibration grid are ``estimate_shape`` (int(max - min) per axis) of all and of the selected
  coordinates;
* model: y_c = sqrt(w) NUFFT(IFFT(mps_ker_c (*) img_ker)) with (*) the 'valid' convolution of a
  Dirac-initialised image kernel (calibration grid + W - 1) and zero-initialised W^3 map kernels;
* alternating minimisation (sigpy.alg.AltMin): max_iter rounds of max_inner_iter conjugate-gradient
  steps (sigpy.alg.ConjugateGradient, warm started) on the normal equations, map kernels first;
* output (JsenseRecon._output): each map kernel zero-padded to ``img_shape`` in k-space, centred
  inverse FFT, divided by the root sum of squares over coils; then the reference resizes each map
  through k-space to the requested image shape (jsense.py:93-96).

The calibration grid is small (about 23^3 for the default width 24), so its centred (I)DFTs and
the k-space resizes are dense matrices applied along one axis at a time (``tr_js_axis_apply``);
the convolutions are ``tr_js_conv*``; the NUFFT is a one-bin plan of the library on the
calibration samples.  Sums in the conjugate-gradient steps are float64.  No CPU fallback: without
the HIP library every entry point raises.
"""

from __future__ import annotations

import math
from dataclasses import dataclass
from pathlib import Path
from typing import Any, Sequence

import numpy as np

from . import _lib
from .data import cfl
from .data.trajectory import CoordinateScaling, prepare_trajectory

# SigPy 0.1.27 JsenseRecon defaults (used when the config leaves a value as None).
DEFAULT_MPS_KER_WIDTH = 16
DEFAULT_KSP_CALIB_WIDTH = 24
DEFAULT_LAMDA = 0.0
DEFAULT_MAX_ITER = 10
DEFAULT_MAX_INNER_ITER = 10

_CG_BLOCKS = 256


@dataclass(frozen=True)
class JsenseConfig:
    """JSENSE settings (jsense.py:16-30); ``None`` keeps SigPy's JsenseRecon default."""

    readout_fraction: float = 0.95
    coordinate_scaling: CoordinateScaling = "legacy_max"
    mps_kernel_width: int | None = None
    kspace_calibration_width: int | None = None
    regularization: float | None = None
    max_iter: int | None = None
    max_inner_iter: int | None = None


def estimate_shape(coord) -> list:
    """int(max - min) per axis of the coordinates, evaluated in their own dtype."""
    c = np.asarray(coord)
    return [int(c[..., i].max() - c[..., i].min()) for i in range(c.shape[-1])]


def centered_dft_matrix(n: int, inverse: bool = False) -> np.ndarray:
    """(n, n) matrix of the centred orthonormal DFT (fftshift . fft . ifftshift, norm 'ortho'):
    F[j, m] = exp(-+2 pi i (j - n//2)(m - n//2) / n) / sqrt(n), phases reduced mod n exactly."""
    j = np.arange(n, dtype=np.int64) - n // 2
    ph = np.mod(np.outer(j, j), n).astype(np.float64) / n
    sign = 1.0 if inverse else -1.0
    return np.exp(sign * 2j * np.pi * ph) / math.sqrt(n)


def resize_matrix(n_out: int, n_in: int) -> np.ndarray:
    """(n_out, n_in) 0/1 matrix of a centred zero-pad or crop (sp.resize along one axis)."""
    si = max(n_in // 2 - n_out // 2, 0)
    so = max(n_out // 2 - n_in // 2, 0)
    cnt = min(n_in - si, n_out - so)
    r = np.zeros((n_out, n_in))
    r[so + np.arange(cnt), si + np.arange(cnt)] = 1.0
    return r


def _sum(partial) -> float:
    return float(partial.cpu().numpy().sum())  # block sums added in block order on the host


class JsenseHip:
    """JsenseRecon on one GPU: device state of one calibration problem."""

    def __init__(self, y, coord, weights, device: Any = 0, mps_ker_width: int = DEFAULT_MPS_KER_WIDTH,
                 ksp_calib_width: float = DEFAULT_KSP_CALIB_WIDTH, lamda: float = DEFAULT_LAMDA,
                 max_iter: int = DEFAULT_MAX_ITER, max_inner_iter: int = DEFAULT_MAX_INNER_ITER):
        import torch
        from .nufft_backend import NufftPlan, resolve_device
        _lib.load()  # fail loudly before touching the inputs
        self.dev = resolve_device(device)
        y = np.asarray(y)
        coord = np.asarray(coord)
        if y.ndim < 2 or coord.shape[:-1] != y.shape[1:] or coord.shape[-1] != 3:
            raise ValueError(f"k-space {y.shape} and coordinates {coord.shape} do not match.")
        w_all = None if weights is None else np.asarray(weights)
        self.W = int(mps_ker_width)
        self.lamda = float(lamda)
        self.max_iter = int(max_iter)
        self.max_inner_iter = int(max_inner_iter)
        self.C = int(y.shape[0])
        self.img_shape = estimate_shape(coord)
        sel = np.amax(np.abs(coord), axis=-1) < ksp_calib_width / 2
        if not np.any(sel):
            raise ValueError(f"no k-space samples inside the calibration region |k| < {ksp_calib_width / 2}.")
        c = np.ascontiguousarray(coord[sel], dtype=np.float32)
        ycal = y[:, sel]
        scale = np.abs(ycal).max()
        if w_all is None:
            w = np.ones(c.shape[0], np.float32)
            ycal = ycal / scale
        else:
            w = np.asarray(w_all[sel], dtype=np.float32)
            ycal = w ** 0.5 * ycal / scale
        self.grd_shape = estimate_shape(c)
        if min(self.grd_shape) < 1:
            raise ValueError(f"calibration samples span an empty grid {self.grd_shape}.")
        self.ker_shape = [g + self.W - 1 for g in self.grd_shape]
        self.M = int(c.shape[0])
        C, M = self.C, self.M
        torch.cuda.set_device(self.dev)
        self.stream = _lib.stream_handle(self.dev)
        s = self.stream
        f32, c64 = torch.float32, torch.complex64
        coord_t = torch.from_numpy(c).to(self.dev)
        self.plan = NufftPlan(self.grd_shape, [coord_t], self.dev)
        y_t = torch.from_numpy(np.ascontiguousarray(ycal, dtype=np.complex64)).to(self.dev)
        self.y_sorted = torch.empty((M, C), dtype=c64, device=self.dev)
        _lib.call("tr_plan_gather_samples", self.plan.handle, 0, y_t.data_ptr(), C, self.y_sorted.data_ptr(), s)
        sw = (w ** 0.5).astype(np.float32)
        self.sqrt_w = torch.empty(M, dtype=f32, device=self.dev)
        self.w = torch.empty(M, dtype=f32, device=self.dev)
        sw_t = torch.from_numpy(sw).to(self.dev)
        ww_t = torch.from_numpy((sw * sw).astype(np.float32)).to(self.dev)
        _lib.call("tr_plan_gather_weights", self.plan.handle, 0, sw_t.data_ptr(), self.sqrt_w.data_ptr(), s)
        _lib.call("tr_plan_gather_weights", self.plan.handle, 0, ww_t.data_ptr(), self.w.data_ptr(), s)
        # dense centred DFT matrices of the calibration grid, per axis
        self._fft = [self._mat(centered_dft_matrix(n)) for n in self.grd_shape]
        self._ifft = [self._mat(centered_dft_matrix(n, inverse=True)) for n in self.grd_shape]
        self.zshape = (C,) + tuple(self.grd_shape)
        self._za = torch.empty(self.zshape, dtype=c64, device=self.dev)
        self._zb = torch.empty(self.zshape, dtype=c64, device=self.dev)
        self._zc = torch.empty(self.zshape, dtype=c64, device=self.dev)
        self._ksp = torch.empty((M, C), dtype=c64, device=self.dev)
        self._ones = torch.ones(tuple(self.grd_shape), dtype=c64, device=self.dev)
        nwork = int(_lib.load().tr_js_work_count(*self.ker_shape, C, self.W))
        if nwork < 0:
            raise ValueError(f"unsupported JSENSE shape: image kernel {self.ker_shape}, width {self.W}: "
                             f"{_lib.last_error()}")
        self._work = torch.empty(nwork, dtype=c64, device=self.dev)
        self._partial = torch.empty(_CG_BLOCKS, dtype=torch.float64, device=self.dev)
        # z0 = FFT(NUFFT^H(sqrt(w) y)): the data side of both right-hand sides
        self.z0 = torch.empty(self.zshape, dtype=c64, device=self.dev)
        _lib.call("tr_data_adjoint", self.plan.handle, 0, self.y_sorted.data_ptr(), self.sqrt_w.data_ptr(), None,
                  C, 1, self._za.data_ptr(), 0, s)
        self._dft3(self._za, self.z0, self._zb, self._fft)
        self.img_ker = None
        self.mps_ker = None

    # -- small helpers ----------------------------------------------------------------------
    def _mat(self, m):
        import torch
        return torch.from_numpy(np.ascontiguousarray(m, dtype=np.complex64)).to(self.dev)

    def _axis(self, x, shape, axis, mat, n_out, out):
        """out = mat applied along spatial `axis` (0 = z, 1 = y, 2 = x) of x viewed as (batch, z, y, x)."""
        outer = int(np.prod(shape[:axis + 1]))  # batch and the spatial axes before `axis`
        inner = int(np.prod(shape[axis + 2:]))  # the spatial axes after it
        _lib.call("tr_js_axis_apply", x.data_ptr(), outer, int(shape[axis + 1]), inner, mat.data_ptr(), int(n_out),
                  out.data_ptr(), self.stream)

    def _dft3(self, src, dst, tmp, mats):
        """dst = centred 3-D (I)DFT of src (C, gz, gy, gx); tmp is scratch; src is left unchanged."""
        shp = self.zshape
        self._axis(src, shp, 0, mats[0], shp[1], dst)
        self._axis(dst, shp, 1, mats[1], shp[2], tmp)
        self._axis(tmp, shp, 2, mats[2], shp[3], dst)

    def _conv(self, img, ker, out):
        _lib.call("tr_js_conv", img.data_ptr(), *self.ker_shape, ker.data_ptr(), self.C, self.W, out.data_ptr(),
                  self._work.data_ptr(), self.stream)

    def _simulate(self, z):
        """self._ksp (M, C) sorted = NUFFT of the C calibration-grid images z.  The images go in as
        the C "coil maps" of one all-ones image: the products 1 * z_c are z_c exactly, and the
        forward x pass then runs 16 channels per FFT chunk instead of one image at a time."""
        _lib.call("tr_data_simulate", self.plan.handle, 0, self._ones.data_ptr(), z.data_ptr(), 1, self.C,
                  self._ksp.data_ptr(), self.stream)

    def _data_normal(self, z, out):
        """out = FFT(NUFFT^H(w NUFFT(IFFT(z)))) on the calibration grid; z is left unchanged."""
        s = self.stream
        self._dft3(z, self._zb, self._zc, self._ifft)
        self._simulate(self._zb)
        _lib.call("tr_data_adjoint", self.plan.handle, 0, self._ksp.data_ptr(), self.w.data_ptr(), None, self.C, 1,
                  self._zb.data_ptr(), 0, s)
        self._dft3(self._zb, out, self._zc, self._fft)

    # -- the two normal operators (sigpy.mri.linop ConvImage / ConvSense, then .H) ---------------
    def normal_mps(self, v, out):
        self._conv(self.img_ker, v, self._za)
        self._data_normal(self._za, self._za)
        _lib.call("tr_js_conv_adj_ker", self._za.data_ptr(), self.img_ker.data_ptr(), *self.ker_shape, self.C,
                  self.W, out.data_ptr(), self._work.data_ptr(), self.stream)

    def normal_img(self, v, out):
        self._conv(v, self.mps_ker, self._za)
        self._data_normal(self._za, self._za)
        _lib.call("tr_js_conv_adj_img", self._za.data_ptr(), *self.ker_shape, self.mps_ker.data_ptr(), self.C,
                  self.W, out.data_ptr(), self._work.data_ptr(), self.stream)

    def rhs_mps(self, out):
        _lib.call("tr_js_conv_adj_ker", self.z0.data_ptr(), self.img_ker.data_ptr(), *self.ker_shape, self.C,
                  self.W, out.data_ptr(), self._work.data_ptr(), self.stream)

    def rhs_img(self, out):
        _lib.call("tr_js_conv_adj_img", self.z0.data_ptr(), *self.ker_shape, self.mps_ker.data_ptr(), self.C,
                  self.W, out.data_ptr(), self._work.data_ptr(), self.stream)

    # -- conjugate gradient (sigpy.alg.ConjugateGradient with P = None, tol = 0) -------------------
    def _apply(self, normal, v, out):
        normal(v, out)
        if self.lamda != 0.0:
            _lib.call("tr_js_axpy", out.data_ptr(), v.data_ptr(), v.numel(), self.lamda, self.stream)

    def _cg(self, normal, rhs, x, iters: int, x_is_zero: bool):
        import torch
        s = self.stream
        n = x.numel()
        part = self._partial
        r = torch.empty_like(x)
        rhs(r)
        ap = torch.empty_like(x)
        p = torch.empty_like(x)
        if x_is_zero:  # A 0 = 0 exactly: r = b
            _lib.call("tr_cdot_re", r.data_ptr(), r.data_ptr(), n, part.data_ptr(), _CG_BLOCKS, s)
        else:
            self._apply(normal, x, ap)
            # r = b - A x (p is scratch for the x-side update of the fused step)
            _lib.call("tr_cg_update", p.data_ptr(), r.data_ptr(), ap.data_ptr(), ap.data_ptr(), n, 1.0,
                      part.data_ptr(), _CG_BLOCKS, s)
        rr = _sum(part)
        _lib.call("tr_cscale", r.data_ptr(), n, 1.0, p.data_ptr(), s)
        for it in range(iters):
            if not rr > 0.0:
                break
            self._apply(normal, p, ap)
            _lib.call("tr_cdot_re", p.data_ptr(), ap.data_ptr(), n, part.data_ptr(), _CG_BLOCKS, s)
            pap = _sum(part)
            if not pap > 0.0:
                break
            alpha = rr / pap
            _lib.call("tr_cg_update", x.data_ptr(), r.data_ptr(), p.data_ptr(), ap.data_ptr(), n, alpha,
                      part.data_ptr(), _CG_BLOCKS, s)
            if it == iters - 1:
                break
            rr_new = _sum(part)
            _lib.call("tr_cg_direction", p.data_ptr(), r.data_ptr(), n, rr_new / rr, s)
            rr = rr_new

    def run(self, show_progress: bool = False, history=None):
        """Alternating minimisation; returns (img_ker, mps_ker) on the device."""
        import torch
        c64 = torch.complex64
        Iz, Iy, Ix = self.ker_shape
        self.img_ker = torch.zeros((Iz, Iy, Ix), dtype=c64, device=self.dev)
        self.img_ker[Iz // 2, Iy // 2, Ix // 2] = 1.0
        self.mps_ker = torch.zeros((self.C, self.W, self.W, self.W), dtype=c64, device=self.dev)
        mps_zero = True
        rounds = range(self.max_iter)
        if show_progress:
            try:
                from tqdm import tqdm
                rounds = tqdm(rounds, desc="JsenseRecon")
            except ImportError:
                pass
        for _ in rounds:
            self._cg(self.normal_mps, self.rhs_mps, self.mps_ker, self.max_inner_iter, mps_zero)
            mps_zero = False
            if history is not None:
                history.append(self.objective())
            self._cg(self.normal_img, self.rhs_img, self.img_ker, self.max_inner_iter, False)
            if history is not None:
                history.append(self.objective())
        return self.img_ker, self.mps_ker

    def objective(self) -> float:
        """||sqrt(w) A(img_ker, mps_ker) - y||^2 (float64, for tests and diagnostics)."""
        import torch
        self._conv(self.img_ker, self.mps_ker, self._za)
        self._dft3(self._za, self._zb, self._zc, self._ifft)
        self._simulate(self._zb)
        a = self._ksp.to(torch.complex128) * self.sqrt_w.to(torch.float64)[:, None]
        return float(torch.sum(torch.abs(a - self.y_sorted.to(torch.complex128)) ** 2))

    # -- output -----------------------------------------------------------------------------
    def maps(self, image_shape_zyx=None):
        """(C, Z, Y, X) complex64 maps on the device: JsenseRecon._output at ``img_shape``, then the
        reference's k-space resize to ``image_shape_zyx`` (jsense.py:93-96) where it differs."""
        import torch
        C, W = self.C, self.W
        n = [int(v) for v in self.img_shape]
        mats = [self._mat(centered_dft_matrix(n[a], inverse=True) @ resize_matrix(n[a], W)) for a in range(3)]
        c64 = torch.complex64
        t1 = torch.empty((C, W, W, n[2]), dtype=c64, device=self.dev)
        self._axis(self.mps_ker, (C, W, W, W), 2, mats[2], n[2], t1)
        t2 = torch.empty((C, W, n[1], n[2]), dtype=c64, device=self.dev)
        self._axis(t1, (C, W, W, n[2]), 1, mats[1], n[1], t2)
        del t1
        out = torch.empty((C, n[0], n[1], n[2]), dtype=c64, device=self.dev)
        self._axis(t2, (C, W, n[1], n[2]), 0, mats[0], n[0], out)
        del t2
        _lib.call("tr_js_rss_normalize", out.data_ptr(), C, n[0] * n[1] * n[2], self.stream)
        if image_shape_zyx is None:
            return out
        tgt = [int(v) for v in image_shape_zyx]
        cur = list(n)
        for a in range(3):
            if tgt[a] == cur[a]:
                continue  # ifft(fft(x)) along this axis is the identity
            m = self._mat(centered_dft_matrix(tgt[a], inverse=True) @ resize_matrix(tgt[a], cur[a])
                          @ centered_dft_matrix(cur[a]))
            nxt = list(cur)
            nxt[a] = tgt[a]
            res = torch.empty([C] + nxt, dtype=c64, device=self.dev)
            self._axis(out, [C] + cur, a, m, tgt[a], res)
            out, cur = res, nxt
        return out

    def close(self):
        if getattr(self, "plan", None) is not None:
            self.plan.close()
            self.plan = None


def estimate_sensitivity_maps(kspace, coordinates_zyx, density, *, image_shape_zyx: Sequence[int],
                              device: Any = 0, show_progress: bool = False,
                              config: JsenseConfig | None = None) -> np.ndarray:
    """Canonical ``[coil, z, y, x]`` complex64 sensitivity maps from echo 0 of ``[echo, coil,
    trajectory, readout]`` k-space (jsense.py:33-97)."""
    data = np.asarray(kspace)
    if data.ndim != 4:
        raise ValueError(f"k-space must have shape [echo,coil,trajectory,readout], got {data.shape}.")
    coordinates = np.asarray(coordinates_zyx)
    weights = np.asarray(density)
    if coordinates.shape[:-1] != data.shape[-2:]:
        raise ValueError(f"trajectory shape {coordinates.shape} does not match k-space {data.shape[-2:]}.")
    if weights.shape != data.shape[-2:]:
        raise ValueError(f"density shape {weights.shape} does not match k-space {data.shape[-2:]}.")
    cfg = config or JsenseConfig()
    target = tuple(int(v) for v in image_shape_zyx)

    def pick(v, default):
        return default if v is None else v

    solver = JsenseHip(data[0], coordinates, weights, device,
                       mps_ker_width=pick(cfg.mps_kernel_width, DEFAULT_MPS_KER_WIDTH),
                       ksp_calib_width=pick(cfg.kspace_calibration_width, DEFAULT_KSP_CALIB_WIDTH),
                       lamda=pick(cfg.regularization, DEFAULT_LAMDA),
                       max_iter=pick(cfg.max_iter, DEFAULT_MAX_ITER),
                       max_inner_iter=pick(cfg.max_inner_iter, DEFAULT_MAX_INNER_ITER))
    try:
        solver.run(show_progress=show_progress)
        maps = solver.maps(target)
        out = maps.cpu().numpy().astype(np.complex64, copy=False)
    finally:
        solver.close()
    return out


def estimate_sensitivity_from_dataset(input_directory, output_directory, *, device: Any = 0,
                                      fov_scale_zyx: Sequence[float] = (1.0, 1.0, 1.0),
                                      show_progress: bool = False, config: JsenseConfig | None = None) -> Path:
    """Estimate and write ``mps.hdr/.cfl`` (1, 1, 1, 1, 1, C, Z, Y, X) without modifying the raw data
    (jsense.py:100-140)."""
    cfg = config or JsenseConfig()
    root = Path(input_directory)
    names = ("ksp.hdr", "ksp.cfl", "ktraj.hdr", "ktraj.cfl", "dens.hdr", "dens.cfl", "imageDim.txt",
             "voxelSize.txt")
    missing = [str(root / n) for n in names if not (root / n).is_file()]
    if missing:
        raise FileNotFoundError(f"Dataset is missing required files: {', '.join(missing)}")
    shape_zyx = [int(v) for v in cfl.read_zyx_txt(root / "imageDim.txt")]
    voxel_zyx = cfl.read_zyx_txt(root / "voxelSize.txt")
    traj = prepare_trajectory(np.squeeze(cfl.read_cfl(root / "ktraj")).real,
                              np.squeeze(cfl.read_cfl(root / "dens")).real,
                              image_shape_zyx=shape_zyx, voxel_size_zyx=voxel_zyx,
                              readout_fraction=cfg.readout_fraction, fov_scale_zyx=fov_scale_zyx,
                              scaling=cfg.coordinate_scaling)
    E, C, T, R = cfl.kspace_layout(cfl.read_cfl_header(root / "ksp"))
    ksp = cfl.read_cfl(root / "ksp", mmap=True).reshape(E, C, T, R)[..., :traj.readout]
    maps = estimate_sensitivity_maps(ksp, traj.coordinates_zyx, traj.density, image_shape_zyx=traj.image_shape_zyx,
                                     device=device, show_progress=show_progress, config=cfg)
    out = Path(output_directory) / "mps"
    cfl.write_cfl(out, maps.reshape((1, 1, 1, 1, 1) + maps.shape))
    return out
