"""Pins of the float64 JSENSE oracle (oracle/jsense.py) and of the host logic around the HIP path.

SigPy (whose JsenseRecon the reference calls) is not installed, so the oracle is pinned by what the
mathematics fixes: the 'valid' convolution against SciPy's ``convolve`` and a scalar loop; both
convolution adjoints and the whole weighted operator by <A x, y> = <x, A^H y>; the centred FFT
against its defining sum; conjugate gradients against a direct solve; the objective never rising
over the alternating half-steps; data generated inside the model fitted; the first half-step
against the explicit least-squares solution; closed forms of the
output step (a centred Dirac kernel gives a constant map; a constant map resized through k-space
scales by sqrt(N / N')); a hand-built calibration selection.
"""

import math

import numpy as np
import pytest
from scipy import signal

from oracle import jsense as oj
from oracle import nufft as on


def crandn(rng, *shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


def vdot(a, b):
    return np.vdot(a.ravel(), b.ravel())


def conv_loop(img, ker):
    """Scalar restatement of the 'valid' convolution: out[n] = sum_k ker[k] img[n + W - 1 - k]."""
    W = ker.shape
    G = tuple(i - w + 1 for i, w in zip(img.shape, W))
    out = np.zeros(G, complex)
    for n in np.ndindex(*G):
        acc = 0j
        for k in np.ndindex(*W):
            acc += ker[k] * img[tuple(n[d] + W[d] - 1 - k[d] for d in range(3))]
        out[n] = acc
    return out


def test_conv_valid_matches_scipy_and_scalar_loop():
    rng = np.random.default_rng(0)
    img = crandn(rng, 9, 8, 7)
    ker = crandn(rng, 3, 4, 4, 4)
    out = oj.conv_valid(img[None], ker)
    assert out.shape == (3, 6, 5, 4)
    for c in range(3):
        np.testing.assert_allclose(out[c], signal.convolve(img, ker[c], mode="valid"), rtol=1e-12, atol=1e-12)
    small_img, small_ker = crandn(rng, 5, 4, 6), crandn(rng, 1, 2, 3, 2)
    np.testing.assert_allclose(oj.conv_valid(small_img[None], small_ker)[0], conv_loop(small_img, small_ker[0]),
                               rtol=1e-12, atol=1e-12)


def test_conv_adjoints():
    rng = np.random.default_rng(1)
    img = crandn(rng, 10, 9, 11)
    ker = crandn(rng, 4, 5, 5, 5)
    z = crandn(rng, 4, 6, 5, 7)
    lhs = vdot(oj.conv_valid(img[None], ker), z)
    np.testing.assert_allclose(vdot(img, oj.conv_valid_adj_img(z, ker, img.shape)), lhs, rtol=1e-12)
    np.testing.assert_allclose(vdot(ker, oj.conv_valid_adj_ker(z, img, ker.shape)), lhs, rtol=1e-12)


@pytest.mark.parametrize("shape", [(5, 6, 7), (4, 4, 9)])
def test_centered_fft_defining_sum(shape):
    rng = np.random.default_rng(2)
    x = crandn(rng, *shape)
    pos = [np.arange(n) - n // 2 for n in shape]
    pz, py, px = np.meshgrid(*pos, indexing="ij")
    want = np.zeros(shape, complex)
    for j in np.ndindex(*shape):
        ph = (pos[0][j[0]] * pz / shape[0] + pos[1][j[1]] * py / shape[1] + pos[2][j[2]] * px / shape[2])
        want[j] = np.sum(x * np.exp(-2j * np.pi * ph)) / math.sqrt(np.prod(shape))
    np.testing.assert_allclose(oj.fft_c(x), want, rtol=1e-10, atol=1e-12)
    np.testing.assert_allclose(oj.ifft_c(oj.fft_c(x)), x, rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(np.linalg.norm(oj.fft_c(x)), np.linalg.norm(x), rtol=1e-12)


def test_cg_matches_direct_solve():
    rng = np.random.default_rng(3)
    n = 7
    a = crandn(rng, n, n)
    A = a.conj().T @ a + 0.5 * np.eye(n)
    b = crandn(rng, n)
    x = oj.cg(lambda v: A @ v, b, np.zeros(n, complex), n)
    np.testing.assert_allclose(x, np.linalg.solve(A, b), rtol=1e-8, atol=1e-10)
    # warm start and lamda: (A + lamda I) x = b
    x = oj.cg(lambda v: A @ v, b, crandn(rng, n), n, lamda=0.3)
    np.testing.assert_allclose(x, np.linalg.solve(A + 0.3 * np.eye(n), b), rtol=1e-8, atol=1e-10)
    # two steps equal the 2-D Krylov minimiser of the A-norm error
    x2 = oj.cg(lambda v: A @ v, b, np.zeros(n, complex), 2)
    K = np.stack([b, A @ b], axis=1)
    c = np.linalg.solve(K.conj().T @ A @ K, K.conj().T @ b)
    np.testing.assert_allclose(x2, K @ c, rtol=1e-8, atol=1e-10)


def kooshball(spokes, readout, kmax, seed):
    """Centre-out 3-D radial trajectory (spokes, readout, 3) in [z, y, x] with |k| <= kmax."""
    rng = np.random.default_rng(seed)
    d = rng.standard_normal((spokes, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    r = np.linspace(0, 1, readout)
    return (r[None, :, None] * d[:, None, :] * np.asarray(kmax, float)).astype(np.float32)


def small_problem(seed=4, coils=3, width=5, calib=10, shape=(12, 12, 12)):
    """A calibration problem with data generated inside the JSENSE model."""
    rng = np.random.default_rng(seed)
    coord = kooshball(300, 16, np.asarray(shape) / 2, seed)
    dens = (0.2 + np.linalg.norm(coord, axis=-1) ** 2).astype(np.float32)
    y0 = np.zeros((coils,) + coord.shape[:-1], np.complex64)
    prob = oj.calibration(y0 + 1, coord, dens, mps_ker_width=width, ksp_calib_width=calib)
    img_ker, mps_ker = oj.init_vars(prob, coils)
    g = np.meshgrid(*[np.arange(s) - s // 2 for s in mps_ker.shape[1:]], indexing="ij")
    env = np.exp(-sum(v ** 2 for v in g) / 4.0)
    true_mps = crandn(rng, *mps_ker.shape) * env
    gi = np.meshgrid(*[np.arange(s) - s // 2 for s in img_ker.shape], indexing="ij")
    true_img = crandn(rng, *img_ker.shape) * np.exp(-sum(v ** 2 for v in gi) / 8.0)
    prob.y = oj.forward(prob, true_img, true_mps)
    return prob, true_img, true_mps


def test_whole_operator_adjoints():
    prob, img, mps = small_problem()
    rng = np.random.default_rng(5)
    r = crandn(rng, *prob.y.shape)
    a = oj.forward(prob, img, mps)
    np.testing.assert_allclose(vdot(img, oj.adjoint_img(prob, r, mps, img.shape)), vdot(a, r), rtol=1e-10)
    np.testing.assert_allclose(vdot(mps, oj.adjoint_ker(prob, r, img, mps.shape)), vdot(a, r), rtol=1e-10)


def test_objective_monotone_and_model_data_fitted():
    prob, img, mps = small_problem()
    hist = []
    oj.jsense(prob, max_iter=6, max_inner_iter=8, history=hist)
    y2 = float(np.vdot(prob.y, prob.y).real)
    assert all(b <= a * (1 + 1e-12) for a, b in zip(hist, hist[1:])), hist
    assert hist[0] < 0.5 * y2 and hist[-1] < 0.02 * y2, (hist, y2)


def test_first_half_step_is_the_map_kernel_least_squares_solution():
    """With as many CG steps as map-kernel unknowns (C W^3 = 24), the first half-step of the
    alternating minimisation (map kernels, image kernel still the Dirac) is the exact
    least-squares solution, built here column by column from the forward operator."""
    prob, _, _ = small_problem(width=2)
    img0, mps0 = oj.init_vars(prob, 3)
    n = mps0.size
    cols = []
    for j in range(n):
        e = np.zeros(n, complex)
        e[j] = 1.0
        cols.append(oj.forward(prob, img0, e.reshape(mps0.shape)).ravel())
    A = np.stack(cols, axis=1)
    want = np.linalg.lstsq(A, prob.y.ravel(), rcond=None)[0].reshape(mps0.shape)
    _, mps1 = oj.jsense(prob, max_iter=1, max_inner_iter=n)
    # the second half-step only moves the image kernel
    np.testing.assert_allclose(mps1, want, rtol=1e-6, atol=1e-8 * np.abs(want).max())


def test_maps_from_centred_dirac_kernels_are_constant():
    a = np.array([1.0 + 2.0j, -0.5j, 3.0])
    ker = np.zeros((3, 6, 6, 6), complex)
    ker[:, 3, 3, 3] = a
    m = oj.maps_from_kernel(ker, [5, 8, 7])
    want = a / np.linalg.norm(a)
    np.testing.assert_allclose(m, np.broadcast_to(want[:, None, None, None], m.shape), rtol=1e-12, atol=1e-14)


def test_to_shape_of_a_constant_scales_by_sqrt_of_sizes():
    m = np.full((2, 5, 6, 7), 0.3 - 0.4j)
    out = oj.to_shape(m, (6, 6, 9))
    assert out.shape == (2, 6, 6, 9)
    np.testing.assert_allclose(out, (0.3 - 0.4j) * math.sqrt(5 * 6 * 7 / (6 * 6 * 9)), rtol=1e-12, atol=1e-14)
    np.testing.assert_allclose(oj.to_shape(m, (5, 6, 7)), m, rtol=1e-12, atol=1e-14)


def test_calibration_selection_known_answer():
    coord = np.array([[[0.0, 0.0, 0.0], [1.0, -2.0, 3.0], [11.9, 0.0, 0.0]],
                      [[12.0, 0.0, 0.0], [-5.0, 6.0, -7.5], [0.0, -30.0, 0.0]]], np.float32)
    y = np.zeros((2, 2, 3), np.complex64)
    y[0] = [[1, 2j, -4], [100, 3, 7]]
    y[1] = [[0.5, 1, 1], [-100j, 2, 8j]]
    w = np.array([[4.0, 9.0, 1.0], [1.0, 16.0, 1.0]], np.float32)
    p = oj.calibration(y, coord, w, mps_ker_width=4, ksp_calib_width=24)
    # |k|_inf < 12 keeps (0,0), (0,1), (0,2), (1,1); |y| max over those = 4 (the 100s are dropped)
    np.testing.assert_array_equal(p.coord, coord.reshape(-1, 3)[[0, 1, 2, 4]])
    want = np.array([[1 * 2, 2j * 3, -4 * 1, 3 * 4], [0.5 * 2, 1 * 3, 1 * 1, 2 * 4]]) / 4
    np.testing.assert_allclose(p.y, want, rtol=1e-7)
    assert p.img_shape == [int(np.float32(12.0) - np.float32(-5.0)), 36, int(np.float32(3.0) - np.float32(-7.5))]
    assert p.grd_shape == [16, 8, 10]


# ---- host logic of the HIP path (no GPU calls) --------------------------------------------------

def test_host_dft_and_resize_matrices_match_oracle():
    from toporecon_hip import jsense as hj
    rng = np.random.default_rng(6)
    for n in (1, 2, 7, 8, 23):
        x = crandn(rng, n)
        f = hj.centered_dft_matrix(n)
        fi = hj.centered_dft_matrix(n, inverse=True)
        ref = np.fft.fftshift(np.fft.fft(np.fft.ifftshift(x), norm="ortho"))
        np.testing.assert_allclose(f @ x, ref, rtol=1e-12, atol=1e-12)
        np.testing.assert_allclose(fi @ (f @ x), x, rtol=1e-12, atol=1e-12)
    for n_out, n_in in ((9, 4), (4, 9), (8, 5), (5, 8), (6, 6)):
        x = crandn(rng, 3, 3, n_in)
        np.testing.assert_allclose(x @ hj.resize_matrix(n_out, n_in).T,
                                   on._resize_center(x, (3, 3, n_out)), rtol=0, atol=0)


def test_host_estimate_shape_matches_oracle():
    from toporecon_hip import jsense as hj
    c = kooshball(50, 10, (40.0, 96.0, 96.0), 7)
    assert hj.estimate_shape(c) == oj.estimate_shape(c)


def test_estimate_sensitivity_maps_rejects_bad_shapes():
    from toporecon_hip.jsense import estimate_sensitivity_maps
    k = np.zeros((1, 2, 5, 4), np.complex64)
    with pytest.raises(ValueError, match="echo,coil,trajectory,readout"):
        estimate_sensitivity_maps(k[0], np.zeros((5, 4, 3)), np.ones((5, 4)), image_shape_zyx=(4, 4, 4))
    with pytest.raises(ValueError, match="trajectory shape"):
        estimate_sensitivity_maps(k, np.zeros((5, 3, 3)), np.ones((5, 4)), image_shape_zyx=(4, 4, 4))
    with pytest.raises(ValueError, match="density shape"):
        estimate_sensitivity_maps(k, np.zeros((5, 4, 3)), np.ones((5, 3)), image_shape_zyx=(4, 4, 4))


def test_prepare_trajectory_known_answers():
    from toporecon_hip.data.trajectory import prepare_trajectory
    raw = np.zeros((2, 10, 3), np.float32)  # [x, y, z]
    raw[:, :, 0] = np.linspace(-0.5, 0.25, 10)
    raw[:, :, 1] = 0.1
    raw[:, :, 2] = np.linspace(0.0, 0.4, 10)
    clean = raw.copy()
    raw[1, 8, :] = 2000.0  # beyond the instability threshold: 9 kept readouts shrink to 8
    dens = np.arange(20, dtype=np.float32).reshape(2, 10)
    t = prepare_trajectory(raw, dens, image_shape_zyx=(20, 30, 40), voxel_size_zyx=(2.0, 1.0, 0.5),
                           readout_fraction=0.95, scaling="legacy_max", fov_scale_zyx=(1.0, 1.5, 1.0))
    assert t.readout == 8 and t.coordinates_zyx.shape == (2, 8, 3)
    assert t.image_shape_zyx == (20, 45, 40) and t.original_image_shape_zyx == (20, 30, 40)
    c = t.coordinates_zyx
    # legacy_max: each axis scaled to max = shape / 2, then the FOV scale
    np.testing.assert_allclose(c[..., 0].max(), 10.0, rtol=1e-6)
    np.testing.assert_allclose(c[..., 1].max(), 15.0 * 1.5, rtol=1e-6)
    np.testing.assert_allclose(c[..., 2].max(), 20.0, rtol=1e-6)
    xs = np.linspace(-0.5, 0.25, 10)[:8]  # max -0.5 + 0.75 * 7 / 9 = 1 / 12
    np.testing.assert_allclose(c[0, :, 2], xs * 20.0 * 12.0, rtol=1e-5)
    np.testing.assert_array_equal(t.density, dens[:, :8])
    p = prepare_trajectory(clean, dens, image_shape_zyx=(20, 30, 40), voxel_size_zyx=(2.0, 1.0, 0.5),
                           readout_fraction=1.0, scaling="physical")
    assert p.readout == 10
    np.testing.assert_allclose(p.coordinates_zyx, clean[:, :, ::-1] * np.array([40.0, 30.0, 20.0], np.float32),
                               rtol=1e-6)
    with pytest.raises(ValueError, match="fov_scale"):
        prepare_trajectory(raw, dens, image_shape_zyx=(2, 2, 2), voxel_size_zyx=(1, 1, 1), readout_fraction=1.0,
                           fov_scale_zyx=(0.5, 1, 1))
    with pytest.raises(ValueError, match="readout_fraction"):
        prepare_trajectory(raw, dens, image_shape_zyx=(2, 2, 2), voxel_size_zyx=(1, 1, 1), readout_fraction=0.0)


def test_cli_jsense_arguments_and_missing_dataset(tmp_path):
    from toporecon_hip.mains import cli_main
    with pytest.raises(FileNotFoundError, match="missing required files"):
        cli_main(["jsense", str(tmp_path / "nothing"), "--work-dir", str(tmp_path / "w"), "--device", "0",
                  "--readout-fraction", "0.9", "--coordinate-scaling", "physical", "--fov-scale", "1", "1.5", "1",
                  "--show-progress"])
    with pytest.raises(SystemExit):
        cli_main(["jsense", str(tmp_path), "--work-dir", str(tmp_path), "--coordinate-scaling", "other"])


def test_prepare_trajectory_reference_known_answers():
    """The reference's own cases (tests/unit/test_trajectory.py:11-44): physical scaling reverses
    [x, y, z] once and applies voxel x shape x FOV ((1, 2, 3) -> (30, 10, 3) with the FOV doubling z),
    readout fraction 0.75 of 4 -> 3, the FOV-scaled shape; a density of the wrong shape is refused."""
    from toporecon_hip.data.trajectory import prepare_trajectory
    raw = np.zeros((2, 4, 3), np.float32)
    raw[..., 0], raw[..., 1], raw[..., 2] = 1.0, 2.0, 3.0
    r = prepare_trajectory(raw, np.ones((2, 4), np.float32), image_shape_zyx=(10, 20, 30),
                           voxel_size_zyx=(0.5, 0.25, 0.1), readout_fraction=0.75, fov_scale_zyx=(2.0, 1.0, 1.0),
                           scaling="physical", unstable_coordinate_threshold=None)
    assert r.readout == 3 and r.image_shape_zyx == (20, 20, 30)
    np.testing.assert_allclose(r.coordinates_zyx[0, 0], (30.0, 10.0, 3.0), rtol=1e-6)
    with pytest.raises(ValueError, match="density shape"):
        prepare_trajectory(np.zeros((2, 4, 3), np.float32), np.zeros((2, 3), np.float32), image_shape_zyx=(4, 4, 4),
                           voxel_size_zyx=(1, 1, 1), readout_fraction=1.0)
