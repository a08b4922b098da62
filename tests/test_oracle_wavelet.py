"""Pins the db6 zero-mode wavelet oracle against PyWavelets 1.1.1 golden vectors."""

import os

import numpy as np
import pytest

from oracle import wavelet as ow

GOLD = os.path.join(os.path.dirname(__file__), "golden", "pywt_db6_zero.npz")


@pytest.fixture(scope="module")
def gold():
    return np.load(GOLD, allow_pickle=False)


def test_derived_filter_bank_matches_pywavelets(gold):
    dec_lo, dec_hi = ow.filters()
    assert np.abs(ow.daubechies_rec_lo(6) - gold["rec_lo"]).max() < 1e-14
    assert np.abs(dec_lo - gold["dec_lo"]).max() < 1e-14
    assert np.abs(dec_hi - gold["dec_hi"]).max() < 1e-14


def test_filter_orthonormality():
    h = ow.daubechies_rec_lo(6)
    assert np.sum(h) == pytest.approx(np.sqrt(2))
    for k in range(0, 6):
        assert np.dot(h[: len(h) - 2 * k], h[2 * k:]) == pytest.approx(1.0 if k == 0 else 0.0, abs=1e-14)


@pytest.mark.parametrize("i", [0, 1])
def test_wavedec3_matches_pywavelets(gold, i):
    x = gold[f"x{i}"].astype(np.float64)
    level = int(gold[f"level{i}"])
    assert ow.max_level(x.shape) == level
    c = ow.wavedec3(x)
    assert len(c) == level + 1
    err = np.abs(c[0] - gold[f"cA{i}"]).max()
    for lv in range(level):
        for k, v in c[1 + lv].items():
            err = max(err, np.abs(v - gold[f"d{i}_{lv}_{k}"]).max())
    assert err < 1e-12


def test_waverec3_matches_pywavelets_on_random_coefficients(gold):
    keys = ["aad", "ada", "add", "daa", "dad", "dda", "ddd"]
    c = [gold["rc_cA"], {k: gold[f"rc_d_0_{k}"] for k in keys}]
    rec = gold["rc_rec"]
    shape = gold["x0"].shape
    out = ow.waverec3(c, shape)
    assert np.abs(out - rec[: shape[0], : shape[1], : shape[2]]).max() < 1e-12


@pytest.mark.parametrize("shape", [(22, 23, 24), (44, 45, 46), (45, 24, 50)])
def test_perfect_reconstruction_and_adjointness(shape):
    rng = np.random.default_rng(sum(shape))
    x = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    c = ow.wavedec3(x)
    assert np.abs(ow.waverec3(c, shape) - x).max() < 1e-12
    d = [rng.standard_normal(c[0].shape) + 0j] + [
        {k: rng.standard_normal(v.shape) + 1j * rng.standard_normal(v.shape) for k, v in dd.items()}
        for dd in c[1:]]
    lhs = sum(np.vdot(a, b) for a, b in zip(ow.tree_leaves(c), ow.tree_leaves(d)))
    rhs = np.vdot(x, ow.waverec3(d, shape))
    assert abs(lhs - rhs) < 1e-10 * abs(lhs)


def test_level_rule_at_full_size():
    assert ow.max_level((80, 192, 192)) == 2
    assert ow.max_level((32, 64, 64)) == 1
    assert ow.level_shapes((80, 192, 192), 2) == [(80, 192, 192), (45, 101, 101), (28, 56, 56)]
