"""CPU checks of the learned warm start's packaging (no GPU needed)."""

import os

import numpy as np


def test_committed_weights_load_safely_and_have_the_documented_shapes():
    from toporecon_hip.learned.net import CIN_REAL, COUT_REAL, default_weights_path
    path = default_weights_path()
    assert os.path.exists(path)
    with np.load(path, allow_pickle=False) as f:
        for li in range(5):
            assert f[f"w{li}"].shape == (27, CIN_REAL[li], COUT_REAL[li])
            assert f[f"b{li}"].shape == (COUT_REAL[li],)
            assert np.all(np.isfinite(f[f"w{li}"]))
        assert list(f["meta_train_shape"]) == [64, 96, 96]
        # training beat the identity map it starts from
        assert float(f["meta_final_loss"]) < 0.6 * float(f["meta_identity_loss"])
