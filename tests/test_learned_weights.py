"""Weight files of the learned preconditioner (host only)."""
import json

import numpy as np

from stream_amd import learned


def test_four_feature_files_widen_with_a_zero_z_row(tmp_path):
    rng = np.random.default_rng(0)
    old = rng.standard_normal(learned._COUNT_4FEAT).astype(np.float32)
    prefix = str(tmp_path / "w")
    old.astype("<f4").tofile(prefix + ".bin")
    with open(prefix + ".json", "w") as f:
        json.dump({"count": int(old.shape[0])}, f)
    w = learned.PrecLearned.load_params(prefix)
    H = learned.HIDDEN
    assert w.shape[0] == sum(int(np.prod(s)) for _, s in learned.param_layout())
    enc = w[:learned.NODE_FEATS * H].reshape(learned.NODE_FEATS, H)
    assert np.array_equal(enc[:2], old[:2 * H].reshape(2, H))
    assert np.all(enc[2] == 0)
    assert np.array_equal(enc[3:], old[2 * H:4 * H].reshape(2, H))
    assert np.array_equal(w[learned.NODE_FEATS * H:], old[4 * H:])
