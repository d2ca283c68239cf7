"""CPU checks of the training driver's host logic (fsx_train.py): the seeded
draws of the training cases and the weight files."""
import json

import numpy as np
import pytest


def test_draws_are_seeded_and_in_range(fsx_lib):
    import fsx_train
    a = fsx_train.draw_cases(2024, 6)
    assert a == fsx_train.draw_cases(2024, 6)
    assert a != fsx_train.draw_cases(2025, 6)
    for lid, re in a:
        assert fsx_train.LID_RANGE[0] <= lid <= fsx_train.LID_RANGE[1]
        assert fsx_train.RE_RANGE[0] <= re <= fsx_train.RE_RANGE[1]
    # the timed case (lid 1, Re 1000) is not among the draws
    assert all(abs(lid - 1.0) > 1e-6 or abs(re - 1000.0) > 1e-6 for lid, re in a)


def test_weight_file_round_trip_and_digest(fsx_lib, tmp_path):
    import fsx_train
    w = np.random.default_rng(0).normal(size=1000).astype(np.float32)
    p = str(tmp_path / "w.npy")
    meta = fsx_train.save_weights(p, w, {"seed": 3})
    w2, meta2 = fsx_train.load_weights(p)
    assert np.array_equal(w, w2) and meta2["sha256"] == meta["sha256"] and meta2["seed"] == 3
    with open(p + ".json") as f:
        m = json.load(f)
    m["sha256"] = "0" * 64
    with open(p + ".json", "w") as f:
        json.dump(m, f)
    with pytest.raises(ValueError):
        fsx_train.load_weights(p)
