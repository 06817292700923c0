"""The autotuner's cache lookup (no GPU): nearest tuned n of the same element
type within a factor of 1.5, and LHL_TUNE=0 disables it."""
import json

import numpy as np


def test_lookup_nearest(tmp_path, monkeypatch, lhl_hip):
    from lhl_hip import autotune
    p = tmp_path / "tuned.json"
    p.write_text(json.dumps({
        "f64:2048": {"knobs": {"reduce_switch": 300}},
        "f64:4096": {"knobs": {"reduce_switch": 400, "reduce_nb": 8}},
        "f32:4096": {"knobs": {"reduce_blocked_min": 1 << 30}},
    }))
    monkeypatch.setenv("LHL_TUNE_CACHE", str(p))
    monkeypatch.setattr(autotune, "_cache", None)
    assert autotune.lookup(np.float64, 4096) == {"reduce_switch": 400, "reduce_nb": 8}
    assert autotune.lookup(np.float64, 3500) == {"reduce_switch": 400, "reduce_nb": 8}
    assert autotune.lookup(np.float64, 2500) == {"reduce_switch": 300}
    assert autotune.lookup(np.float64, 9000) == {}
    assert autotune.lookup(np.float32, 4000) == {"reduce_blocked_min": 1 << 30}
    assert autotune.lookup(np.float32, 2048) == {}
    monkeypatch.setenv("LHL_TUNE", "0")
    assert autotune.lookup(np.float64, 4096) == {}
