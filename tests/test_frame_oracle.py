"""Pins the viewer-frame oracle (backend/crates/server/src/ws.rs:232-294): layout round
trip, stride/count rule, and the saturating density-ratio cast."""
import numpy as np

import frame_oracle as fo


def particles(n, seed=0):
    rng = np.random.default_rng(seed)
    return {"x": rng.random(n, np.float32), "y": rng.random(n, np.float32), "z": rng.random(n, np.float32),
            "temperature": np.full(n, 293.15, np.float32), "fluid_type": (rng.random(n) < 0.3).astype(np.uint8),
            "density": (1000.0 + rng.normal(0, 20, n)).astype(np.float32)}


def test_layout_round_trip_and_subsampling():
    p = particles(1000)
    buf = fo.build_frame(p, 50, 7, 0.125)
    tag, fn, count, t, rows = fo.parse_frame(buf)
    assert (tag, fn, count, t) == (2, 7, 50, 0.125)
    assert len(buf) == 21 + 20 * 50
    for k, row in enumerate(rows):
        i = 20 * k  # stride = 1000 // 50
        assert row[:4] == (p["x"][i], p["y"][i], p["z"][i], np.float32(293.15))
        assert row[4] == p["fluid_type"][i] and row[6] == 0


def test_count_rule():
    p = particles(1003)
    assert fo.parse_frame(fo.build_frame(p, 50, 0, 0.0))[2] == 50    # stride 20, 1003 // 20 = 50
    assert fo.parse_frame(fo.build_frame(p, 5000, 0, 0.0))[2] == 1003  # stride 1
    assert fo.build_frame(p, 0, 3, 0.0) == bytes([2, 3, 0, 0, 0, 0, 0, 0, 0]) + bytes(12)  # empty frame


def test_density_ratio_cast():
    assert fo.density_ratio(1000.0, 0) == 1000
    assert fo.density_ratio(1000.9, 0) == 1000        # truncation
    assert fo.density_ratio(-5.0, 0) == 0             # clamp below
    assert fo.density_ratio(1.0e9, 0) == 65535        # clamp above
    assert fo.density_ratio(float("nan"), 0) == 0     # NaN -> 0
    assert fo.density_ratio(1.204, 1) == 1000
