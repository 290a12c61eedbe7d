"""Viewer frame built on the device (backend/crates/server/src/ws.rs:232-294) against the
oracle's restatement on the same state, byte for byte."""
import sys

import numpy as np
import pytest

import frame_oracle as fo
import ultraballpit_amd as ubp
from ultraballpit_amd import scenes
from parity_util import oracle

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    if not ubp.gpu_available():
        pytest.fail("no gfx950 device / engine library not loadable")


@pytest.mark.parametrize("subsample", [0, 1, 97, 400, 100000])
def test_device_frame_equals_reference_frame(subsample):
    sc = scenes.dam_break_small(lat=(12, 12, 24))
    say("case: frame, dam_break_small, subsample", subsample)
    e = ubp.HipKernel.from_scene(sc)
    before = e.build_frame(subsample, 0, 0.0)  # no step yet: the input arrays
    assert before == fo.build_frame(e.particles(), subsample, 0, 0.0)
    e.step_n(7, use_graph=False)
    got = e.build_frame(subsample, 7, 1.5e-3)
    assert got == fo.build_frame(e.particles(), subsample, 7, 1.5e-3)
    o = oracle(sc)
    o.step_auto(7)
    p = {k: o.get(k).astype(np.float32) for k in ("x", "y", "z", "temperature", "density")}
    p["fluid_type"] = np.zeros(len(p["x"]), np.uint8)
    assert got == fo.build_frame(p, subsample, 7, 1.5e-3)
