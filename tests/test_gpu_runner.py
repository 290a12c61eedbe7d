"""The orchestrator's optimistic stepping loop (backend/crates/orchestrator/
src/runner.rs:157-277): the device loop (``HipKernel.run_optimistic``) against
the host-dt restatement over the trait calls, on the engine and on the
oracle, bit for bit, with every try accepted, every try rejected, a mix, and
the reference's own limits."""
import sys

import numpy as np
import pytest

import ultraballpit_amd as ubp
from ultraballpit_amd import runner, scenes
from parity_util import assert_bitwise, first_mismatch, oracle

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    if not ubp.gpu_available():
        pytest.fail("no gfx950 device / engine library not loadable")


def scene():
    return scenes.dam_break_small(lat=(12, 12, 24))


@pytest.mark.parametrize("limit", [0.0, 0.03, 0.385, 1e9])
def test_device_loop_equals_host_loop_and_oracle(limit):
    sc = scene()
    say(f"case: optimistic loop, dam_break_small 3456, 30 steps, density limit {limit}")
    e = ubp.HipKernel.from_scene(sc)
    got = e.run_optimistic(max_steps=30, alpha=1.5, density_limit=limit, cooldown=4)
    say("   device", got)
    h = ubp.HipKernel.from_scene(sc)
    want = runner.run_simulation_loop(h, h.dt, lambda: h.error_metrics().max_density_variation,
                                      max_timesteps=30, alpha=1.5, limit=limit, cooldown=4)
    assert got == want
    assert first_mismatch(e.particles(), h.particles()) is None
    o = oracle(sc)
    ow = runner.run_simulation_loop(o, o.timestep, lambda: o.error_metrics()[0], max_timesteps=30,
                                    alpha=1.5, limit=limit, cooldown=4)
    assert ow == got
    assert_bitwise(e.particles(), o, f"optimistic loop, limit {limit}")
    if limit == 0.0:
        assert got["accepted"] == 0 and got["rejected"] > 0
    if limit == 1e9:
        assert got["rejected"] == 0 and got["accepted"] == 30
    if limit == 0.385:  # free-surface particles keep the variation near 0.36-0.39
        assert got["accepted"] > 0 and got["rejected"] > 0


def test_device_loop_reference_constants_and_max_time():
    sc = scene()
    say("case: optimistic loop with the reference's constants, stop at 2 ms of simulated time")
    e = ubp.HipKernel.from_scene(sc)
    got = e.run_optimistic(max_time=2e-3)
    o = oracle(sc)
    want = runner.run_simulation_loop(o, o.timestep, lambda: o.error_metrics()[0], max_time=2e-3)
    say("   device", got, "oracle", want)
    assert got == want and got["sim_time"] >= 2e-3
    assert_bitwise(e.particles(), o, "optimistic loop, max_time 2 ms")


def test_device_loop_rejects_bad_arguments():
    e = ubp.HipKernel.from_scene(scene())
    with pytest.raises(ubp.EngineError):
        e.run_optimistic()  # neither max_steps nor max_time
    with pytest.raises(ubp.EngineError):
        e.run_optimistic(max_steps=3, alpha=0.0)
