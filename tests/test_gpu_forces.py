"""FP64 force extraction on the device (orchestrator/src/force.rs:39-129
sampling) against the oracle's restatement on the same stepped state.  The
per-pair FP32 values are identical; only the FP64 summation order differs
(cell order on the device, particle index order in the oracle)."""
import sys

import numpy as np
import pytest

import ultraballpit_amd as ubp
from ultraballpit_amd import scenes
from parity_util import assert_bitwise, oracle

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    if not ubp.gpu_available():
        pytest.fail("no gfx950 device / engine library not loadable")


@pytest.mark.parametrize("mode", [0, 1])
def test_surface_forces_match_oracle(mode):
    sc = scenes.flow_past_sphere(lat=(32, 16, 16), sdf_nodes=(64, 32, 32))
    fs = scenes.sphere_sdf(sc["sphere"]["centre"], sc["sphere"]["radius"], sc["sdf"])
    say(f"case: flow past sphere {len(sc['x'])} particles, 6 steps, surface forces mode {mode}")
    e = ubp.HipKernel.from_scene(sc)
    o = oracle(sc)
    e.step_n(6, use_graph=False)
    o.step_auto(6)
    assert_bitwise(e.particles(), o, "flow past sphere after 6 steps")
    e.set_force_sdf(fs)
    o.set_force_sdf(fs)
    mu = float(np.float32(2.0e-3))  # the C ABI takes mu as a float
    got = e.surface_forces(mu=mu, mode=mode)
    want = o.surface_forces(mu=mu, mode=mode)
    say("  engine", got["net_force"], "oracle", want["net_force"], "samples", got["samples"])
    assert got["samples"] == want["samples"] > 100
    tol = 1e-9 * want["abs_sum"]
    for k in ("pressure_force", "viscous_force", "net_force"):
        assert np.all(np.abs(got[k] - want[k]) <= tol), (k, got[k], want[k])
    assert np.all(np.abs(got["net_moment"] - want["net_moment"]) <= tol * 0.4)
    assert np.linalg.norm(got["viscous_force"]) > 0


def test_surface_forces_need_geometry_and_state():
    sc = scenes.flow_past_sphere(lat=(16, 8, 8), sdf_nodes=(32, 16, 16))
    e = ubp.HipKernel.from_scene(sc)
    with pytest.raises(ubp.EngineError):
        e.surface_forces()
    e.set_force_sdf(scenes.sphere_sdf(sc["sphere"]["centre"], sc["sphere"]["radius"], sc["sdf"]))
    with pytest.raises(ubp.EngineError):
        e.surface_forces()  # no step yet
    e.step_n(1, use_graph=False)
    f = e.surface_forces(mode=1)
    assert np.isfinite(f["net_force"]).all() and np.all(f["viscous_force"] == 0.0)


@pytest.mark.timeout(900)
def test_config3_flow_past_sphere_full_size():
    # the flow-past-sphere workload at its full size: 256x128x128 lattice
    # minus the sphere (4.16M particles), sphere SDF 256x128x128
    sc = scenes.flow_past_sphere()
    fs = scenes.sphere_sdf(sc["sphere"]["centre"], sc["sphere"]["radius"], sc["sdf"])
    say(f"case: flow past sphere, full size ({len(sc['x'])} particles), 5 device-dt steps, FP64 forces")
    e = ubp.HipKernel.from_scene(sc)
    e.step_n(5, use_graph=True)
    e.set_force_sdf(fs)
    mu = float(np.float32(3.05e-2))  # the C ABI takes mu as a float
    got = e.surface_forces(mu=mu, mode=0)
    state = e.particles()
    e.close()
    say("  engine done; oracle (OpenMP) ...")
    o = oracle(sc)
    o.step_auto(5)
    assert_bitwise(state, o, "flow past sphere, full size, after 5 steps")
    o.set_force_sdf(fs)
    want = o.surface_forces(mu=mu, mode=0)
    say("  engine", got["net_force"], "oracle", want["net_force"], "samples", got["samples"])
    assert got["samples"] == want["samples"] > 5000
    tol = 1e-9 * want["abs_sum"]
    for k in ("pressure_force", "viscous_force", "net_force"):
        assert np.all(np.abs(got[k] - want[k]) <= tol), (k, got[k], want[k])
