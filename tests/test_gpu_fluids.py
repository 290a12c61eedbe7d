"""Air and water/air scenes on the engine against the oracle, bit for bit:
the Tait EOS with gamma 1.4 and rho0 1.204 for air (eos.rs:7-19), the
per-type boundary density (sph.rs:376-408), Mixed set-up (domain.rs:
126-137).  Scenes come from the device domain set-up."""
import sys

import numpy as np
import pytest

import ultraballpit_amd as ubp
from ultraballpit_amd import domain, scenes
from parity_util import assert_bitwise, engine, oracle

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    if not ubp.gpu_available():
        pytest.fail("no gfx950 device / engine library not loadable")


def scene(fluid, box, spacing, c0, gravity, walls_as_particles=False, sdf_nodes=(24, 24, 24)):
    cfg = {"domain": {"min": [0.0, 0.0, 0.0], "max": list(box)}, "particle_spacing": spacing,
           "fluid_type": fluid, "initial_temperature": 293.15}
    sdf = scenes.box_sdf((0, 0, 0), box, sdf_nodes)
    F, B = domain.setup_domain(cfg, sdf)
    sc = {k: F[k] for k in ("x", "y", "z", "vx", "vy", "vz", "mass", "density", "temperature", "fluid_type")}
    sc.update({"h": np.float32(1.3 * spacing), "gravity": gravity, "speed_of_sound": np.float32(c0),
               "cfl_number": np.float32(0.25), "viscosity": 0.0, "domain_min": (0.0, 0.0, 0.0),
               "domain_max": tuple(box), "alpha": np.float32(0.02), "beta": np.float32(0.0)})
    if walls_as_particles:
        sc["boundary"] = B
    else:
        sc["sdf"] = sdf
    return sc


def run(sc, steps, name):
    say(f"case: {name}, {len(sc['x'])} particles, types {np.bincount(sc['fluid_type'], minlength=2).tolist()}, "
        f"{steps} device-dt steps")
    e = engine(sc)
    o = oracle(sc)
    e.step_n(steps, use_graph=False)
    o.step_auto(steps)
    p = e.particles()
    assert all(np.isfinite(p[k]).all() for k in ("x", "y", "z", "vx", "vy", "vz", "density"))
    assert_bitwise(p, o, f"{name} after {steps} steps")
    return p


def test_air_box_bitwise():
    p = run(scene("Air", (0.012, 0.012, 0.012), 0.001, 10.0, (0.0, 0.0, -9.81)), 40, "air box")
    assert np.all(p["density"] < 5.0)


def test_mixed_water_air_bitwise():
    p = run(scene("Mixed", (0.012, 0.016, 0.012), 0.001, 10.0, (0.0, -9.81, 0.0)), 40, "mixed box")
    assert (p["density"] > 500).any() and (p["density"] < 5).any()


def test_mixed_with_reference_wall_particles_bitwise():
    run(scene("Mixed", (0.008, 0.012, 0.008), 0.001, 10.0, (0.0, -9.81, 0.0), walls_as_particles=True), 30,
        "mixed box, wall + geometry boundary particles")
