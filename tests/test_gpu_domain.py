"""Domain set-up on the device against the oracle's restatement of
setup_domain (backend/crates/orchestrator/src/domain.rs:35-402), bit for
bit, and the device's analytic SDF node grids against the host builders."""
import sys
import time

import numpy as np
import pytest

import pyoracle as po
import ultraballpit_amd as ubp
from ultraballpit_amd import domain, scenes

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    if not ubp.gpu_available():
        pytest.fail("no gfx950 device / engine library not loadable")


def flat_sdf(dims, cell, value=1.0, origin=(0.0, 0.0, 0.0)):
    return {"origin": np.asarray(origin, np.float32), "cell": np.float32(cell),
            "dims": np.asarray(dims, np.uint32), "data": np.full(int(np.prod(dims)), value, np.float32)}


def config(dmin, dmax, spacing, fluid="Water", temperature=293.15, bc=None):
    return {"domain": {"min": list(dmin), "max": list(dmax)}, "particle_spacing": spacing,
            "fluid_type": fluid, "initial_temperature": temperature, "boundary_conditions": bc or {}}


def oracle_of(cfg, sdf):
    return po.setup_domain(cfg["domain"]["min"], cfg["domain"]["max"], cfg["particle_spacing"],
                           domain.FLUID_TYPES[cfg["fluid_type"]], cfg["initial_temperature"],
                           domain._wall_mask(cfg["boundary_conditions"]), sdf)


def same_bits(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    if a.dtype == np.float32:
        return a.shape == b.shape and np.array_equal(a.view(np.uint32), b.view(np.uint32))
    return np.array_equal(a, b)


def check(cfg, sdf, name):
    say("case: setup_domain", name)
    t0 = time.perf_counter()
    F, B = domain.setup_domain(cfg, sdf)
    t1 = time.perf_counter()
    OF, OB = oracle_of(cfg, sdf)
    t2 = time.perf_counter()
    say(f"   {len(F['x'])} fluid, {len(B['x'])} boundary; device {1e3 * (t1 - t0):.1f} ms, "
        f"oracle {1e3 * (t2 - t1):.1f} ms")
    assert len(F["x"]) == len(OF["x"]) and len(B["x"]) == len(OB["x"])
    for k in ("x", "y", "z", "mass", "density", "temperature", "fluid_type"):
        assert same_bits(F[k], OF[k]), f"{name}: fluid {k} differs"
    for k in ("vx", "vy", "vz", "ax", "ay", "az", "pressure"):
        assert not F[k].any()
    for k in po.BOUNDARY_FIELDS:
        assert same_bits(B[k], OB[k]), f"{name}: boundary {k} differs"
    return F, B


def test_reference_domain_test_configs():
    F, B = check(config((0, 0, 0), (0.01, 0.01, 0.01), 0.005), flat_sdf((3, 3, 3), 0.005), "domain.rs:880")
    assert len(F["x"]) == 8 and len(B["x"]) == 72
    F, _ = check(config((0, 0, 0), (0.01, 0.02, 0.01), 0.005, "Mixed"), flat_sdf((3, 5, 3), 0.005), "domain.rs:923")
    assert np.bincount(F["fluid_type"]).tolist() == [8, 8]


def test_water_box_1cm_config():
    # configs/water-box-1cm.json: 1 cm closed box, spacing 1 mm, the box as geometry
    sdf = scenes.box_sdf((0, 0, 0), (0.01, 0.01, 0.01), (24, 24, 24))
    check(config((0, 0, 0), (0.01, 0.01, 0.01), 0.001), sdf, "water-box-1cm")


def test_channel_around_a_sphere_mixed_open_ends_ragged_extent():
    sdf = scenes.sphere_in_box_sdf((0, 0, 0), (0.4, 0.2, 0.2), (0.1, 0.1, 0.1), 0.03, (96, 48, 48))
    bc = {"x_min": "Outflow", "x_max": "Outflow"}
    F, B = check(config((0, 0, 0), (0.4, 0.2, 0.2017), 0.004, "Mixed", 300.0, bc), sdf, "sphere channel")
    assert (F["fluid_type"] == 1).any() and (F["fluid_type"] == 0).any()
    nwall = 3 * (100 * 51) * 2 + 3 * (100 * 50) * 2
    assert len(B["x"]) > nwall  # plus the layers under the sphere surface


def test_sdf_grid_smaller_than_the_domain():
    check(config((0, 0, 0), (0.0101, 0.002, 0.002), 0.001, "Air", bc={"y_min": "Outflow"}),
          flat_sdf((4, 4, 4), 0.001), "small SDF grid")


def test_flow_past_sphere_scale():
    # the 4M-site channel of the flow-past-sphere workload at its full lattice
    sdf = scenes.sphere_in_box_sdf((0, 0, 0), (0.4, 0.2, 0.2), (0.1, 0.1, 0.1), 0.03, (256, 128, 128))
    F, B = check(config((0, 0, 0), (0.4, 0.2, 0.2), 0.4 / 256), sdf, "flow past sphere, 256x128x128 sites")
    assert len(F["x"]) > 4_000_000


@pytest.mark.parametrize("dims", [(24, 24, 24), (256, 64, 160)])
def test_device_box_sdf_equals_host(dims):
    say("case: device box SDF", dims)
    a = domain.box_sdf((0, 0, 0), (1.0, 0.25, 0.6), dims)
    b = scenes.box_sdf((0, 0, 0), (1.0, 0.25, 0.6), dims)
    for k in ("origin", "cell", "dims"):
        assert np.array_equal(np.asarray(a[k]), np.asarray(b[k]))
    assert same_bits(a["data"], b["data"])


def test_device_sphere_sdfs_equal_host():
    say("case: device sphere-in-box and sphere SDFs, 256x128x128")
    a = domain.sphere_in_box_sdf((0, 0, 0), (0.4, 0.2, 0.2), (0.1, 0.1, 0.1), 0.03, (256, 128, 128))
    b = scenes.sphere_in_box_sdf((0, 0, 0), (0.4, 0.2, 0.2), (0.1, 0.1, 0.1), 0.03, (256, 128, 128))
    assert same_bits(a["data"], b["data"])
    c = domain.sphere_sdf((0.1, 0.1, 0.1), 0.03, b)
    d = scenes.sphere_sdf((0.1, 0.1, 0.1), 0.03, b)
    assert same_bits(c["data"], d["data"])


def test_invalid_domain_fails_cleanly():
    with pytest.raises(ubp.EngineError):
        domain.setup_domain(config((0, 0, 0), (1, 1, 1), -0.1), flat_sdf((3, 3, 3), 0.5))
    with pytest.raises(ubp.EngineError):
        domain.setup_domain(config((0, 0, 0), (1, 1, 1), 0.1, "Oil"), flat_sdf((3, 3, 3), 0.5))
