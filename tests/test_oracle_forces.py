"""Force on a geometry (orchestrator/src/force.rs:39-129 sampling, FP64 sums):
the oracle's restatement against closed forms on synthetic fields."""
import numpy as np
import pytest

import pyoracle
from ultraballpit_amd import scenes


def sphere_case(r_sphere=0.01, dx=1.0e-3, box=0.04, nodes=49):
    """Lattice filling a cube around a sphere (sites inside removed), no
    gravity, the sphere alone as the force geometry."""
    n_side = int(round(box / dx))
    pos = scenes.lattice(n_side, n_side, n_side, dx)
    c = np.full(3, box / 2)
    pos = pos[np.linalg.norm(pos - c, axis=1) > r_sphere]
    sc = scenes._particles(pos, dx)
    sc.update({"h": np.float32(1.3 * dx), "gravity": (0.0, 0.0, 0.0), "speed_of_sound": np.float32(10.0),
               "cfl_number": np.float32(0.25), "viscosity": 0.0, "domain_min": (0.0, 0.0, 0.0),
               "domain_max": (box, box, box)})
    like = scenes.box_sdf((0, 0, 0), (box, box, box), (nodes, nodes, nodes))
    fs = scenes.sphere_sdf(c, r_sphere, like)
    o = pyoracle.Oracle(sc, precision=64, flags=pyoracle.ENGINE)
    o.step(1e-9)  # builds the grid and the state fields
    o.set_force_sdf(fs)
    return o, pos, c


def test_uniform_pressure_gives_no_net_force():
    o, pos, c = sphere_case()
    o.set("pressure", np.full(o.n, 1000.0))
    o.set("density", np.full(o.n, 1000.0))
    for a in "xyz":
        o.set("v" + a, np.zeros(o.n))
    for mode in (0, 1):
        f = o.surface_forces(mu=0.0, mode=mode)
        assert f["samples"] > 1000
        # exact symmetry is broken only by FP32 node positions and distances
        assert np.max(np.abs(f["pressure_force"])) < 1e-6 * f["abs_sum"]
        assert np.all(f["viscous_force"] == 0.0)


def test_linear_pressure_gives_archimedes_buoyancy():
    # p = p0 - k z: the fluid pushes the body along +z with k V (V = 4/3 pi r^3)
    r = 0.01
    o, pos, c = sphere_case(r_sphere=r)
    k = 2.0e5
    z = o.get("z")
    o.set("pressure", 5000.0 - k * (z - c[2]))
    o.set("density", np.full(o.n, 1000.0))
    f = o.surface_forces(mu=0.0, mode=1)
    want = k * 4.0 / 3.0 * np.pi * r ** 3
    assert f["pressure_force"][2] == pytest.approx(want, rel=0.05)
    assert abs(f["pressure_force"][0]) < 1e-3 * want and abs(f["pressure_force"][1]) < 1e-3 * want
    # mode 0 (every node of the 2h band weighted cell^2) counts the band
    # thickness in cells over again: several times the surface integral
    f0 = o.surface_forces(mu=0.0, mode=0)
    assert f0["pressure_force"][2] > 2.0 * want


def test_moment_of_offset_force():
    # the same buoyancy about the origin: M = c x F
    o, pos, c = sphere_case()
    z = o.get("z")
    o.set("pressure", 5000.0 - 2.0e5 * (z - c[2]))
    o.set("density", np.full(o.n, 1000.0))
    f = o.surface_forces(mu=0.0, mode=1)
    want = np.cross(c, f["pressure_force"])
    assert np.allclose(f["net_moment"], want, rtol=1e-2, atol=1e-3 * np.linalg.norm(want))


def test_viscous_traction_of_a_shear_flow_is_linear_in_mu():
    o, pos, c = sphere_case()
    z = o.get("z")
    o.set("pressure", np.zeros(o.n))
    o.set("density", np.full(o.n, 1000.0))
    o.set("vx", 0.5 * (z - c[2]))
    f1 = o.surface_forces(mu=1.0e-3, mode=1)
    f2 = o.surface_forces(mu=2.0e-3, mode=1)
    assert np.linalg.norm(f1["viscous_force"]) > 0
    assert np.allclose(f2["viscous_force"], 2.0 * f1["viscous_force"], rtol=1e-12)
    assert np.all(f1["pressure_force"] == 0.0)
