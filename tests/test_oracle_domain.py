"""Pins the oracle's restatement of setup_domain
(backend/crates/orchestrator/src/domain.rs:35-402): the reference's own tests
(domain.rs:880-958), closed forms for the wall layers, a planar geometry
whose normal and layers are exact, a sphere (normals point out of the solid,
layers lie inside it), lattice sites past domain_max, and the f32::MAX rule
for sites outside the SDF grid."""
import numpy as np

import pyoracle as po

ALL = 0b111111


def flat_sdf(dims, cell, value=1.0, origin=(0.0, 0.0, 0.0)):
    return {"origin": np.asarray(origin, np.float32), "cell": np.float32(cell),
            "dims": np.asarray(dims, np.uint32), "data": np.full(int(np.prod(dims)), value, np.float32)}


def test_reference_setup_domain_basic():
    # domain.rs:880-920: 1 cm box, spacing 5 mm, all-positive 3^3 SDF
    F, B = po.setup_domain((0, 0, 0), (0.01, 0.01, 0.01), 0.005, 0, 293.15, ALL, flat_sdf((3, 3, 3), 0.005))
    assert len(F["x"]) == 8 and len(B["x"]) == 6 * 3 * 2 * 2
    assert np.all(F["fluid_type"] == 0)
    s = np.float32(0.005)
    assert np.all(F["mass"] == np.float32(1000.0) * (s * (s * s)))
    assert np.all(F["density"] == np.float32(1000.0)) and np.all(F["temperature"] == np.float32(293.15))
    # i outer, k inner: x changes slowest
    assert list(F["x"]) == [np.float32(0.0025)] * 4 + [np.float32(0.0075)] * 4
    assert list(F["z"][:2]) == [np.float32(0.0025), np.float32(0.0075)]


def test_reference_mixed_fluid_type():
    # domain.rs:922-958: taller domain, Mixed -> bottom half water, top half air
    F, B = po.setup_domain((0, 0, 0), (0.01, 0.02, 0.01), 0.005, 2, 293.15, ALL, flat_sdf((3, 5, 3), 0.005))
    water = F["fluid_type"] == 0
    assert water.sum() == 8 and (~water).sum() == 8
    assert np.all(F["y"][water] < np.float32(0.01)) and np.all(F["y"][~water] > np.float32(0.01))
    s = np.float32(0.005)
    vol = s * (s * s)
    assert np.all(F["mass"][~water] == np.float32(1.2) * vol) and np.all(F["density"][~water] == np.float32(1.2))
    assert np.all(B["mass"] == np.float32(1000.0) * vol)  # Mixed walls take the water mass


def test_wall_layers_closed_form():
    dmin, dmax, s = (0.0, 0.0, 0.0), (0.004, 0.003, 0.005), np.float32(0.001)
    walls = 0b100001  # x_min and z_max
    F, B = po.setup_domain(dmin, dmax, s, 1, 300.0, walls, flat_sdf((8, 8, 8), 0.001, origin=(-0.002,) * 3))
    n = {"x": 4, "y": 3, "z": 5}
    assert len(B["x"]) == 3 * n["y"] * n["z"] + 3 * n["x"] * n["y"]
    k = 3 * n["y"] * n["z"]
    half = lambda c: (np.arange(c, dtype=np.float32) + np.float32(0.5)) * s  # noqa: E731
    # x_min: layer, then j (y), then k (z)
    exp_x = np.repeat(-(np.arange(3, dtype=np.float32) + np.float32(0.5)) * s, n["y"] * n["z"])
    assert np.array_equal(B["x"][:k], exp_x)
    assert np.array_equal(B["y"][:k], np.tile(np.repeat(half(3), 5), 3))
    assert np.array_equal(B["z"][:k], np.tile(half(5), 9))
    assert np.all(B["nx"][:k] == 1.0) and np.all(B["ny"][:k] == 0.0) and np.all(B["nz"][:k] == 0.0)
    # z_max: layer, then i (x), then j (y), normal -z
    exp_z = np.repeat(np.float32(0.005) + (np.arange(3, dtype=np.float32) + np.float32(0.5)) * s, 12)
    assert np.array_equal(B["z"][k:], exp_z)
    assert np.array_equal(B["x"][k:], np.tile(np.repeat(half(4), 3), 3))
    assert np.all(B["nz"][k:] == -1.0)
    assert np.all(B["mass"] == np.float32(1.2) * (s * (s * s)))  # Air walls take the air mass
    assert np.all(F["fluid_type"] == 1)


def test_planar_geometry_layers_and_normals():
    # SDF d = z - 0.3 on a 0.25-spaced grid: exact in f32, trilinear exact,
    # gradient exactly (0, 0, 1)
    dims = (9, 9, 9)
    origin = (-0.5, -0.5, -0.5)
    z = origin[2] + 0.25 * np.arange(9)
    data = np.broadcast_to(z - 0.3, dims).astype(np.float32).ravel()
    sdf = {"origin": np.float32(origin), "cell": np.float32(0.25), "dims": np.uint32(dims), "data": data}
    s = np.float32(0.125)
    F, B = po.setup_domain((0, 0, 0), (1.0, 1.0, 1.0), s, 0, 293.15, 0, sdf)
    zs = (np.arange(8, dtype=np.float32) + np.float32(0.5)) * s
    d = zs - np.float32(0.3)
    surf = (d > 0) & (d < s / 2)   # one lattice plane: z = 0.3125
    assert surf.sum() == 1
    assert len(B["x"]) == 64 * 3
    assert np.all(B["nz"] == 1.0) and np.all(B["nx"] == 0.0) and np.all(B["ny"] == 0.0)
    zsite = zs[surf][0]
    assert np.array_equal(B["z"][:3], zsite - (np.arange(3, dtype=np.float32) + np.float32(0.5)) * s)
    # fluid: the sites with d >= 0
    assert len(F["x"]) == 64 * int((d >= 0).sum())
    assert np.all(F["z"] >= np.float32(0.3))


def test_sphere_geometry_normals_point_out_of_the_solid():
    c, r = np.array([0.5, 0.5, 0.5]), 0.2
    dims = (41, 41, 41)
    g = -0.5 + 0.05 * np.arange(41)
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    data = (np.sqrt((X - c[0]) ** 2 + (Y - c[1]) ** 2 + (Z - c[2]) ** 2) - r).astype(np.float32).ravel()
    sdf = {"origin": np.float32((-0.5,) * 3), "cell": np.float32(0.05), "dims": np.uint32(dims), "data": data}
    F, B = po.setup_domain((0, 0, 0), (1, 1, 1), 0.02, 0, 293.15, 0, sdf)
    n = np.stack([B["nx"], B["ny"], B["nz"]], axis=1).astype(np.float64)
    p = np.stack([B["x"], B["y"], B["z"]], axis=1).astype(np.float64)
    assert len(p) > 0 and len(p) % 3 == 0
    assert np.allclose(np.linalg.norm(n, axis=1), 1.0, atol=1e-6)
    site = p[0::3] + 0.5 * 0.02 * n[0::3]
    assert np.all(np.sum(n[0::3] * (site - c), axis=1) > 0)  # outward normal
    assert np.all(np.linalg.norm(p - c, axis=1) < r + 0.02)   # layers go into the solid
    # fluid outside the sphere only (trilinear distance within a cell of the true one)
    q = np.stack([F["x"], F["y"], F["z"]], axis=1).astype(np.float64)
    assert np.all(np.linalg.norm(q - c, axis=1) > r - 0.05)


def test_sites_past_domain_max_and_outside_the_sdf_grid():
    # extents not a multiple of the spacing: ceil gives one site past dmax,
    # which the fluid skips; an SDF grid smaller than the domain reads
    # f32::MAX there (>= 0: fluid kept; not < spacing/2: no geometry layers)
    s = np.float32(0.001)
    F, B = po.setup_domain((0, 0, 0), (0.0101, 0.002, 0.002), s, 0, 293.15, 0b000011,
                           flat_sdf((4, 4, 4), 0.001))
    assert len(F["x"]) == 10 * 2 * 2
    assert F["x"].max() < np.float32(0.0101)
    assert len(B["x"]) == 2 * 3 * 2 * 2
