"""Pin the oracle's neighbour search, binning and GridSDF against brute force
and closed forms, plus the reference's own unit tests
(backend/crates/kernel/src/neighbor.rs:185-269,
backend/crates/orchestrator/src/geometry.rs:184-250)."""
import numpy as np
import pytest

import pyoracle as po
from ultraballpit_amd import scenes


def brute_neighbors(x, y, z, i, radius):
    """O(N) scan in float32 with the reference association
    ((dx*dx + dy*dy) + dz*dz) <= r*r, j != i."""
    f = np.float32
    dx = f(x[i]) - x.astype(f)
    dy = f(y[i]) - y.astype(f)
    dz = f(z[i]) - z.astype(f)
    d2 = (dx * dx + dy * dy) + dz * dz
    ok = d2 <= f(radius) * f(radius)
    ok[i] = False
    return set(np.nonzero(ok)[0].tolist())


# ---- reference unit tests (neighbor.rs:189-268) ----

def test_single_particle_no_neighbors():
    assert len(po.neighbors([0.5], [0.5], [0.5], 0.2, [0] * 3, [1] * 3, 0.2, 0)) == 0


def test_two_close_particles():
    x, y, z = [0.5, 0.51], [0.5, 0.5], [0.5, 0.5]
    assert list(po.neighbors(x, y, z, 0.2, [0] * 3, [1] * 3, 0.2, 0)) == [1]
    assert list(po.neighbors(x, y, z, 0.2, [0] * 3, [1] * 3, 0.2, 1)) == [0]


def test_two_far_particles():
    assert len(po.neighbors([0.1, 0.9], [0.1, 0.9], [0.1, 0.9], 0.2, [0] * 3, [1] * 3, 0.2, 0)) == 0


def test_across_cell_boundary():
    assert list(po.neighbors([0.19, 0.21], [0.5] * 2, [0.5] * 2, 0.2, [0] * 3, [1] * 3, 0.2, 0)) == [1]


def test_cluster():
    x = 0.5 + np.arange(10) * 0.01
    assert len(po.neighbors(x, [0.5] * 10, [0.5] * 10, 0.2, [0] * 3, [1] * 3, 0.2, 0)) == 9


# ---- brute force: set and order ----

@pytest.mark.parametrize("seed", [1, 2, 3])
def test_neighbor_set_equals_brute_force_and_order(seed):
    rng = np.random.default_rng(seed)
    n = 600
    pts = rng.random((n, 3)).astype(np.float32)
    # duplicates and points outside the domain (clamped into edge cells)
    pts[5] = pts[6]
    pts[7] = [1.05, 0.5, -0.02]
    cell = 0.1
    dmin, dmax = [0, 0, 0], [1, 1, 1]
    x, y, z = pts[:, 0], pts[:, 1], pts[:, 2]
    dims = np.ceil(np.float32(1.0) / np.float32(cell)).astype(int)

    def cellof(p):
        c = np.floor((p.astype(np.float32) - np.float32(0)) / np.float32(cell))
        return np.clip(c, 0, dims - 1).astype(int)

    cx, cy, cz = cellof(x), cellof(y), cellof(z)
    for i in range(0, n, 7):
        got = po.neighbors(x, y, z, cell, dmin, dmax, cell, i)
        ref = brute_neighbors(x, y, z, i, cell)
        # the grid search only looks at the 27 adjacent (clamped) cells, like
        # the reference; every brute-force neighbour lies there
        assert set(got.tolist()) == ref
        # order: cells in (dz, dy, dx) order, then ascending index
        key = [((cz[j] - cz[i]) * 9 + (cy[j] - cy[i]) * 3 + (cx[j] - cx[i]), j) for j in got]
        assert key == sorted(key)


# ---- Morton keys and the stable permutation ----

def morton_ref(cx, cy, cz, bits):
    code, pos = 0, 0
    for lvl in range(max(bits)):
        for a, c in enumerate((cx, cy, cz)):
            if lvl < bits[a]:
                code |= ((c >> lvl) & 1) << pos
                pos += 1
    return code


def test_morton_closed_form():
    bits = [3, 3, 3]
    assert po.morton(1, 0, 0, bits) == 1
    assert po.morton(0, 1, 0, bits) == 2
    assert po.morton(0, 0, 1, bits) == 4
    assert po.morton(7, 7, 7, bits) == 511
    assert po.morton(2, 0, 0, bits) == 8
    # unequal widths (config-2 grid 125 x 32 x 75 -> 7, 5, 7 bits): dense,
    # bijective onto [0, 2^19) over the padded box
    bits = [7, 5, 7]
    rng = np.random.default_rng(0)
    seen = set()
    for _ in range(3000):
        c = (int(rng.integers(128)), int(rng.integers(32)), int(rng.integers(128)))
        m = po.morton(*c, bits)
        assert m == morton_ref(*c, bits) and m < (1 << 19)
        seen.add((m, c))
    assert len({m for m, _ in seen}) == len({c for _, c in seen})


def test_bins_perm_is_stable_argsort():
    sc = scenes.dam_break_small(lat=(8, 8, 16))
    o = po.Oracle(sc)
    o.step_auto(3)
    keys, perm, counts = o.bins(morton=True)
    assert np.array_equal(perm, np.argsort(keys, kind="stable"))
    # keys from positions: independent cell + Morton computation
    dims = o.grid_dims()
    bits = [int(d - 1).bit_length() for d in dims]
    cell = np.float32(2.0) * np.float32(sc["h"])
    p = {k: o.get(k).astype(np.float32) for k in ("x", "y", "z")}
    c = [np.clip(np.floor((p[a] - np.float32(0)) / cell), 0, dims[i] - 1).astype(int)
         for i, a in enumerate("xyz")]
    want = [morton_ref(int(c[0][i]), int(c[1][i]), int(c[2][i]), bits) for i in range(len(keys))]
    assert np.array_equal(keys, np.array(want, np.uint32))
    # neighbour counts from the last force pass == brute-force counts
    for i in range(0, len(keys), 37):
        assert counts[i] == len(brute_neighbors(p["x"], p["y"], p["z"], i, np.float32(2) * np.float32(sc["h"])))


# ---- GridSDF (geometry.rs:99-175) ----

def test_grid_sdf_reference_sample():
    # geometry.rs:188-221
    d = np.array([-1, -1, -1, -1, 0, 0, 0, 0, 1, 1, 1, 1], np.float32)
    sdf = {"origin": [0, 0, 0], "cell": 1.0, "dims": [3, 2, 2], "data": d}
    # at the upper-j/k nodes the reference returns "outside"; test interior
    v, ok = po.sdf_sample(sdf, [0.5, 0.0, 0.0])
    assert ok and abs(v + 0.5) < 0.1
    v, ok = po.sdf_sample(sdf, [1.5, 0.0, 0.0])
    assert ok and abs(v - 0.5) < 0.1
    v, ok = po.sdf_sample(sdf, [0.0, 0.0, 0.0])
    assert ok and abs(v + 1.0) < 0.1


def test_grid_sdf_reference_gradient():
    # geometry.rs:223-249
    d = np.repeat(np.array([-1.0, 0.0, 1.0], np.float32), 9)
    sdf = {"origin": [0, 0, 0], "cell": 1.0, "dims": [3, 3, 3], "data": d}
    g, ok = po.sdf_gradient(sdf, [1.0, 1.0, 1.0])
    assert ok and abs(g[0] - 1) < 0.1 and abs(g[1]) < 0.1 and abs(g[2]) < 0.1


def test_trilinear_exact_on_affine_field():
    # trilinear interpolation reproduces an affine field exactly (up to FP32)
    dims = (9, 7, 8)
    o = np.array([-0.1, 0.2, 0.05])
    c = 0.05
    a = np.array([0.3, -0.7, 0.55])
    I, J, K = np.meshgrid(*[np.arange(d) for d in dims], indexing="ij")
    P = np.stack([o[0] + c * I, o[1] + c * J, o[2] + c * K], -1)
    f = (P @ a + 0.01).astype(np.float32).ravel()
    sdf = {"origin": o, "cell": c, "dims": dims, "data": f}
    rng = np.random.default_rng(4)
    for _ in range(50):
        p = o + c * (rng.random(3) * (np.array(dims) - 1.01))
        v, ok = po.sdf_sample(sdf, p)
        assert ok and abs(v - (np.float32(p) @ a + 0.01)) < 2e-6
        g, ok2 = po.sdf_gradient(sdf, p)
        if ok2:
            assert np.allclose(g, a / np.linalg.norm(a), atol=1e-4)


def test_outside_grid():
    sdf = scenes.box_sdf((0, 0, 0), (1, 1, 1), (16, 16, 16))
    assert not po.sdf_sample(sdf, [-5, 0.5, 0.5])[1]
    assert not po.sdf_sample(sdf, [0.5, 0.5, 9])[1]


def test_box_sdf_nodes_are_exact_distances():
    sdf = scenes.box_sdf((0, 0, 0), (1.0, 0.25, 0.6), (64, 24, 40))
    dims = sdf["dims"].astype(int)
    o = sdf["origin"].astype(np.float64)
    c = float(sdf["cell"])
    d = sdf["data"].reshape(dims)
    for (i, j, k) in [(0, 0, 0), (10, 5, 7), (32, 12, 20), (63, 23, 39)]:
        p = o + c * np.array([i, j, k])
        inside = np.all(p >= 0) and np.all(p <= [1.0, 0.25, 0.6])
        q = np.maximum(-p, p - np.array([1.0, 0.25, 0.6]))
        want = -min(q.max(), 0.0) - np.linalg.norm(np.maximum(q, 0))
        assert abs(d[i, j, k] - want) < 1e-6
        assert (d[i, j, k] > 0) == bool(inside and q.max() < 0)
