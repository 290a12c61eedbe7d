"""The reference's dam-break benchmark (reference-tests/src/benchmarks.rs:51-220,
`#[ignore]` there) on its own set-up: configs/dam-break-2d.json (domain
0.05 x 0.10 x 0.004 m, spacing 2 mm, c0 = 20, CFL 0.4, walls at x_min,
y_min, z_min, z_max), boundary particles from the domain set-up, a water
column of width a = domain width / 4 and height 2a placed by hand, the
reference's fixed alpha = 1, beta = 2, host CFL dt, and a snapshot every
10 ms.  The front (largest x) is compared with Martin & Moyce (1952),
tests/golden/martin_moyce_1952.json (the reference's data file), through the
reference's interpolation (reference-tests/src/analytical.rs:211-234).

At this stock resolution the front lags the experiment by up to 12.1%, so
the reference's +-10% bar is not met by its own method either (the oracle
restates that method); the CPU test pins that outcome at 15%.  The GPU test
checks that the engine steps the same set-up bit for bit like the oracle."""
import json
import math
import os

import numpy as np
import pytest

import pyoracle as po

HERE = os.path.dirname(os.path.abspath(__file__))
DMIN, DMAX, SPACING = (0.0, 0.0, 0.0), (0.05, 0.10, 0.004), np.float32(0.002)
# x_min, y_min, z_min, z_max walls; x_max and y_max outflow (dam-break-2d.json)
WALLS = 0b110101


def front_reference():
    d = json.load(open(os.path.join(HERE, "golden", "martin_moyce_1952.json")))["data"]
    T, Z = d["dimensionless_time"], d["dimensionless_front_position"]

    def ref(t):  # analytical.rs:211-234
        if t <= T[0]:
            return Z[0]
        if t >= T[-1]:
            return Z[-1] + (Z[-1] - Z[-2]) / (T[-1] - T[-2]) * (t - T[-1])
        return float(np.interp(t, T, Z))
    return ref


def scene():
    """benchmarks.rs:59-123: boundary particles from the domain set-up (no
    obstacle: a positive SDF everywhere), the fluid column placed by hand."""
    null_sdf = {"origin": np.float32([-0.01, -0.01, -0.01]), "cell": np.float32(0.5 * SPACING),
                "dims": np.uint32([72, 122, 26]), "data": np.full(72 * 122 * 26, 1.0, np.float32)}
    _, B = po.setup_domain(DMIN, DMAX, SPACING, 0, 293.15, WALLS, null_sdf)
    f32 = np.float32
    a = f32(DMAX[0] - DMIN[0]) / f32(4.0)
    hgt = f32(2.0) * a
    m = f32(1000.0) * (SPACING * SPACING * SPACING)
    nx, ny, nz = int(math.ceil(a / SPACING)), int(math.ceil(hgt / SPACING)), int(math.ceil(f32(DMAX[2]) / SPACING))
    pts = []
    for i in range(nx):
        for j in range(ny):
            for k in range(nz):
                x = f32(DMIN[0]) + (f32(i) + f32(0.5)) * SPACING
                y = f32(DMIN[1]) + (f32(j) + f32(0.5)) * SPACING
                z = f32(DMIN[2]) + (f32(k) + f32(0.5)) * SPACING
                if x <= f32(DMIN[0]) + a and y <= f32(DMIN[1]) + hgt and z <= f32(DMAX[2]):
                    pts.append((x, y, z))
    P = np.array(pts, np.float32)
    n = len(P)
    zz = np.zeros(n, np.float32)
    sc = {"x": P[:, 0], "y": P[:, 1], "z": P[:, 2], "vx": zz, "vy": zz, "vz": zz,
          "mass": np.full(n, m, np.float32), "density": np.full(n, 1000.0, np.float32),
          "temperature": np.full(n, 293.15, np.float32), "fluid_type": np.zeros(n, np.uint8),
          "h": f32(1.3) * SPACING, "gravity": (0.0, -9.81, 0.0), "speed_of_sound": f32(20.0),
          "cfl_number": f32(0.4), "viscosity": 0.001, "domain_min": DMIN, "domain_max": DMAX,
          "alpha": f32(1.0), "beta": f32(2.0), "boundary": B}
    return sc, float(a)


def run_fronts(a, step, timestep, front):
    """Host-dt stepping with a snapshot every 10 ms until T* > 3.5
    (benchmarks.rs:136-204); returns (T*, Z*) per snapshot."""
    t, nxt, out = 0.0, 0.01, []
    while True:
        dt = timestep()
        step(np.float32(dt))
        t += float(np.float32(dt))
        if t >= nxt:
            nxt += 0.01
            ts = t * math.sqrt(2 * 9.81 / a)
            if ts > 3.5:
                return out
            out.append((ts, front() / a))


def test_oracle_front_against_martin_moyce():
    sc, a = scene()
    assert len(sc["x"]) == 156
    ref = front_reference()
    o = po.Oracle(sc, precision=32, flags=po.ENGINE)
    rows = run_fronts(a, o.step, o.timestep, lambda: float(max(o.get("x").max(), 0.0)))
    assert len(rows) >= 8
    errs = [abs(z - ref(t)) / ref(t) for t, z in rows]
    assert max(errs) <= 0.15, rows
    assert all(z1 >= z0 - 1e-6 for (_, z0), (_, z1) in zip(rows, rows[1:])), "the front moves forward"


@pytest.mark.gpu
def test_engine_steps_the_benchmark_like_the_oracle():
    import ultraballpit_amd as ubp
    from parity_util import assert_bitwise
    if not ubp.gpu_available():
        pytest.fail("no gfx950 device / engine library not loadable")
    sc, a = scene()
    print("case: Martin-Moyce benchmark set-up, 156 fluid + boundary particles, host CFL dt to T* = 3.5", flush=True)
    e = ubp.HipKernel.from_scene(sc)
    o = po.Oracle(sc, precision=32, flags=po.ENGINE)

    def step(dt):
        e.step(float(dt))
        o.step(dt)

    def timestep():
        d_e, d_o = np.float32(e.dt()), np.float32(o.timestep())
        assert d_e == d_o
        return d_o

    rows = run_fronts(a, step, timestep, lambda: float(max(e.particles()["x"].max(), 0.0)))
    assert len(rows) >= 8
    assert_bitwise(e.particles(), o, "Martin-Moyce set-up at T* = 3.5")
