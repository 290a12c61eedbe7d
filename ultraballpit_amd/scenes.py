"""Seeded synthetic scenes (host-side set-up, the part of the orchestrator's
domain set-up the step engine needs).

Particle placement follows the lattice rules of the reference's
``place_fluid_particles`` (backend/crates/orchestrator/src/domain.rs:88-159:
sites at ``(i + 0.5) * dx``, mass ``rho0 * dx^3``, density ``rho0``,
temperature 293.15 K, zero velocity) and the wall-particle rule of
``generate_wall_boundary_particles`` (domain.rs:162-348: three layers at
``(layer + 0.5) * dx`` outside each wall).  Randomness comes from a
counter-based splitmix64 stream so every consumer (engine, oracle, batch
runner) sees identical bytes for the same seed.

A scene is a plain dict:  ``x y z vx vy vz mass density temperature
fluid_type`` (float32/uint8 arrays in original particle order) plus the
``CpuKernel::new`` scalars (lib.rs:138-148) ``h gravity speed_of_sound
cfl_number viscosity domain_min domain_max``, the Monaghan ``alpha``/``beta``
(sph.rs:440-441 fixes 1 and 2), an optional ``sdf`` node grid (GridSDF
layout, geometry.rs:11-20: z fastest) and optional ``boundary`` arrays.
"""
from __future__ import annotations

import math

import numpy as np

WATER_RHO0 = 1000.0  # eos.rs:7
GOLDEN = np.uint64(0x9E3779B97F4A7C15)


def splitmix64(counter: np.ndarray) -> np.ndarray:
    """splitmix64 finaliser applied to ``seed_state + k * golden``."""
    with np.errstate(over="ignore"):
        z = counter.astype(np.uint64) * GOLDEN
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return z ^ (z >> np.uint64(31))


def uniform(seed: int, stream: int, n: int) -> np.ndarray:
    """n doubles in [0, 1) from stream ``stream`` of ``seed`` (53-bit)."""
    base = np.uint64((seed * 0x100000001B3 + stream * 0x1F3D5B79) & 0xFFFFFFFFFFFFFFFF)
    with np.errstate(over="ignore"):
        k = np.arange(1, n + 1, dtype=np.uint64) + base * np.uint64(1 << 20)
    return (splitmix64(k) >> np.uint64(11)).astype(np.float64) * (1.0 / (1 << 53))


def normal(seed: int, stream: int, n: int) -> np.ndarray:
    u1 = uniform(seed, stream, n)
    u2 = uniform(seed, stream + 1, n)
    u1 = np.maximum(u1, 1e-300)
    return np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * np.pi * u2)


def box_sdf(lo, hi, dims, cell=None, margin=None):
    """Signed distance to the inside of the box [lo, hi] (positive inside),
    sampled on a GridSDF node grid of ``dims`` nodes centred on the box.
    ``cell`` defaults to the smallest spacing that leaves at least one cell of
    margin on every side."""
    lo = np.asarray(lo, np.float64)
    hi = np.asarray(hi, np.float64)
    dims = np.asarray(dims, np.int64)
    ext = hi - lo
    if cell is None:
        cell = float(np.max((ext) / (dims - 3)))
    span = (dims - 1) * cell
    origin = lo - 0.5 * (span - ext)
    gx = origin[0] + cell * np.arange(dims[0])
    gy = origin[1] + cell * np.arange(dims[1])
    gz = origin[2] + cell * np.arange(dims[2])
    X, Y, Z = np.meshgrid(gx, gy, gz, indexing="ij")  # i (x) slowest, k (z) fastest
    P = np.stack([X, Y, Z], axis=-1)
    q = np.maximum(lo - P, P - hi)  # per-axis outside distance (negative inside)
    outside = np.linalg.norm(np.maximum(q, 0.0), axis=-1)
    inside = np.minimum(np.max(q, axis=-1), 0.0)
    d = -(outside + inside)  # positive inside the box
    return {"origin": origin.astype(np.float32), "cell": np.float32(cell),
            "dims": dims.astype(np.uint32), "data": d.astype(np.float32).ravel()}


def sphere_in_box_sdf(lo, hi, centre, radius, dims, cell=None):
    """min(box interior distance, distance outside a solid sphere)."""
    s = box_sdf(lo, hi, dims, cell)
    dims = s["dims"].astype(np.int64)
    o = s["origin"].astype(np.float64)
    c = float(s["cell"])
    gx = o[0] + c * np.arange(dims[0])
    gy = o[1] + c * np.arange(dims[1])
    gz = o[2] + c * np.arange(dims[2])
    X, Y, Z = np.meshgrid(gx, gy, gz, indexing="ij")
    ds = np.sqrt((X - centre[0]) ** 2 + (Y - centre[1]) ** 2 + (Z - centre[2]) ** 2) - radius
    d = np.minimum(s["data"].astype(np.float64).reshape(dims), ds)
    s["data"] = d.astype(np.float32).ravel()
    return s


def sphere_sdf(centre, radius, like: dict) -> dict:
    """Analytic sphere distance |x - c| - r on the node grid of `like`."""
    dims = np.asarray(like["dims"]).astype(np.int64)
    o = np.asarray(like["origin"], np.float64)
    c = float(like["cell"])
    gx = o[0] + c * np.arange(dims[0])
    gy = o[1] + c * np.arange(dims[1])
    gz = o[2] + c * np.arange(dims[2])
    X, Y, Z = np.meshgrid(gx, gy, gz, indexing="ij")
    d = np.sqrt((X - centre[0]) ** 2 + (Y - centre[1]) ** 2 + (Z - centre[2]) ** 2) - radius
    return {"origin": np.asarray(like["origin"], np.float32), "cell": np.float32(like["cell"]),
            "dims": np.asarray(like["dims"], np.uint32), "data": d.astype(np.float32).ravel()}


def _particles(pos: np.ndarray, dx: float, rho0: float = WATER_RHO0) -> dict:
    n = len(pos)
    m = np.float32(rho0 * dx * dx * dx)
    z = np.zeros(n, np.float32)
    return {
        "x": pos[:, 0].astype(np.float32), "y": pos[:, 1].astype(np.float32),
        "z": pos[:, 2].astype(np.float32),
        "vx": z.copy(), "vy": z.copy(), "vz": z.copy(),
        "mass": np.full(n, m, np.float32),
        "density": np.full(n, rho0, np.float32),
        "temperature": np.full(n, 293.15, np.float32),
        "fluid_type": np.zeros(n, np.uint8),
    }


def lattice(nx, ny, nz, dx, origin=(0.0, 0.0, 0.0), n_keep=None):
    """Sites ``origin + (i + 0.5) dx`` in z-outer / y / x-inner order."""
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    pos = np.stack([i.ravel(), j.ravel(), k.ravel()], axis=1).astype(np.float64)
    pos = (pos + 0.5) * dx + np.asarray(origin, np.float64)
    if n_keep is not None:
        pos = pos[:n_keep]
    return pos


def water_box(n_side=20, dx=0.5e-3, box=(0.01, 0.01, 0.0125), sdf_nodes=64,
              alpha=0.01, beta=0.0, cfl=0.25, with_sdf=True):
    """Water box settling: n_side^3 lattice filling the lower 80% of a closed
    box, g = -9.81 z, c0 = 10 sqrt(2 g H) with H the water height.  The
    16-layer / 12.5 mm ambiguity of the 8000-particle case is resolved by a
    10 x 10 x 12.5 mm box."""
    pos = lattice(n_side, n_side, n_side, dx)
    sc = _particles(pos, dx)
    H = n_side * dx
    g = 9.81
    sc.update({
        "name": f"water_box_{n_side}^3",
        "h": np.float32(1.3 * dx), "dx": dx,
        "gravity": (0.0, 0.0, -g),
        "speed_of_sound": np.float32(10.0 * math.sqrt(2.0 * g * H)),
        "cfl_number": np.float32(cfl), "viscosity": 0.0,
        "domain_min": (0.0, 0.0, 0.0), "domain_max": tuple(float(b) for b in box),
        "alpha": np.float32(alpha), "beta": np.float32(beta),
    })
    if with_sdf:
        sc["sdf"] = box_sdf((0, 0, 0), box, (sdf_nodes,) * 3)
    return sc


def wall_particles(dmin, dmax, dx, layers=3, walls=("x-", "x+", "y-", "y+", "z-", "z+")):
    """Boundary particles of the closed-box walls, ``layers`` layers at
    ``(layer + 0.5) dx`` outside each wall face, covering the face plus the
    edge/corner regions of the layers (domain.rs:162-348)."""
    dmin = np.asarray(dmin, np.float64)
    dmax = np.asarray(dmax, np.float64)
    counts = np.maximum(np.round((dmax - dmin) / dx).astype(np.int64), 1)
    pts, nrm = [], []
    for w in walls:
        a = "xyz".index(w[0])
        sgn = -1.0 if w[1] == "-" else 1.0
        o1, o2 = [b for b in range(3) if b != a]
        r1 = dmin[o1] + (np.arange(-layers, counts[o1] + layers) + 0.5) * dx
        r2 = dmin[o2] + (np.arange(-layers, counts[o2] + layers) + 0.5) * dx
        if a > o1:  # avoid duplicating edges: the earlier axis owns them
            r1 = dmin[o1] + (np.arange(0, counts[o1]) + 0.5) * dx
        if a > o2:
            r2 = dmin[o2] + (np.arange(0, counts[o2]) + 0.5) * dx
        for layer in range(layers):
            off = (layer + 0.5) * dx
            c = dmin[a] - off if sgn < 0 else dmax[a] + off
            A, B = np.meshgrid(r1, r2, indexing="ij")
            P = np.zeros((A.size, 3))
            P[:, a] = c
            P[:, o1] = A.ravel()
            P[:, o2] = B.ravel()
            N = np.zeros((A.size, 3))
            N[:, a] = -sgn  # normal points into the fluid
            pts.append(P)
            nrm.append(N)
    P = np.concatenate(pts)
    N = np.concatenate(nrm)
    m = np.float32(WATER_RHO0 * dx ** 3)
    return {"x": P[:, 0].astype(np.float32), "y": P[:, 1].astype(np.float32),
            "z": P[:, 2].astype(np.float32), "mass": np.full(len(P), m, np.float32),
            "nx": N[:, 0].astype(np.float32), "ny": N[:, 1].astype(np.float32),
            "nz": N[:, 2].astype(np.float32)}


def water_box_walls(n_side=10, dx=1.0e-3, alpha=1.0, beta=2.0):
    """The reference wall model: lattice water in a closed box, three layers
    of wall particles, domain clamp, no SDF (faithful-mode set-up)."""
    box = (n_side * dx, n_side * dx, 1.25 * n_side * dx)
    sc = water_box(n_side=n_side, dx=dx, box=box, alpha=alpha, beta=beta, with_sdf=False)
    sc["boundary"] = wall_particles((0, 0, 0), box, dx)
    sc["name"] = f"water_box_walls_{n_side}^3"
    return sc


def dam_break(lat=(81, 81, 162), n_keep=1 << 20, tank=(1.0, 0.25, 0.6), column=(0.25, 0.25, 0.5),
              c0=30.0, alpha=0.02, beta=0.0, cfl=0.25, jitter=0.05, seed=2024,
              sdf_nodes=(256, 64, 160)):
    """3D dam break: lattice lat[0] x lat[1] x lat[2] over the water column
    (z-outer/y/x-inner order), first ``n_keep`` sites kept, seeded uniform
    jitter of +-jitter*dx per axis, h = 1.3 dx, closed tank with a box SDF."""
    dx = column[0] / lat[0]
    pos = lattice(lat[0], lat[1], lat[2], dx, n_keep=n_keep)
    n = len(pos)
    if jitter > 0:
        for a in range(3):
            pos[:, a] += (2.0 * uniform(seed, a, n) - 1.0) * jitter * dx
    sc = _particles(pos, dx)
    sc.update({
        "name": f"dam_break_{n}",
        "h": np.float32(1.3 * dx), "dx": dx,
        "gravity": (0.0, 0.0, -9.81),
        "speed_of_sound": np.float32(c0),
        "cfl_number": np.float32(cfl), "viscosity": 0.0,
        "domain_min": (0.0, 0.0, 0.0), "domain_max": tuple(float(t) for t in tank),
        "alpha": np.float32(alpha), "beta": np.float32(beta),
        "sdf": box_sdf((0, 0, 0), tank, sdf_nodes),
    })
    return sc


def dam_break_small(lat=(16, 16, 32), **kw):
    """The dam break at a coarser lattice (same tank and proportions)."""
    n = lat[0] * lat[1] * lat[2]
    kw.setdefault("sdf_nodes", (64, 24, 40))
    return dam_break(lat=lat, n_keep=n, **kw)


def flow_past_sphere(lat=(256, 128, 128), channel=(0.4, 0.2, 0.2), centre=(0.1, 0.1, 0.1),
                     radius=0.03, u0=0.5, noise=0.01, alpha=0.01, beta=0.0, seed=7,
                     sdf_nodes=(256, 128, 128), cfl=0.25):
    """Channel flow past a sphere: lattice minus the sites inside the sphere,
    uniform u0 along x plus seeded 1% Gaussian velocity noise, c0 = 10*u*3."""
    dx = channel[0] / lat[0]
    pos = lattice(lat[0], lat[1], lat[2], dx)
    keep = np.linalg.norm(pos - np.asarray(centre), axis=1) > radius + 0.5 * dx
    pos = pos[keep]
    n = len(pos)
    sc = _particles(pos, dx)
    sc["vx"] = (u0 * (1.0 + noise * normal(seed, 0, n))).astype(np.float32)
    sc["vy"] = (u0 * noise * normal(seed, 2, n)).astype(np.float32)
    sc["vz"] = (u0 * noise * normal(seed, 4, n)).astype(np.float32)
    sc.update({
        "name": f"flow_past_sphere_{n}",
        "h": np.float32(1.3 * dx), "dx": dx,
        "gravity": (0.0, 0.0, 0.0),
        "speed_of_sound": np.float32(10.0 * u0 * 3.0),
        "cfl_number": np.float32(cfl), "viscosity": 0.0,
        "domain_min": (0.0, 0.0, 0.0), "domain_max": tuple(float(c) for c in channel),
        "alpha": np.float32(alpha), "beta": np.float32(beta),
        "sdf": sphere_in_box_sdf((0, 0, 0), channel, centre, radius, sdf_nodes),
        "sphere": {"centre": centre, "radius": radius},
    })
    return sc


def flow_past_sphere_slab(slab=2, nslabs=8, halo_h=6.0, **kw):
    """One GPU's share of the flow past a sphere cut into ``nslabs`` slabs
    along x (the channel's longest axis): the particles of the full scene
    (same lattice, same seeded noise) with x in the slab widened by
    ``halo_h`` * h on each side.  6 h = 3 x 2h is the halo a step needs when
    the halo is exchanged once per step (XSPH -> drift -> density -> forces).
    The slab domain and its box-and-sphere SDF end at the widened slab, so
    its faces act as walls here; in a sharded run they are exchange faces.
    ``owned`` marks the slab's own particles."""
    full = flow_past_sphere(**kw)
    channel = full["domain_max"]
    w = channel[0] / nslabs
    x0, x1 = slab * w, (slab + 1) * w
    halo = halo_h * float(full["h"])
    lo, hi = max(0.0, x0 - halo), min(channel[0], x1 + halo)
    x = full["x"].astype(np.float64)
    keep = (x >= lo) & (x < hi)
    sc = {k: (v[keep] if isinstance(v, np.ndarray) and v.shape[:1] == x.shape else v) for k, v in full.items()}
    sc["owned"] = (x[keep] >= x0) & (x[keep] < x1)
    cell = float(full["sdf"]["cell"])
    dims = (int(np.ceil((hi - lo) / cell)) + 3, int(full["sdf"]["dims"][1]), int(full["sdf"]["dims"][2]))
    sc["sdf"] = sphere_in_box_sdf((lo, 0.0, 0.0), (hi, channel[1], channel[2]), full["sphere"]["centre"],
                                  full["sphere"]["radius"], dims, cell)
    sc["domain_min"] = (lo, 0.0, 0.0)
    sc["domain_max"] = (hi, channel[1], channel[2])
    sc["slab"] = {"index": slab, "of": nslabs, "x": (x0, x1), "halo": halo, "domain_x": (lo, hi),
                  "owned": int(sc["owned"].sum()), "halo_particles": int((~sc["owned"]).sum())}
    sc["name"] = f"flow_past_sphere_slab_{slab}_of_{nslabs}"
    return sc


def sweep_scenes(n_scenes=256, n_min=2000, n_max=200000, seed=11, tank=(1.0, 0.25, 0.6),
                 sdf_nodes=(64, 24, 40)):
    """Batched parameter sweep: dam-break scenes with particle counts
    log-uniform in [n_min, n_max], fill height U[0.3, 0.7] of the tank and
    alpha U[0.005, 0.05], all from ``seed``."""
    u_n = uniform(seed, 100, n_scenes)
    u_h = uniform(seed, 101, n_scenes)
    u_a = uniform(seed, 102, n_scenes)
    scenes = []
    for s in range(n_scenes):
        n_target = int(round(math.exp(math.log(n_min) + u_n[s] * (math.log(n_max) - math.log(n_min)))))
        fill = 0.3 + 0.4 * u_h[s]
        col = (0.25, 0.25, fill * tank[2])
        # lattice with ~n_target sites over the column
        vol = col[0] * col[1] * col[2]
        dx = (vol / n_target) ** (1.0 / 3.0)
        lat = tuple(max(1, int(round(c / dx))) for c in col)
        dx = col[0] / lat[0]
        sc = dam_break(lat=lat, n_keep=lat[0] * lat[1] * lat[2], tank=tank, column=col,
                       alpha=float(0.005 + 0.045 * u_a[s]), seed=seed * 1000 + s,
                       sdf_nodes=sdf_nodes)
        sc["name"] = f"sweep_{s}"
        scenes.append(sc)
    return scenes
