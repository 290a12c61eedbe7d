"""Domain set-up and analytic SDF node grids on the device.

``setup_domain`` replaces the orchestrator's set-up
(backend/crates/orchestrator/src/domain.rs:35-402): fluid particles on the
(i + 0.5) spacing lattice minus the sites outside the domain box or inside
the geometry, three wall layers per Wall face, three layers under the
geometry surface.  The configuration is a dict with the fields of
SimulationConfig it reads (config.rs): ``domain`` {``min``, ``max``},
``particle_spacing``, ``fluid_type`` ("Water", "Air", "Mixed"),
``initial_temperature`` and ``boundary_conditions`` (face -> "Wall" or
another boundary type; missing faces are Walls, BoundaryConditions::default,
config.rs:156-167).  Results are ParticleArrays-shaped dicts in the
reference's order.

``box_sdf``, ``sphere_in_box_sdf`` and ``sphere_sdf`` give the same node
grids as the host builders of the same names in ``scenes`` (bit for bit),
computed on the device.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import (EngineError, _BoundaryOut, _Domain, _ParticlesOut, _check, _f32, _fp, _u32p, FIELDS, lib)

FLUID_TYPES = {"Water": 0, "Air": 1, "Mixed": 2}
FACES = ("x_min", "x_max", "y_min", "y_max", "z_min", "z_max")
BOUNDARY_FIELDS = ("x", "y", "z", "mass", "nx", "ny", "nz")


def _wall_mask(bc: dict | None) -> int:
    bc = bc or {}
    mask = 0
    for f, face in enumerate(FACES):
        if bc.get(face, "Wall") == "Wall":
            mask |= 1 << f
    return mask


def setup_domain(config: dict, sdf: dict):
    """(fluid particles, boundary particles) of ``config`` around the
    geometry ``sdf`` (a GridSDF node grid dict: origin, cell, dims, data)."""
    ft = config.get("fluid_type", "Water")
    if ft not in FLUID_TYPES:
        raise EngineError(f"fluid_type must be one of {sorted(FLUID_TYPES)}")
    d = _Domain()
    for a in range(3):
        d.domain_min[a] = float(config["domain"]["min"][a])
        d.domain_max[a] = float(config["domain"]["max"][a])
    d.particle_spacing = float(config["particle_spacing"])
    d.fluid_type = FLUID_TYPES[ft]
    d.initial_temperature = float(config.get("initial_temperature", 293.15))
    d.walls = _wall_mask(config.get("boundary_conditions"))
    o = _f32(sdf["origin"])
    dims = np.ascontiguousarray(sdf["dims"], np.uint32)
    data = _f32(sdf["data"]).ravel()
    nf, nb = C.c_size_t(0), C.c_size_t(0)
    _check(lib().ubp_setup_domain(C.byref(d), _fp(o), float(sdf["cell"]), _u32p(dims), _fp(data), None, 0, None, 0,
                                  C.byref(nf), C.byref(nb)), "setup_domain")
    n, m = nf.value, nb.value
    F = {k: np.zeros(n, np.float32) for k in FIELDS}
    F["fluid_type"] = np.zeros(n, np.uint8)
    B = {k: np.zeros(m, np.float32) for k in BOUNDARY_FIELDS}
    fo = _ParticlesOut()
    for k in FIELDS:
        setattr(fo, k, _fp(F[k]))
    fo.fluid_type = F["fluid_type"].ctypes.data_as(C.POINTER(C.c_uint8))
    bo = _BoundaryOut()
    for k in BOUNDARY_FIELDS:
        setattr(bo, k, _fp(B[k]))
    _check(lib().ubp_setup_domain(C.byref(d), _fp(o), float(sdf["cell"]), _u32p(dims), _fp(data), C.byref(fo), n,
                                  C.byref(bo), m, C.byref(nf), C.byref(nb)), "setup_domain")
    return F, B


def _analytic(kind: int, params, origin, cell: float, dims, out=None) -> np.ndarray:
    prm = np.zeros(6, np.float64)
    prm[:len(params)] = params
    org = np.ascontiguousarray(origin, np.float64)
    d = np.ascontiguousarray(dims, np.uint32)
    n = int(np.prod(d.astype(np.int64)))
    combine = out is not None
    buf = np.array(out, np.float32).ravel() if combine else np.zeros(n, np.float32)
    _check(lib().ubp_sdf_analytic(kind, prm.ctypes.data_as(C.POINTER(C.c_double)),
                                  org.ctypes.data_as(C.POINTER(C.c_double)), float(cell), _u32p(d), int(combine),
                                  _fp(buf)), "sdf_analytic")
    return buf


def box_sdf(lo, hi, dims, cell=None):
    """Signed distance to the inside of the box [lo, hi] (positive inside) on
    ``dims`` nodes centred on the box; ``cell`` defaults to the smallest
    spacing that leaves at least one cell of margin on every side."""
    lo = np.asarray(lo, np.float64)
    hi = np.asarray(hi, np.float64)
    dims = np.asarray(dims, np.int64)
    ext = hi - lo
    if cell is None:
        cell = float(np.max(ext / (dims - 3)))
    origin = lo - 0.5 * ((dims - 1) * cell - ext)
    data = _analytic(0, np.concatenate([lo, hi]), origin, cell, dims)
    return {"origin": origin.astype(np.float32), "cell": np.float32(cell), "dims": dims.astype(np.uint32),
            "data": data}


def sphere_in_box_sdf(lo, hi, centre, radius, dims, cell=None):
    """min(box interior distance, distance outside a solid sphere)."""
    s = box_sdf(lo, hi, dims, cell)
    o = s["origin"].astype(np.float64)
    s["data"] = _analytic(1, list(centre) + [radius], o, float(s["cell"]), s["dims"], out=s["data"])
    return s


def sphere_sdf(centre, radius, like: dict) -> dict:
    """|x - c| - r on the node grid of ``like``."""
    o = np.asarray(like["origin"], np.float32).astype(np.float64)
    data = _analytic(1, list(centre) + [radius], o, float(like["cell"]), like["dims"])
    return {"origin": np.asarray(like["origin"], np.float32), "cell": np.float32(like["cell"]),
            "dims": np.asarray(like["dims"], np.uint32), "data": data}


__all__ = ["setup_domain", "box_sdf", "sphere_in_box_sdf", "sphere_sdf", "FLUID_TYPES", "FACES"]
