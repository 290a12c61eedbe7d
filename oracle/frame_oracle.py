"""CPU ORACLE of the viewer frame (TEST INFRASTRUCTURE ONLY).

Restates build_frame (backend/crates/server/src/ws.rs:232-294) over ParticleArrays-shaped
arrays in original particle order: tag 0x02, frame number u64, particle
count u32, sim time f64, then for original indices k * stride (stride =
max(n / subsample, 1), or 1 for subsample 0, count = min(n / stride,
subsample), so subsample 0 sends no particles): x y z
temperature f32, fluid type u8, density ratio u16 =
((density / rest_density) * 1000).clamp(0, 65535) as u16 (truncation;
NaN -> 0), reserved u8, all little endian.
"""
from __future__ import annotations

import struct

import numpy as np

REST = {0: np.float32(1000.0), 1: np.float32(1.204)}  # backend/crates/kernel/src/eos.rs:7-10


def density_ratio(density, fluid_type) -> int:
    r = np.float32(np.float32(density) / REST[int(fluid_type)]) * np.float32(1000.0)
    r = np.float32(r)
    if np.isnan(r):
        return 0
    r = min(max(float(r), 0.0), 65535.0)
    return int(r)  # truncation toward zero


def build_frame(p: dict, subsample: int, frame_number: int, sim_time: float) -> bytes:
    n = len(p["x"])
    stride = max(n // subsample, 1) if subsample > 0 else 1
    count = min(n // stride, subsample)
    out = bytearray(b"\x02")
    out += struct.pack("<QId", frame_number, count, sim_time)
    for k in range(count):
        i = k * stride
        out += struct.pack("<4f", p["x"][i], p["y"][i], p["z"][i], p["temperature"][i])
        ft = int(p["fluid_type"][i])
        out += struct.pack("<BHB", ft, density_ratio(p["density"][i], ft), 0)
    return bytes(out)


def parse_frame(buf: bytes):
    tag, fn, count, t = struct.unpack_from("<BQId", buf, 0)
    rows = [struct.unpack_from("<4fBHB", buf, 21 + 20 * k) for k in range(count)]
    return tag, fn, count, t, rows
