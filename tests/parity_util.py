"""Shared helpers for the engine-vs-oracle parity tests."""
from __future__ import annotations

import numpy as np

import pyoracle
import ultraballpit_amd as ubp

STATE = ("x", "y", "z", "vx", "vy", "vz", "ax", "ay", "az", "density", "pressure", "mass")


def engine(scene):
    return ubp.HipKernel.from_scene(scene)


def oracle(scene, precision=32, flags=pyoracle.ENGINE):
    return pyoracle.Oracle(scene, precision=precision, flags=flags)


def first_mismatch(a: dict, b: dict, fields=STATE):
    """Return (field, index, a, b) of the first bitwise difference, or None."""
    for k in fields:
        x = np.asarray(a[k], np.float32)
        y = np.asarray(b[k], np.float32)
        bad = np.nonzero(x.view(np.uint32) != y.view(np.uint32))[0]
        if bad.size:
            i = int(bad[0])
            return k, i, float(x[i]), float(y[i]), int(bad.size)
    return None


def assert_bitwise(eng_state: dict, orc: "pyoracle.Oracle", where=""):
    o = {k: orc.get(k).astype(np.float32) for k in STATE}
    mm = first_mismatch(eng_state, o)
    assert mm is None, f"{where}: first mismatch {mm} (field, index, engine, oracle, count)"


def fp32_vs_fp64(orc32, orc64, h):
    """Max position error (in units of h) and max relative density error of
    an FP32 run against its FP64 twin."""
    dpos = max(np.max(np.abs(orc32.get(k) - orc64.get(k))) for k in ("x", "y", "z"))
    r64 = orc64.get("density")
    drho = np.max(np.abs(orc32.get("density") - r64) / np.abs(r64))
    return float(dpos / h), float(drho)
