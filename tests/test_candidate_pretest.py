"""The candidate test's integer pre-test decides like the FP32 test outside its band.

The neighbour search (``k_neigh``) first classifies each candidate by an exact
integer squared distance of 16-bit cell-relative coordinates: squared unit
distances <= qlo are neighbours, > qhi are not, and only the band between
takes the FP32 test ``(dx*dx + dy*dy) + dz*dz <= (2h)^2``
(``kernel/src/neighbor.rs:144``).  The thresholds come from a rounding bound
(``set_qtest`` in ``ubp_engine.hip``, restated below).  This test samples
adversarial pairs -- distances within 0.1% of 2h, particles on and next to
cell faces, coordinates near the far end of the domain -- with the engine's
FP32 arithmetic emulated in numpy, and checks that every pair outside the band
is classified as the FP32 test classifies it.  A band that is too thin, a
wrong sign of the cell shift, or a cell corner computed differently from the
binning fails here.
"""
import math

import numpy as np

QS = 12000
f32 = np.float32


def set_qtest(h, dmin, dmax):
    """Restatement of set_qtest (ubp_engine.hip) on the engine's FP32 params."""
    h = f32(h)
    cell = f32(2.0) * h
    sr = f32(2.0) * h
    rsq = sr * sr
    dims = [max(1, int(math.ceil(float(f32(f32(dmax[a]) - f32(dmin[a])) / cell)))) for a in range(3)]
    qscale = f32(QS / float(cell))
    L = 0.0
    for a in range(3):
        L = max(L, abs(float(dmin[a])), abs(float(dmin[a]) + dims[a] * float(cell)), abs(float(dmax[a])))
    ulpL = math.ldexp(1.0, math.frexp(L)[1] - 1 - 23)
    e = 1.0 + 4.0 * ulpL * float(qscale) + 0.01
    ev = math.sqrt(3.0) * e
    rsq_u = float(rsq) * float(qscale) * float(qscale)
    band = 2.0 * (ev * (2.0 * math.sqrt(rsq_u) + ev) + 1e-6 * rsq_u)
    use = band < 0.01 * rsq_u and rsq_u + band < 1.5e9
    return dict(cell=cell, rsq=rsq, dims=dims, qscale=qscale, use=use,
                qlo=int(math.floor(rsq_u - band)), qhi=int(math.ceil(rsq_u + band)),
                dmin=[f32(v) for v in dmin])


def axis_cell(p, gmin, cell, dim):
    c = np.floor((p - gmin) / cell)          # float32 throughout (neighbor.rs:57-71)
    c = np.where(c >= 0, c, f32(0.0))
    c = np.minimum(c, f32(dim - 1))
    return c.astype(np.int64)


def quantize(p, c, gmin, cell, qscale):
    corner = gmin + c.astype(np.float32) * cell
    return np.rint((p - corner) * qscale).astype(np.int64)


def classify(Q, pi, pj):
    """(integer d2, FP32 accept) for pairs (pi, pj), j's cell a neighbour of i's."""
    ci = [axis_cell(pi[:, a], Q["dmin"][a], Q["cell"], Q["dims"][a]) for a in range(3)]
    cj = [axis_cell(pj[:, a], Q["dmin"][a], Q["cell"], Q["dims"][a]) for a in range(3)]
    near = np.all([np.abs(ci[a] - cj[a]) <= 1 for a in range(3)], axis=0)
    d2i = np.zeros(len(pi), np.int64)
    for a in range(3):
        qi = quantize(pi[:, a], ci[a], Q["dmin"][a], Q["cell"], Q["qscale"])
        qj = quantize(pj[:, a], cj[a], Q["dmin"][a], Q["cell"], Q["qscale"])
        assert np.all((qi >= -64) & (qi <= QS + 64)) and np.all((qj >= -64) & (qj <= QS + 64))
        d = (qi - (cj[a] - ci[a]) * QS) - qj   # lane coordinate shifted to j's cell corner
        assert np.all(np.abs(d) < 2 ** 15)
        d2i += d * d
    dx = pi[:, 0] - pj[:, 0]
    dy = pi[:, 1] - pj[:, 1]
    dz = pi[:, 2] - pj[:, 2]
    d2 = (dx * dx + dy * dy) + dz * dz
    return d2i[near], (d2 <= Q["rsq"])[near]


def pairs(rng, Q, dmin, dmax, n):
    """Pairs at distances within 0.1% of 2h, i on or next to cell faces."""
    lo = np.array(dmin, np.float64)
    hi = np.array(dmax, np.float64)
    cell = float(Q["cell"])
    pi = lo + rng.random((n, 3)) * (hi - lo)
    # a third of the particles snapped onto cell faces (and one ulp either side)
    snap = rng.random(n) < 0.33
    faces = lo + np.round((pi - lo) / cell) * cell
    pi = np.where(snap[:, None], faces, pi).astype(np.float32)
    nud = rng.integers(-1, 2, size=(n, 3)).astype(np.float32)
    pi = np.where(snap[:, None], np.nextafter(pi, pi + nud), pi)
    u = rng.normal(size=(n, 3))
    u /= np.linalg.norm(u, axis=1)[:, None]
    r = cell * (1.0 + (rng.random(n) - 0.5) * 2e-3)
    pj = (pi.astype(np.float64) + u * r[:, None]).astype(np.float32)
    pj = np.clip(pj, np.array(dmin, np.float32), np.array(dmax, np.float32))
    pi = np.clip(pi, np.array(dmin, np.float32), np.array(dmax, np.float32))
    return pi, pj


SCENES = [
    # (h, domain min, domain max): the 1M dam break, the flow past a sphere,
    # the water box, the sweep's coarsest scene and a domain far from the origin
    (1.3 * 0.25 / 81, (0.0, 0.0, 0.0), (1.0, 0.25, 0.6)),
    (1.3 * 1.5625e-3, (0.0, 0.0, 0.0), (0.4, 0.2, 0.2)),
    (0.65e-3, (0.0, 0.0, 0.0), (0.01, 0.01, 0.0125)),
    (1.3 * 0.02, (0.0, 0.0, 0.0), (1.0, 0.25, 0.6)),
    (5e-3, (10.0, -3.0, 7.0), (11.0, -2.0, 8.0)),
]


def test_pretest_outside_band_equals_fp32_test():
    rng = np.random.default_rng(20261204)
    for h, dmin, dmax in SCENES:
        Q = set_qtest(h, dmin, dmax)
        assert Q["use"], (h, dmin)
        assert Q["qlo"] < Q["qhi"]
        tested = 0
        for _ in range(8):
            pi, pj = pairs(rng, Q, dmin, dmax, 250_000)
            d2i, acc = classify(Q, pi, pj)
            sure_in = d2i <= Q["qlo"]
            sure_out = d2i > Q["qhi"]
            assert np.all(acc[sure_in]), (h, dmin, "pre-test accepted a pair the FP32 test rejects")
            assert not np.any(acc[sure_out]), (h, dmin, "pre-test rejected a pair the FP32 test accepts")
            tested += len(d2i)
        assert tested > 1_000_000


def test_pretest_band_is_thin_for_the_benchmark_scenes():
    # candidates re-tested in FP32: a band at most 0.25% of (2h)^2 wide
    for h, dmin, dmax in SCENES[:3]:
        Q = set_qtest(h, dmin, dmax)
        rsq_u = float(Q["rsq"]) * float(Q["qscale"]) ** 2
        assert (Q["qhi"] - Q["qlo"]) / rsq_u < 2.5e-3


def test_pretest_off_when_coordinates_dwarf_the_cell():
    # 1 mm cells at 5 km from the origin: the corners' rounding is too coarse
    Q = set_qtest(0.5e-3, (5000.0, 0.0, 0.0), (5000.5, 0.5, 0.5))
    assert not Q["use"]
