"""Pins the surrogate network's CPU oracle: finite-difference gradients of the
unrounded FP64 network, BF16 rounding against torch's independent cast, and
closeness of the BF16-rounded network to its FP64 twin."""
import numpy as np
import pytest

import mlp_oracle as mo


def small_case(seed=0, rows=40):
    rng = np.random.default_rng(seed)
    m = mo.init_model(seed)
    for k in ("b1", "b2", "b3", "b4"):
        m[k] = rng.normal(0, 0.1, m[k].shape).astype(np.float32)
    x = rng.normal(0, 1, (rows, mo.IN)).astype(np.float32)
    t = rng.normal(0, 1, (rows, mo.OUT)).astype(np.float32)
    return m, x, t


def test_finite_difference_gradients_fp64():
    m, x, t = small_case()
    m64 = {k: np.asarray(v, np.float64) for k, v in m.items()}
    L, g = mo.loss_grad(m64, x, t, rounding=False)
    rng = np.random.default_rng(1)
    eps = 1e-6
    for k, shape in mo.PARAMS:
        for _ in range(12):
            i = tuple(int(rng.integers(0, s)) for s in shape)
            mp = {kk: v.copy() for kk, v in m64.items()}
            mm = {kk: v.copy() for kk, v in m64.items()}
            mp[k][i] += eps
            mm[k][i] -= eps
            fd = (mo.loss_grad(mp, x, t, rounding=False)[0] - mo.loss_grad(mm, x, t, rounding=False)[0]) / (2 * eps)
            an = g[k][i]
            assert abs(fd - an) <= 1e-6 * max(1.0, abs(an)) + 1e-8, (k, i, fd, an)


def test_gradient_catches_planted_mistake():
    # a transposed weight in the backward pass is caught by the FD check
    m, x, t = small_case(seed=3)
    m64 = {k: np.asarray(v, np.float64) for k, v in m.items()}
    _, g = mo.loss_grad(m64, x, t, rounding=False)
    wrong = g["w2"].T
    eps = 1e-6
    mp = {kk: v.copy() for kk, v in m64.items()}
    mm = {kk: v.copy() for kk, v in m64.items()}
    mp["w2"][3, 7] += eps
    mm["w2"][3, 7] -= eps
    fd = (mo.loss_grad(mp, x, t, rounding=False)[0] - mo.loss_grad(mm, x, t, rounding=False)[0]) / (2 * eps)
    assert abs(fd - g["w2"][3, 7]) < 1e-8 and abs(fd - wrong[3, 7]) > 1e-6


def test_bf16_rounding_matches_torch():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(2)
    v = np.concatenate([rng.normal(0, 1, 100000), rng.normal(0, 1e-30, 1000), rng.normal(0, 1e30, 1000),
                        np.array([1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, -1.0 - 2 ** -8, 0.0, -0.0])]).astype(np.float32)
    ours = mo.bf16(v)
    ref = torch.from_numpy(v).to(torch.bfloat16).to(torch.float32).numpy()
    assert np.array_equal(ours.view(np.uint32), ref.view(np.uint32))


def test_rounded_network_close_to_fp64():
    m, x, t = small_case(seed=4, rows=200)
    y32 = mo.forward(m, x, rounding=True)
    y64 = mo.forward({k: np.asarray(v, np.float64) for k, v in m.items()}, x, rounding=False)
    assert np.max(np.abs(y32 - y64)) < 0.05 * np.max(np.abs(y64))
    L32, g32 = mo.loss_grad(m, x, t, rounding=True)
    L64, g64 = mo.loss_grad({k: np.asarray(v, np.float64) for k, v in m.items()}, x, t, rounding=False)
    assert abs(L32 - L64) < 0.02 * L64
    for k, _ in mo.PARAMS:
        # BF16 rounding of three back-propagation stages and flipped ReLU masks
        assert np.linalg.norm(g32[k] - g64[k]) <= 0.1 * np.linalg.norm(g64[k]) + 1e-9, k


def test_adam_matches_closed_form_first_step():
    # first Adam step moves every parameter by lr * sign(g) (bias-corrected)
    m, x, t = small_case(seed=5)
    _, g = mo.loss_grad(m, x, t, rounding=False)
    st = {}
    m2 = mo.adam_step(m, g, st, lr=1e-3, eps=1e-12)
    for k, _ in mo.PARAMS:
        d = m2[k].astype(np.float64) - m[k].astype(np.float64)
        nz = np.abs(g[k]) > 1e-6
        assert np.allclose(d[nz], -1e-3 * np.sign(g[k][nz]), rtol=1e-3, atol=1e-7), k
