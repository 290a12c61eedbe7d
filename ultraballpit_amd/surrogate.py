"""Learned surrogate: training-data generation, training and evaluation on
the GPU engine (no reference counterpart).

A correction cycle marks a plain state (its acceleration a_last and density
deviation), advances K network steps of ``factor`` x dt and one plain step.
The network predicts the effective acceleration over a network step,
a_eff = 2 (x(t + w) - x(t) - w v(t)) / w^2 with w = factor x dt: the constant
acceleration that carries a particle along the plain trajectory over the
step (acoustic oscillations that a step of w cannot resolve average out).
Training rows come from plain rollouts at fixed dt: after each mark, the
states reached after ``factor * s`` plain steps (s = 0..K) are exactly where
the network is evaluated; their features, and targets (a_eff - a_last)/|g|
completed w later, are collected on the device.  Everything runs through
the HIP library; nothing here computes physics on the CPU.
"""
from __future__ import annotations

import time

import numpy as np

from . import HipKernel, Trainer, MLP_FIELDS

K_DEFAULT = 8
FACTOR_DEFAULT = 4


def fixed_dt(scene: dict, safety: float = 0.9) -> float:
    """A fixed plain dt: `safety` x the CFL dt of the initial state."""
    e = HipKernel.from_scene(scene)
    dt = e.dt()
    e.close()
    return float(safety * dt)


def collect_rollout(trainer: Trainer, scene: dict, cycles: int, dt: float, K: int = K_DEFAULT,
                    factor: int = FACTOR_DEFAULT, frac: float = 1.0, seed: int = 0, warm_cycles: int = 0):
    """Plain rollout of `scene` at fixed dt in correction cycles (mark, then
    K landing states `factor` steps apart, then the correction step).  At
    every landing state (and the mark state) the rows of `frac` of the
    particles get their features; their targets are the effective
    accelerations over the next `factor` plain steps."""
    e = HipKernel.from_scene(scene)
    n = e.particle_count()
    rng = np.random.default_rng(seed)
    win = float(np.float32(dt)) * factor
    e.step(dt)
    for _ in range(warm_cycles * (K * factor + 1)):
        e.step_async(dt)
    e.sync()
    pending = []  # (due step, first row, ids)
    step = 0

    def advance(k):
        nonlocal step
        for _ in range(k):
            e.step_async(dt)
            step += 1
            while pending and pending[0][0] == step:
                _, first, ids = pending.pop(0)
                trainer.finish(e, first, ids, win)

    for c in range(cycles):
        e.surrogate_mark()
        for s in range(0, K + 1):
            if s > 0:
                advance(factor)
            ids = np.arange(n, dtype=np.uint32) if frac >= 1.0 else \
                np.sort(rng.choice(n, size=max(1, int(frac * n)), replace=False)).astype(np.uint32)
            first = trainer.collect(e, ids, s / K)
            pending.append((step + factor, first, ids))
        advance(1)  # the correction step of the cycle
    while pending:
        advance(1)
    e.sync()
    e.close()


def collect_corrective(trainer: Trainer, scene: dict, model: dict, cycles: int, dt: float, K: int = K_DEFAULT,
                       factor: int = FACTOR_DEFAULT, frac: float = 1.0, seed: int = 0, warm_cycles: int = 0):
    """Rows on the network's own states: from every mark the network (current
    `model`) runs K steps and its states are collected; the engine then
    returns to the mark (device checkpoint) and advances with plain steps,
    and each row's target is the pair of effective accelerations that would
    carry the network's state onto the plain trajectory one window later.
    Training on these rows teaches the network to correct its own drift."""
    e = HipKernel.from_scene(scene)
    e.surrogate_set_model(model)
    n = e.particle_count()
    rng = np.random.default_rng(seed)
    win = float(np.float32(dt)) * factor
    dtw = float(np.float32(win))
    e.step(dt)
    for _ in range(warm_cycles * (K * factor + 1)):
        e.step_async(dt)
    e.sync()
    for c in range(cycles):
        e.surrogate_mark()
        e.save_checkpoint()
        rows = []
        for s in range(0, K):
            if s > 0:
                e.surrogate_step(dtw, s / K)
            ids = np.arange(n, dtype=np.uint32) if frac >= 1.0 else \
                np.sort(rng.choice(n, size=max(1, int(frac * n)), replace=False)).astype(np.uint32)
            rows.append((trainer.collect(e, ids, s / K), ids))
        e.restore_checkpoint()
        for s in range(0, K):
            for _ in range(factor):
                e.step_async(dt)
            first, ids = rows[s]
            trainer.finish(e, first, ids, win)
        e.step_async(dt)  # the correction step of the cycle
    e.sync()
    e.close()


def train(scenes: list, cycles: int, steps: int, batch: int = 32768, lr: float = 2e-3, lr_end: float = 1e-4,
          K: int = K_DEFAULT, factor: int = FACTOR_DEFAULT, frac: float = 1.0, seed: int = 1,
          warm_cycles: int = 0, rounds: int = 1, log=None):
    """Train a network on plain rollouts of `scenes` (each at its own fixed
    dt); each further round adds rows on the network's own states
    (collect_corrective) and retrains on all rows.  Returns (model, losses,
    info)."""
    per_round = sum(int(len(s["x"]) * min(frac, 1.0)) + 1 for s in scenes) * cycles * (K + 1)
    rows = per_round * max(1, rounds)
    tr = Trainer(rows, seed)
    t0 = time.perf_counter()
    fit_s = 0.0
    model = None
    losses = np.zeros(0, np.float32)
    for rnd in range(max(1, rounds)):
        for k, sc in enumerate(scenes):
            if rnd == 0:
                collect_rollout(tr, sc, cycles, fixed_dt(sc), K, factor, frac, seed + k, warm_cycles)
            else:
                collect_corrective(tr, sc, model, cycles, fixed_dt(sc), K, factor, frac, seed + 100 * rnd + k,
                                   warm_cycles)
            if log:
                log(f"[surrogate] round {rnd} rollout {k + 1}/{len(scenes)}: {tr.rows()} rows")
        t1 = time.perf_counter()
        tr.init_model(None)
        losses = tr.fit(steps, batch, lr, lr_end)
        fit_s += time.perf_counter() - t1
        model = tr.model()
        if log:
            log(f"[surrogate] round {rnd}: loss {float(losses[:10].mean()):.4f} -> {float(losses[-10:].mean()):.4f}")
    t2 = time.perf_counter()
    tr.close()
    return model, losses, {"rows": tr_rows_or(rows), "rounds": max(1, rounds), "collect_s": t2 - t0 - fit_s,
                           "fit_s": fit_s,
                           "loss_first": float(losses[:10].mean()) if len(losses) else None,
                           "loss_last": float(losses[-10:].mean()) if len(losses) else None}


def tr_rows_or(rows):
    return int(rows)


def _positions(e: HipKernel) -> np.ndarray:
    p = e.particles()
    return np.stack([p["x"], p["y"], p["z"]], axis=1).astype(np.float64)


def evaluate(scene: dict, model: dict, cycles: int, dt: float | None = None, K: int = K_DEFAULT,
             factor: int = FACTOR_DEFAULT, threshold: float = 1e30, verify_every: int = 0,
             warm_steps: int = 0) -> dict:
    """Surrogate rollout vs the plain solver over the same simulated time at
    the same fixed dt schedule, both after `warm_steps` plain steps
    (model=None: the network-free baseline a = a_last).
    Position error = mean |x_net - x_plain| / mean |x_plain - x_start|; wall
    times are device-synchronised."""
    if dt is None:
        dt = fixed_dt(scene)
    nplain = 1 + cycles * (K * factor + 1)
    # plain reference
    p = HipKernel.from_scene(scene)
    for _ in range(warm_steps):
        p.step_async(dt)
    p.sync()
    x_init = _positions(p)
    p.step(dt)
    p.sync()
    t0 = time.perf_counter()
    for _ in range(nplain - 1):
        p.step_async(dt)
    p.sync()
    t_plain = time.perf_counter() - t0
    x_plain = _positions(p)
    p.close()
    # surrogate
    s = HipKernel.from_scene(scene)
    s.surrogate_set_model(model)
    for _ in range(warm_steps):
        s.step_async(dt)
    s.step(dt)
    s.sync()
    t0 = time.perf_counter()
    st = s.surrogate_run(cycles, K, dt, factor, threshold, verify_every)
    s.sync()
    t_sur = time.perf_counter() - t0
    x_sur = _positions(s)
    s.close()
    disp = float(np.mean(np.linalg.norm(x_plain - x_init, axis=1)))
    err = float(np.mean(np.linalg.norm(x_sur - x_plain, axis=1)))
    return {"dt": dt, "sim_time": nplain * dt, "plain_steps": nplain, "warm_steps": warm_steps,
            "plain_s": t_plain, "surrogate_s": t_sur,
            "speedup": t_plain / t_sur if t_sur > 0 else None, "pos_err": err, "disp": disp,
            "rel_err": err / disp if disp > 0 else None, "stats": st}


__all__ = ["fixed_dt", "collect_rollout", "train", "evaluate", "MLP_FIELDS"]
