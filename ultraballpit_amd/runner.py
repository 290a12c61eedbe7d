"""The orchestrator's runner loop over the backend interface.

``run_simulation_loop`` restates the stepping of the reference's
``run_simulation_loop`` (backend/crates/orchestrator/src/runner.rs:157-277)
for any kernel with the trait's calls (``step``, ``save_checkpoint``,
``restore_checkpoint``) plus the CFL dt of its current state
(``compute_timestep``, kernel/src/sph.rs:591-628) and the density variation
of ``error_metrics`` (kernel/src/lib.rs:341-377): the host-dt form, one trait
call at a time, as the reference runs it.  ``HipKernel.run_optimistic`` is
the same loop with the dt kept on the device.
"""
from __future__ import annotations

import numpy as np

OPTIMISTIC_ALPHA = 1.5    # runner.rs:189
DENSITY_VAR_LIMIT = 0.03  # runner.rs:190
COOLDOWN_STEPS = 10       # runner.rs:191


def run_simulation_loop(kernel, compute_timestep, density_variation, max_timesteps=None, max_time=None,
                        alpha=OPTIMISTIC_ALPHA, limit=DENSITY_VAR_LIMIT, cooldown=COOLDOWN_STEPS) -> dict:
    """Step ``kernel`` until ``max_timesteps`` steps or ``max_time`` seconds of
    simulated time: outside a cooldown, save a checkpoint, step at
    alpha * dt_safe, and on a density variation above ``limit`` restore, step
    at dt_safe and cool down for ``cooldown`` steps."""
    if max_timesteps is None and max_time is None:
        raise ValueError("give max_timesteps or max_time")
    f32 = np.float32
    sim_time = 0.0
    steps = accepted = rejected = 0
    cool = 0
    while True:
        dt_safe = f32(compute_timestep())
        if cool > 0:
            cool -= 1
            kernel.step(float(dt_safe))
            dt_used = dt_safe
        elif kernel.save_checkpoint():
            dt_try = f32(dt_safe * f32(alpha))
            kernel.step(float(dt_try))
            if f32(density_variation()) > f32(limit):
                kernel.restore_checkpoint()
                kernel.step(float(dt_safe))
                dt_used = dt_safe
                rejected += 1
                cool = cooldown
            else:
                dt_used = dt_try
                accepted += 1
        else:
            kernel.step(float(dt_safe))
            dt_used = dt_safe
        sim_time += float(dt_used)
        steps += 1
        if max_timesteps is not None and steps >= max_timesteps:
            break
        if max_time is not None and sim_time >= max_time:
            break
    return {"steps": steps, "accepted": accepted, "rejected": rejected, "sim_time": sim_time}


__all__ = ["run_simulation_loop", "OPTIMISTIC_ALPHA", "DENSITY_VAR_LIMIT", "COOLDOWN_STEPS"]
