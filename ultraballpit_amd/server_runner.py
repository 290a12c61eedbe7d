"""The server's simulation runner over the backend interface.

``SimulationRunner`` restates the stepping part of the server's runner
(backend/crates/server/src/runner.rs:27-232) on a kernel with the trait's
calls, the way the server would drive this engine once its kernel field goes
through ``create_kernel`` (runner.rs:29, 82).  Per step: the CFL dt of the
current state (compute_timestep, kernel/src/sph.rs:591-628), the step,
auto-pause on a density variation above 100 or on any non-finite position,
velocity, density or pressure, the counters, and the force on the geometry
recorded every step.  The instability test and the force run on the device
(``state_finite``, ``surface_forces``) and the viewer frame is built there
(``build_frame``): no step reads all particles back.
"""
from __future__ import annotations

import enum
from dataclasses import dataclass

import numpy as np


class SimStatus(enum.Enum):
    """server/src/ws.rs SimStatus (Created, Running, Paused, Stopped)."""
    CREATED = 0
    RUNNING = 1
    PAUSED = 2
    STOPPED = 3


@dataclass
class ForceRecord:
    """runner.rs:15-24."""
    timestep: int
    sim_time: float
    net_force: np.ndarray
    net_moment: np.ndarray


class SimulationRunner:
    """runner.rs:27-232 without the locks: ``kernel`` is a HipKernel (or any
    object with dt, step, error_metrics, state_finite, surface_forces,
    set_force_sdf and build_frame)."""

    DENSITY_PAUSE = 100.0  # runner.rs:180

    def __init__(self, kernel, h: float, speed_of_sound: float, cfl_number: float, force_sdf: dict | None = None):
        f32 = np.float32
        self.kernel = kernel
        self.status = SimStatus.CREATED
        self.timestep = 0
        self.sim_time = 0.0
        self.dt = f32(f32(cfl_number) * f32(h) / f32(speed_of_sound))  # runner.rs:98
        n = kernel.particle_count()
        self.subsample_count = int(max(f32(n) * f32(0.05), f32(1.0)))  # runner.rs:95
        self.force_history: list[ForceRecord] = []
        self.has_force_geometry = force_sdf is not None
        if self.has_force_geometry:
            kernel.set_force_sdf(force_sdf)

    def start(self):
        self.status = SimStatus.RUNNING

    def pause(self):
        self.status = SimStatus.PAUSED

    def resume(self):
        self.status = SimStatus.RUNNING

    def stop(self):
        self.status = SimStatus.STOPPED

    def step(self):
        """One timestep if running (runner.rs:150-232)."""
        if self.status != SimStatus.RUNNING:
            return
        dt = np.float32(self.kernel.dt())
        self.dt = dt
        self.kernel.step(float(dt))
        if np.float32(self.kernel.error_metrics().max_density_variation) > np.float32(self.DENSITY_PAUSE):
            self.status = SimStatus.PAUSED
            return
        if not self.kernel.state_finite():
            self.status = SimStatus.PAUSED
            return
        self.timestep += 1
        self.sim_time += float(dt)
        if self.has_force_geometry:
            # compute_surface_forces (orchestrator/src/force.rs:39-129): the
            # pressure force with the reference's cell^2 weighting, FP64 sums
            f = self.kernel.surface_forces(mu=0.0, mode=0)
            self.force_history.append(ForceRecord(self.timestep, self.sim_time, np.asarray(f["net_force"]),
                                                  np.asarray(f["net_moment"])))

    def build_frame(self) -> bytes:
        """The viewer's Frame message (ws.rs:232-294) of the current state."""
        return self.kernel.build_frame(self.subsample_count, self.timestep, self.sim_time)


__all__ = ["SimulationRunner", "SimStatus", "ForceRecord"]
