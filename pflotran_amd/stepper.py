"""The transport time-step controller of the reference, restated for the host
loop around the device step:

  TimestepperBaseSetTargetTime  timestepper_base.F90:320-521  the next step:
      the current size, cut to land on a waypoint that forces a match
      (output times, the final time) when the step would reach it within
      the tolerance 0.1 of a step, never above the waypoint's maximum; a
      step shortened by a waypoint is reverted to its previous size after it
  TimestepperBEStepDT           timestepper_BE.F90:344-399    a step that is
      cut restarts from the old state at half the size (the device does the
      retry: pfx_gi_solver_timestep), and its cut flag holds the growth
  TimestepperBEUpdateDT         timestepper_BE.F90:200-246    no growth for
      5 steps after a cut (constant_time_step_threshold, base.F90:146)
  PMRTUpdateTimestep            pm_rt.F90:620-647             growth from the
      step's Newton count: tfac(n) dt up to TS_ACCELERATION iterations, half
      the step beyond, at most twice the step and the maximum

Waypoints are built as the reference builds them: the TIME card's maximum
step from time 0 (WaypointListFillIn carries it forward), the final time,
OUTPUT TIMES, and OUTPUT PERIODIC TIME by repeated addition of the increment
up to the final time plus 1e-10 of it (init_common.F90:695-725).  The outer
synchronisation time is the final time (a subsurface run without sync
waypoints, waypoint.F90:767-810).
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import List, Optional, Sequence, Tuple

# timestepper_BE.F90:124-133
DEFAULT_TFAC = (2.0, 2.0, 2.0, 2.0, 2.0, 1.8, 1.6, 1.4, 1.2, 1.0, 1.0, 1.0, 1.0)


@dataclass
class Waypoint:
    time: float
    dt_max: float
    force: bool = False   # output, condition update or the final time: steps land on it


def build_waypoints(final: float, dt_max: float, output_times: Sequence[float] = (),
                    output_period: Optional[float] = None) -> List[Waypoint]:
    """Sorted waypoints with equal times merged (WaypointInsertInList)."""
    times = {0.0: False, final: True}
    for t in output_times:
        times[float(t)] = True
    if output_period and output_period > 0.0:
        limit = final + final * 1e-10
        t = 0.0
        while True:
            t += output_period
            if t > limit:
                break
            times[t] = True
    wps = []
    for t in sorted(times):
        if wps and abs(wps[-1].time - t) <= 0.0:
            wps[-1].force = wps[-1].force or times[t]
            continue
        wps.append(Waypoint(t, dt_max, times[t]))
    return wps


@dataclass
class TimeStepper:
    final: float
    dt_init: float
    dt_max: float
    waypoints: List[Waypoint]
    iaccel: int = 5
    tfac: Tuple[float, ...] = DEFAULT_TFAC
    dt_min: float = 1e-20                 # timestepper_base.F90:159
    tolerance: float = 0.1                # time_step_tolerance, :153
    constant_time_step_threshold: int = 5
    # state
    time: float = 0.0
    dt: float = 0.0
    prev_dt: float = 0.0
    target_time: float = 0.0
    cur: int = 0
    prev_wp: int = 0
    revert_dt: bool = False
    num_contig_revert_due_to_sync: int = 0
    cut_flag: bool = False
    num_constant_time_steps: int = 0
    steps: int = 0
    done: bool = False
    history: List[Tuple[float, int, int]] = field(default_factory=list)   # (dt, Newton its, cuts)

    def __post_init__(self):
        self.dt = self.dt_init

    @property
    def sync_time(self) -> float:
        return self.final

    def set_target_time(self) -> float:
        """TimestepperBaseSetTargetTime: the size of the next step."""
        if self.cut_flag:
            self.cut_flag = False
            self.cur = self.prev_wp
        elif self.revert_dt and self.num_contig_revert_due_to_sync < 2:
            self.dt = min(self.prev_dt, self.dt_max)
        self.revert_dt = False
        revert_wp = revert_sync = False
        dt = self.dt
        self.prev_dt = dt
        cur = self.cur
        self.prev_wp = cur
        W = self.waypoints
        sync = self.sync_time
        target = self.time + dt
        while True:
            wp = W[cur]
            dt_max = wp.dt_max
            dt = min(dt, dt_max)
            target = self.time + dt
            force = False
            if cur + 1 < len(W):
                nxt = W[cur + 1]
                if dt_max > nxt.dt_max and dt > nxt.dt_max and target > wp.time:
                    if self.time + nxt.dt_max < wp.time:
                        force = True
                    else:
                        dt = nxt.dt_max
                        target = self.time + dt
            force = wp.force or force
            eq_wp = target + self.tolerance * dt >= wp.time
            eq_sync = target + self.tolerance * dt >= sync
            if eq_sync and sync < wp.time:
                eq_wp = False
            if eq_sync or (eq_wp and force):
                max_time = min(sync, wp.time) if force else sync
                target = target - dt
                dt = max_time - target
                if dt > dt_max and abs(dt - dt_max) > 1.0:   # 1 s against cancellation
                    dt = dt_max
                    target = target + dt
                else:
                    target = max_time
                    if eq_wp and force:
                        revert_wp = True
                    if eq_sync and dt <= 0.5 * self.prev_dt:
                        revert_sync = True
                    if max_time >= wp.time:
                        cur += 1
                break
            elif target > wp.time:
                cur += 1
            else:
                break
        if revert_sync or revert_wp:
            self.revert_dt = True
            if revert_sync:
                self.num_contig_revert_due_to_sync += 1
        else:
            self.num_contig_revert_due_to_sync = 0
        if cur < len(W):
            self.dt_max = W[cur].dt_max
        else:
            self.done = True            # stop after this step (TS_STOP_END_SIMULATION)
        self.dt = dt
        self.target_time = target
        self.cur = cur
        return dt

    def accept(self, dt_taken: float, newton_iterations: int, cuts: int) -> None:
        """The step converged at dt_taken after `cuts` cuts: advance time and
        update the size (TimestepperBEUpdateDT, PMRTUpdateTimestep)."""
        if cuts > 0:
            # StepDT moved the target time back with each cut
            self.target_time = self.target_time - self.dt + dt_taken
            self.dt = dt_taken
            self.cut_flag = True
            self.done = False     # a cut step no longer reaches the waypoint
        self.time = self.target_time
        self.steps += 1
        self.history.append((dt_taken, int(newton_iterations), int(cuts)))
        update = True
        if self.cut_flag:
            self.num_constant_time_steps = 1
        elif self.num_constant_time_steps > 0:
            self.num_constant_time_steps += 1
        if self.num_constant_time_steps > self.constant_time_step_threshold:
            self.num_constant_time_steps = 0
        elif self.num_constant_time_steps > 0:
            update = False
        if update and self.iaccel != 0:
            self.dt = self.grow(self.dt, newton_iterations)

    def grow(self, dt: float, n: int) -> float:
        """PMRTUpdateTimestep's original rule (volume-fraction governor off);
        a count below 1 is taken as 1 (the reference would index tfac(0))."""
        n = max(int(n), 1)
        if n <= self.iaccel:
            dtt = self.tfac[n - 1] * dt if n <= len(self.tfac) else 0.5 * dt
        else:
            dtt = 0.5 * dt
        dtt = min(dtt, 2.0 * dt, self.dt_max)
        return max(dtt, self.dt_min)
