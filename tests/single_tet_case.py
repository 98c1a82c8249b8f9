"""The reference's single-tet regression case (test/regression_tests/single_tet/fem)
as data plus the time loop of ExplicitSolver::Solve (src/Solvers/Explicit.cpp:308-452)
with the reference's DirectTimeStepper clamp (include/TimeStepper.h:97-128) and output
scheduler (include/Scheduler.h:37-63).  `step(dt)` advances whichever solver is driven
(the oracle or the GPU library) by one step; `fields()` returns (u, v, a, f)."""
import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
FIELDS = {"displacement": 0, "velocity": 1, "acceleration": 2, "force": 3}


def load():
    with open(os.path.join(HERE, "golden", "single_tet_fem_maxima.json")) as fh:
        g = json.load(fh)
    g["X"] = np.array(g["nodes"], dtype=np.float64)
    g["conn"] = np.array(g["connectivity"], dtype=np.int32)
    g["fixed_mask"] = np.array(g["fixed"], dtype=np.uint8)
    g["v0"] = np.tile(np.array(g["initial_velocity"]), (len(g["nodes"]), 1))
    return g


class OutputScheduler:
    """An event every `increment` from `start`; events passed in one call count once."""

    def __init__(self, start, end, increment):
        self.next, self.end, self.increment = start, end, increment

    def at_next_event(self, time):
        if time + 1.0e-15 >= self.next:
            self.next += self.increment
            self.at_next_event(time)
            if self.next > self.end:
                self.next = np.inf
            return True
        return False


def run(case, step, fields, get_increment=None):
    """Outputs [(time, u, v, a, f)] of the reference's time loop; the first increment is
    requested before the loop (and the caller's initial force), later ones per step."""
    t_end, dt0 = case["time_end"], case["time_increment"]
    o = case["output"]
    sched = OutputScheduler(o["time_start"], o["time_end"], o["time_increment"])

    def increment(time, n):
        if get_increment is not None:
            return get_increment(time, n)
        if time + dt0 <= t_end:
            return dt0
        return 0.0 if time > t_end else t_end - time

    time, n = 0.0, 0
    increment(time, n)
    outs = []
    if sched.at_next_event(time):
        outs.append((time,) + tuple(np.array(x, copy=True) for x in fields()))
    while not time >= t_end:
        dt = increment(time, n)
        step(dt)
        time += dt
        n += 1
        if sched.at_next_event(time):
            outs.append((time,) + tuple(np.array(x, copy=True) for x in fields()))
    return outs


def maxima(outs):
    """{name: (max |value|, 1-offset output, 1-offset node)} like exodiff's summary."""
    res = {}
    for name, k in FIELDS.items():
        arr = np.stack([o[1 + k] for o in outs])  # outputs x nodes x 3
        for d, c in enumerate("xyz"):
            a = np.abs(arr[:, :, d])
            i, j = np.unravel_index(np.argmax(a), a.shape)
            res[f"{name}_{c}"] = (float(a[i, j]), int(i) + 1, int(j) + 1)
    return res


class PowerMethodStepper:
    """scale x the power-method increment, recomputed every `interval` steps from step
    0 and clamped to end at time_end (include/TimeStepper.h:138-216).  `power(previous)`
    returns the full increment given the previous full one (0 on the first call)."""

    def __init__(self, power, scale, interval, time_end):
        self.power, self.scale, self.interval, self.time_end = power, scale, interval, time_end
        self.next_update, self.increment, self.updates = 0, 0.0, []

    def __call__(self, time, n):
        if n == self.next_update:
            full = self.power(self.increment / self.scale)
            self.increment = full * self.scale
            self.next_update += self.interval
            self.updates.append((n, full))
        if time + self.increment <= self.time_end:
            return self.increment
        return 0.0 if time > self.time_end else self.time_end - time
