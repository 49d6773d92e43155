"""Force-coefficient history: Cd/Cl samples, window means, the shedding frequency (Strouhal
number) and the convergence tests. Python face of the library's C++ implementation
(lbm_history.cpp), mirroring src/aero/lbm/include/orrery/aero/lbm/force_history.h:11-96."""
from __future__ import annotations

import ctypes
import dataclasses

import numpy as np

from . import abi


@dataclasses.dataclass
class Coefficients:
    drag: float = 0.0
    lift: float = 0.0


@dataclasses.dataclass
class Oscillation:
    period: float
    frequency: float
    cycles: int
    amplitude: float
    period_spread: float


@dataclasses.dataclass
class SettledCriteria:
    """force_history.h:65-75"""
    steady_tolerance: float = 1e-3
    max_period_spread: float = 0.05
    stationarity: float = 0.02


def force_coefficients(force_x: float, force_y: float, speed: float, length: float) -> Coefficients:
    d = ctypes.c_double()
    l = ctypes.c_double()
    abi.load().lbm_force_coefficients(force_x, force_y, speed, length, ctypes.byref(d), ctypes.byref(l))
    return Coefficients(d.value, l.value)


def _arr(v):
    return np.ascontiguousarray(v, dtype=np.float64)


def mean_since(time, values, start: float) -> float:
    t, v = _arr(time), _arr(values)
    n = min(len(t), len(v))
    return abi.load().lbm_mean_since(abi.ptr(t, ctypes.c_double), abi.ptr(v, ctypes.c_double), n, start)


def _osc(o: abi.COscillation) -> Oscillation:
    return Oscillation(o.period, o.frequency, int(o.cycles), o.amplitude, o.period_spread)


def dominant_oscillation(time, values, start: float) -> Oscillation | None:
    t, v = _arr(time), _arr(values)
    n = min(len(t), len(v))
    o = abi.COscillation()
    ok = abi.load().lbm_dominant_oscillation(abi.ptr(t, ctypes.c_double), abi.ptr(v, ctypes.c_double), n, start,
                                             ctypes.byref(o))
    return _osc(o) if ok else None


def steady_since(time, drag, lift, start: float, criteria: SettledCriteria = SettledCriteria()) -> bool:
    t, d, l = _arr(time), _arr(drag), _arr(lift)
    n = min(len(t), len(d), len(l))
    return bool(abi.load().lbm_steady_since(abi.ptr(t, ctypes.c_double), abi.ptr(d, ctypes.c_double),
                                            abi.ptr(l, ctypes.c_double), n, start, criteria.steady_tolerance))


def settled_oscillation(time, drag, lift, start: float,
                        criteria: SettledCriteria = SettledCriteria()) -> Oscillation | None:
    t, d, l = _arr(time), _arr(drag), _arr(lift)
    n = min(len(t), len(d), len(l))
    o = abi.COscillation()
    ok = abi.load().lbm_settled_oscillation(abi.ptr(t, ctypes.c_double), abi.ptr(d, ctypes.c_double),
                                            abi.ptr(l, ctypes.c_double), n, start, criteria.max_period_spread,
                                            criteria.stationarity, ctypes.byref(o))
    return _osc(o) if ok else None


class CoefficientHistory:
    """Samples of (time, Cd, Cl) with a moving average over the last `window` samples
    (force_history.h:22-42)."""

    def __init__(self, window: int):
        self._window = max(1, int(window))
        self.time: list[float] = []
        self.drag: list[float] = []
        self.lift: list[float] = []

    def push(self, time: float, sample: Coefficients) -> None:
        self.time.append(float(time))
        self.drag.append(float(sample.drag))
        self.lift.append(float(sample.lift))

    def size(self) -> int:
        return len(self.time)

    def window(self) -> int:
        return self._window

    def moving_average(self) -> Coefficients:
        n = min(self._window, len(self.time))
        if n == 0:
            return Coefficients()
        d = 0.0
        l = 0.0
        for k in range(len(self.time) - n, len(self.time)):
            d += self.drag[k]
            l += self.lift[k]
        return Coefficients(d / n, l / n)


def forces_settled(time, drag, lift, side, start: float, criteria: SettledCriteria = SettledCriteria()) -> bool:
    """forces_settled_3d (src/aero/lbm3d/lbm_aero_solver3d.cpp:60-83): every force component is
    steady or settled periodic shedding over the samples at time >= start."""
    t, d, l, s = _arr(time), _arr(drag), _arr(lift), _arr(side)
    n = min(len(t), len(d), len(l), len(s))
    return bool(abi.load().lbm_forces_settled(abi.ptr(t, ctypes.c_double), abi.ptr(d, ctypes.c_double),
                                              abi.ptr(l, ctypes.c_double), abi.ptr(s, ctypes.c_double), n, start,
                                              criteria.steady_tolerance, criteria.max_period_spread,
                                              criteria.stationarity))


@dataclasses.dataclass
class Verdict:
    converged: bool
    kind: str          # "steady", "settled" or "" (not converged)
    drag: float        # mean drag coefficient over the judged samples
    lift: float
    start: float       # first sample time the means cover


def judge_solve(time, drag, lift, side, min_time: float, window: int = 50, tolerance: float = 1e-4,
                criteria: SettledCriteria = SettledCriteria(), stationarity: float = 5e-3) -> Verdict:
    """Convergence of a running solve from its coefficient history (one sample per entry), by
    the first of three tests that holds, none judged before `min_time`:

    - steady: the mean drag of the last `window` samples within `tolerance` of the mean of the
      `window` before; the means cover the last window;
    - settled: forces_settled over the last half of the run (the reference's adapter test,
      src/aero/lbm3d/lbm_aero_solver3d.cpp:564-584): a wake shedding at a locked period with
      stationary amplitude and mean; the means cover the last half;
    - stationary (no reference counterpart; for irregular wakes that neither test accepts):
      the last half of the run split into two halves of equal sample count, whose mean drag
      agree within `stationarity` x |mean drag| and whose mean lift agree within
      `stationarity` x max(|mean drag|, |mean lift|); the means cover the last half."""
    t, d, l, s = _arr(time), _arr(drag), _arr(lift), _arr(side)
    n = min(len(t), len(d), len(l), len(s))
    if n == 0 or t[n - 1] < min_time:
        return Verdict(False, "", float("nan"), float("nan"), float("nan"))
    tail = max(0, n - window)
    if n >= 2 * window and abs(d[tail:n].mean() - d[n - 2 * window:tail].mean()) < tolerance:
        return Verdict(True, "steady", float(d[tail:n].mean()), float(l[tail:n].mean()), float(t[tail]))
    start = 0.5 * float(t[n - 1])
    if forces_settled(t[:n], d[:n], l[:n], s[:n], start, criteria):
        return Verdict(True, "settled", mean_since(t[:n], d[:n], start), mean_since(t[:n], l[:n], start), start)
    half = n // 2
    if n - half >= 2 * window:
        mid = half + (n - half) // 2
        cd, cl = float(d[half:n].mean()), float(l[half:n].mean())
        scale = max(abs(cd), abs(cl))
        if abs(d[half:mid].mean() - d[mid:n].mean()) <= stationarity * abs(cd) and \
                abs(l[half:mid].mean() - l[mid:n].mean()) <= stationarity * scale:
            return Verdict(True, "stationary", cd, cl, float(t[half]))
    return Verdict(False, "", float(d[tail:n].mean()), float(l[tail:n].mean()), float(t[tail]))
