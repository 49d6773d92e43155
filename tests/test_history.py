"""Force-coefficient history: the reference's own tests (tests/aero/lbm/test_force_history.cpp)
restated, and bitwise agreement with the reference's force_history.cpp (oracle/_ref)."""
import ctypes
import math

import numpy as np
import pytest

import refcheck
from orrery_amd.history import (CoefficientHistory, Coefficients, dominant_oscillation, force_coefficients,
                                mean_since, settled_oscillation, steady_since)


def test_force_coefficients_use_the_dynamic_pressure():
    c = force_coefficients(0.0135, -0.002, 0.05, 20.0)
    assert c.drag == pytest.approx(0.0135 / (0.5 * 0.0025 * 20.0))
    assert c.lift == pytest.approx(-0.002 / (0.5 * 0.0025 * 20.0))


def test_moving_average_covers_the_last_window():
    h = CoefficientHistory(3)
    assert h.moving_average().drag == 0.0
    h.push(0.0, Coefficients(1.0, 0.0))
    assert h.moving_average().drag == 1.0
    h.push(1.0, Coefficients(2.0, 1.0))
    h.push(2.0, Coefficients(3.0, -1.0))
    h.push(3.0, Coefficients(7.0, 3.0))
    assert h.size() == 4
    assert h.moving_average().drag == pytest.approx(4.0)
    assert h.moving_average().lift == pytest.approx(1.0)
    assert CoefficientHistory(0).window() == 1


def shedding_like():
    f = 1.0 / 2425.0
    t = np.arange(0, 40000, 10, dtype=np.float64)
    v = 1.3 + 0.3 * np.sin(2.0 * math.pi * f * t + 0.4) + np.exp(-t / 500.0)
    return f, t, v


def test_dominant_oscillation_of_a_shedding_like_signal():
    f, t, v = shedding_like()
    o = dominant_oscillation(t, v, 10000.0)
    assert o is not None
    assert o.frequency == pytest.approx(f, rel=2e-4)
    assert o.cycles >= 10
    assert o.amplitude == pytest.approx(0.3, rel=1e-3)
    assert o.period_spread < 0.01
    assert mean_since(t, v, 10000.0) == pytest.approx(1.3, rel=1e-2)
    assert dominant_oscillation(t, np.ones_like(t), 0.0) is None


def signal(amplitude, drag, phase):
    t = np.arange(0, 20000, 10, dtype=np.float64)
    return t, np.array([drag(x) for x in t]), np.array([amplitude(x) * math.sin(2 * math.pi * phase(x)) for x in t])


def settled_cases():
    constant = lambda x: 0.4  # noqa: E731
    mean_drag = lambda x: 1.4 + 0.01 * math.sin(4.0 * math.pi * x / 1000.0)  # noqa: E731
    regular = lambda x: x / 1000.0  # noqa: E731
    return {
        "locked": signal(constant, mean_drag, regular),
        "growing": signal(lambda x: 0.4 * math.exp((x - 20000.0) / 10500.0), mean_drag, regular),
        "uneven": signal(constant, mean_drag, lambda x: x / 1000.0 + 0.15 * math.sin(2 * math.pi * x / 2700.0)),
        "drifting": signal(constant, lambda x: 1.5 - 1.5e-5 * x, regular),
    }


def test_settled_shedding_needs_consistent_periods_and_stationarity():
    cases = settled_cases()
    t, d, l = cases["locked"]
    o = settled_oscillation(t, d, l, 10000.0)
    assert o is not None and o.period == pytest.approx(1000.0, rel=1e-3)
    for name in ("growing", "uneven", "drifting"):
        t, d, l = cases[name]
        assert dominant_oscillation(t, l, 10000.0) is not None
        assert settled_oscillation(t, d, l, 10000.0) is None, name
    t, d, l = cases["locked"]
    assert settled_oscillation(t, d, l, 18500.0) is None


def test_steady_history():
    t = np.arange(2000, dtype=np.float64)
    drag = 1.2 + 0.5 * np.exp(-t / 100.0)
    lift = 1e-12 * np.sin(t)
    assert steady_since(t, drag, lift, 1500.0)
    assert not steady_since(t, drag, lift, 500.0)
    assert not steady_since(t, drag, lift, 1999.5)
    assert not steady_since(t, drag, 0.01 * np.sin(t / 50.0), 1500.0)


@pytest.mark.skipif(not refcheck.available(), reason="oracle/_ref not built (reference tree absent)")
def test_history_matches_the_reference_bitwise():
    ref = refcheck.load()
    P = ctypes.POINTER(ctypes.c_double)
    for name in ("ref_dominant_oscillation", "ref_settled_oscillation", "ref_steady_since", "ref_mean_since"):
        getattr(ref, name).restype = ctypes.c_double if name == "ref_mean_since" else ctypes.c_int
    out = np.zeros(5)
    p = refcheck.ptr
    _, t, v = shedding_like()
    for start in (0.0, 5000.0, 10000.0, 39000.0):
        ok = ref.ref_dominant_oscillation(p(t), p(v), ctypes.c_size_t(len(t)), ctypes.c_double(start), p(out))
        o = dominant_oscillation(t, v, start)
        assert bool(ok) == (o is not None)
        if ok:
            assert [o.period, o.frequency, float(o.cycles), o.amplitude, o.period_spread] == out.tolist()
        assert ref.ref_mean_since(p(t), p(v), ctypes.c_size_t(len(t)), ctypes.c_double(start)) == mean_since(t, v, start)
    for name, (t, d, l) in settled_cases().items():
        for start in (0.0, 10000.0, 18500.0):
            ok = ref.ref_settled_oscillation(p(t), p(d), p(l), ctypes.c_size_t(len(t)), ctypes.c_double(start), p(out))
            o = settled_oscillation(t, d, l, start)
            assert bool(ok) == (o is not None), (name, start)
            if ok:
                assert [o.period, o.frequency, float(o.cycles), o.amplitude, o.period_spread] == out.tolist()
            assert bool(ref.ref_steady_since(p(t), p(d), p(l), ctypes.c_size_t(len(t)), ctypes.c_double(start))) == \
                steady_since(t, d, l, start)


def random_histories(seed):
    """Noisy, drifting, locked, steady and chirped signals (40-400 samples), with starts inside
    and outside the history."""
    rng = np.random.default_rng(seed)
    for case in range(60):
        n = int(rng.integers(40, 400))
        t = np.cumsum(rng.choice([10.0, 10.0, 10.0, 5.0], size=n))
        kind = case % 5
        period = float(rng.uniform(60.0, 600.0))
        phase = 2 * math.pi * t / period
        if kind == 0:
            lift = 0.3 * np.sin(phase) + 1e-3 * rng.standard_normal(n)
        elif kind == 1:
            lift = 0.3 * np.sin(phase * (1 + t / t[-1])) + 0.05 * rng.standard_normal(n)
        elif kind == 2:
            lift = 1e-6 * rng.standard_normal(n)
        elif kind == 3:
            lift = np.round(0.3 * np.sin(phase), 2)  # ties with the mean
        else:
            lift = np.cumsum(rng.standard_normal(n))
        wobble = 1e-5 if kind == 2 else 0.01
        drag = 1.2 + wobble * np.sin(2 * phase) + float(rng.uniform(0, 1e-3)) * wobble * rng.standard_normal(n)
        side = (0.2 if case % 3 == 0 else 1e-7) * np.sin(phase + 1.0)
        for start in (0.0, float(t[n // 2]), float(rng.uniform(t[0], t[-1])), float(t[-1]) + 1.0):
            yield t, drag, lift, side, start


@pytest.mark.skipif(not refcheck.available(), reason="oracle/_ref not built (reference tree absent)")
def test_history_matches_the_reference_on_random_signals():
    ref = refcheck.load()
    for name in ("ref_dominant_oscillation", "ref_settled_oscillation", "ref_steady_since"):
        getattr(ref, name).restype = ctypes.c_int
    ref.ref_mean_since.restype = ctypes.c_double
    p = refcheck.ptr
    out = np.zeros(5)
    seen = {"dominant": 0, "settled": 0, "steady": 0}
    for t, d, l, s, start in random_histories(11):
        n = ctypes.c_size_t(len(t))
        c = ctypes.c_double(start)
        assert ref.ref_mean_since(p(t), p(l), n, c) == mean_since(t, l, start)
        ok = ref.ref_dominant_oscillation(p(t), p(l), n, c, p(out))
        o = dominant_oscillation(t, l, start)
        assert bool(ok) == (o is not None)
        if ok:
            seen["dominant"] += 1
            assert [o.period, o.frequency, float(o.cycles), o.amplitude, o.period_spread] == out.tolist()
        ok = ref.ref_settled_oscillation(p(t), p(d), p(l), n, c, p(out))
        o = settled_oscillation(t, d, l, start)
        assert bool(ok) == (o is not None)
        if ok:
            seen["settled"] += 1
            assert [o.period, o.frequency, float(o.cycles), o.amplitude, o.period_spread] == out.tolist()
        ok = bool(ref.ref_steady_since(p(t), p(d), p(l), n, c))
        seen["steady"] += ok
        assert ok == steady_since(t, d, l, start)
    # every branch was exercised
    assert min(seen.values()) > 0, seen


def forces_settled_from_reference(ref, t, d, l, s, start):
    """forces_settled_3d (lbm_aero_solver3d.cpp:60-83) composed from the reference's own
    mean_since and settled_oscillation (oracle/_ref): per component, two or more samples at
    time >= start within 1e-3 x |mean drag|, or settled shedding paired with the drag."""
    p = refcheck.ptr
    n = ctypes.c_size_t(len(t))
    bound = 1e-3 * abs(ref.ref_mean_since(p(t), p(d), n, ctypes.c_double(start)))
    out = np.zeros(5)
    for v in (d, l, s):
        kept = v[t >= start]
        if len(kept) >= 2 and kept.max() - kept.min() <= bound:
            continue
        if not ref.ref_settled_oscillation(p(t), p(d), p(v), n, ctypes.c_double(start), p(out)):
            return False
    return True


@pytest.mark.skipif(not refcheck.available(), reason="oracle/_ref not built (reference tree absent)")
def test_forces_settled_matches_the_reference_composition():
    from orrery_amd.history import forces_settled
    ref = refcheck.load()
    ref.ref_settled_oscillation.restype = ctypes.c_int
    ref.ref_mean_since.restype = ctypes.c_double
    verdicts = set()
    for t, d, l, s, start in random_histories(5):
        want = forces_settled_from_reference(ref, t, d, l, s, start)
        verdicts.add(want)
        assert forces_settled(t, d, l, s, start) == want
    assert verdicts == {True, False}


def test_forces_settled_accepts_settled_shedding():
    """A wake shedding at a locked period with a steady side force counts as converged, as the
    reference's forces_settled_3d has it; a growing or drifting one does not."""
    from orrery_amd.history import forces_settled
    cases = settled_cases()
    t, d, l = cases["locked"]
    side = np.full_like(t, 1e-9)
    assert forces_settled(t, d, l, side, 10000.0)
    assert not steady_since(t, d, l, 10000.0)
    for name in ("growing", "uneven", "drifting"):
        t, d, l = cases[name]
        assert not forces_settled(t, d, l, side, 10000.0), name
    # shedding in the side force too (a wake rotating its plane) is still settled
    t, d, l = cases["locked"]
    assert forces_settled(t, d, l, l[::-1].copy(), 10000.0)
    # steady flow: all three within 1e-3 of the drag
    steady = np.full_like(t, 1.4)
    assert forces_settled(t, steady, 1e-6 * np.ones_like(t), side, 10000.0)


def test_judge_solve_kinds():
    """The three convergence kinds of a running solve, on synthetic histories sampled every
    10 steps: a decaying drag (steady), locked shedding (settled), an irregular wake about a
    settled mean (stationary), and a drifting one (not converged)."""
    from orrery_amd.history import judge_solve
    t = np.arange(10.0, 20010.0, 10.0)
    zero = np.zeros_like(t)
    steady = 1.2 + 0.5 * np.exp(-t / 800.0)
    v = judge_solve(t, steady, 1e-9 * np.ones_like(t), zero, 4000.0)
    assert v.converged and v.kind == "steady" and v.drag == pytest.approx(1.2, rel=1e-4)
    assert not judge_solve(t[t < 3000], steady[t < 3000], zero[t < 3000], zero[t < 3000], 4000.0).converged
    d = 1.4 + 0.01 * np.sin(2 * np.pi * t / 350.0)  # no whole number of periods in a 500-step window
    l = 0.4 * np.sin(2 * np.pi * t / 700.0)
    v = judge_solve(t, d, l, zero, 4000.0)
    assert v.converged and v.kind == "settled" and v.drag == pytest.approx(1.4, rel=1e-3)
    rng = np.random.default_rng(1)
    noise = np.convolve(rng.standard_normal(len(t) + 20), np.ones(20) / 20, mode="valid")[:len(t)]
    irregular = 0.9 + 0.02 * noise + 0.01 * np.sin(t / 37.0) * np.sin(t / 151.0)
    v = judge_solve(t, irregular, 0.3 + 0.02 * noise[::-1], zero, 4000.0)
    assert v.converged and v.kind == "stationary"
    assert v.drag == pytest.approx(irregular[len(t) // 2:].mean(), rel=1e-12)
    drifting = irregular + 2e-6 * t
    assert not judge_solve(t, drifting, 0.3 + 0.02 * noise[::-1], zero, 4000.0).converged
