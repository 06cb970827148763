"""Pins of the pressure-drop convergence exit (SIMPLE.cpp:338-409) in the
oracle, on the CPU. The reference's own tests hold no fixture for it, so the
oracle's stop iteration is checked against the rule restated here in numpy
over the core-plane drop history recorded step by step: the window starts
at iteration rampSteps, holds dpConvergenceWindow drops, and the run stops
at the first full window whose standard deviation and latest deviation from
the mean are below dpConvergencePercent % of max(|mean|, 1), and, with the
slope gate, whose least-squares slope over the last dpSlopeWindowIters drops
is at most dpSlopeMaxDelta / dpSlopeWindowIters.
"""
from __future__ import annotations

import numpy as np
import pytest

import oracle_py as op
import workloads as wl

X_IN, X_OUT = 0.5e-3, 2.5e-3


def channel(window=30, percent=1.0, ramp=10, gate=False, slope_window=0, slope_max=10.0):
    p, ct = wl.straight_200x50()
    p.M, p.N = 20, 60
    ct = wl.geom_straight(20, 60, 4, 15)
    p.epsilon = 1e-14  # the residual test never fires first
    p.maxIterations = 3000
    p.usePressureDropConvergence = 1
    p.dpConvergenceWindow = window
    p.dpConvergencePercent = percent
    p.rampSteps = ramp
    p.enableInletRamp = 0
    p.usePressureDropSlopeGate = 1 if gate else 0
    p.dpSlopeWindowIters = slope_window
    p.dpSlopeMaxDelta = slope_max
    p.xPlaneCoreInlet, p.xPlaneCoreOutlet = X_IN, X_OUT
    return p, ct


def rule_stop(dps, window, percent, ramp, slope_window=0, slope_max=None):
    """First iteration index whose check passes, restated from the rule."""
    for it in range(len(dps)):
        if it < ramp or it - ramp + 1 < window:
            continue
        w = np.array(dps[it - window + 1:it + 1])
        mean = w.mean()
        den = max(abs(mean), 1.0)
        ok = w.std() / den * 100 < percent and abs(w[-1] - mean) / den * 100 < percent
        if slope_window:
            if it + 1 >= slope_window:
                y = np.array(dps[it - slope_window + 1:it + 1])
                slope = np.polyfit(np.arange(slope_window, dtype=float), y, 1)[0]
                ok = ok and abs(slope) <= slope_max / slope_window
        if ok:
            return it
    return None


def record(p, ct, limit=3000):
    o = op.Oracle(p, ct)
    dps, stop = [], None
    for it in range(limit):
        st = o.step()
        dps.append(o.plane(X_IN)["avgTotal"] - o.plane(X_OUT)["avgTotal"])
        if st["converged"]:
            stop = it
            break
    return dps, stop


@pytest.mark.parametrize("window,percent", [(30, 1.0), (60, 0.2)])
def test_window_exit_matches_rule(window, percent):
    p, ct = channel(window, percent)
    dps, stop = record(p, ct)
    assert stop is not None
    assert stop == rule_stop(dps, window, percent, 10)


def test_slope_gate_delays_exit():
    p, ct = channel(30, 1.0, gate=True, slope_window=80, slope_max=0.5)
    dps, stop = record(p, ct)
    assert stop is not None
    assert stop == rule_stop(dps, 30, 1.0, 10, slope_window=80, slope_max=0.5)
    plain = rule_stop(dps, 30, 1.0, 10)
    assert stop >= plain


def test_exit_off_runs_to_residual_test():
    p, ct = channel()
    p.usePressureDropConvergence = 0
    p.massResidualScope = 1  # the reference scope floors residMass on this channel
    p.epsilon = 1e-8
    p.maxIterations = 20000
    o = op.Oracle(p, ct)
    n, st, _ = o.run(20000)
    assert st["converged"] == 1 and st["error"] <= 1e-8
    assert n > 100  # well past where the window exit stops
