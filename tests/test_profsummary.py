"""profsummary.py's selection of the timing bursts (CPU only): in a kernel
trace, the launches bench.py times are the first 16 of every run of at
least 16 back-to-back dispatches of the kernel; a SIMPLE step never makes
such a run, since every launch of the kernel is followed by another kernel."""
from __future__ import annotations

import sqlite3

import profsummary as ps

K = "void cpdev::k_pcg_q<float, 4>(cpdev::KArgs)"
F = "cpdev::k_finalize(cpdev::KArgs, int, int)"


def trace(tmp_path, seq):
    d = tmp_path / "trace"
    d.mkdir()
    c = sqlite3.connect(d / "run_results.db")
    c.execute("create table kernels (name text, start integer, duration integer)")
    t = 0
    for name, dur in seq:
        c.execute("insert into kernels values (?, ?, ?)", (name, t, dur))
        t += dur + 10
    c.commit()
    c.close()
    return str(d)


def test_bursts_are_the_long_runs_only(tmp_path):
    step = [(K, 30000), (F, 4000)] * 5                 # ordinary steps: never two in a row
    burst = [(K, 20000 + k) for k in range(16)]       # one timing burst
    tail = [(K, 50000)]                                # another stream's launch right after it
    d = trace(tmp_path, step + burst + tail + step + [(F, 1)] + burst + step)
    name, dur = ps.burst_durations(d, "k_pcg_q<float", 16)
    assert name == K
    assert dur == [20000.0 + k for k in range(16)] * 2  # first 16 of each run, the trailing launch excluded


def test_short_runs_are_not_bursts(tmp_path):
    d = trace(tmp_path, [(K, 1000)] * 15 + [(F, 1)] + [(K, 1000)] * 3)
    _, dur = ps.burst_durations(d, "k_pcg_q", 16)
    assert dur == []
