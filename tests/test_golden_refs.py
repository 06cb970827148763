"""The committed converged pressure drops of the 1024 x 512 network
workload: wherever both the CPU FP64 oracle and the device FP64 path were
run to a fixed point, their pressure drops agree far inside the 1e-4
accuracy bar the bench applies (the two differ only in summation order)."""
from __future__ import annotations

import json
from pathlib import Path

GOLD = Path(__file__).resolve().parent / "golden" / "network_1024x512_dp.json"


def test_oracle_and_device_fp64_agree():
    d = json.loads(GOLD.read_text())
    both = [e for e in d["seeds"].values() if "dp_oracle" in e and "dp_fp64_gpu" in e]
    assert len(both) >= 4
    for e in both:
        assert abs(e["dp_fp64_gpu"] / e["dp_oracle"] - 1) < 1e-7, e


def test_every_bench_seed_has_an_oracle_reference():
    """The bench grades every seed against the FP64 oracle's pressure drop,
    never against the device's own FP64 result."""
    d = json.loads(GOLD.read_text())
    for s in range(8):
        e = d["seeds"][str(s)]
        assert e["dp_oracle"] > 0 and e["iters_oracle"] > 0
