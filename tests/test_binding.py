"""The reference-side binding (coldplate-solver_amd/binding/hip_backend.cpp)
compiles against a stand-in for the reference's SIMPLE class that uses the
SIMPLE.h member names and types, maps every parameter, converts Eigen's
column-major fields, and reports a missing device as an error; on an MI355X
the same program runs outer iterations through it. The C layout of every
ABI struct equals the ctypes layout the tests and the bench use."""
from __future__ import annotations

import ctypes as C
import re
import subprocess
from pathlib import Path

import pytest

import cpbind as cb

REPO = Path(__file__).resolve().parent.parent
BIND = REPO / "coldplate-solver_amd" / "binding"
STUB = REPO / "tests" / "binding"
LIB_DIR = REPO / "coldplate-solver_amd" / "lib"
BUILT = STUB / "_build" / "check_binding"
REF_SIMPLE_H = Path("/root/reference/SIMPLE.h")


def compile_check(out: Path) -> None:
    cmd = ["g++", "-std=c++17", "-O1", "-Wall", "-Wextra", "-Werror", "-I", str(STUB), "-I", str(REPO / "include"),
           "-I", str(BIND), str(STUB / "check_binding.cpp"), str(BIND / "hip_backend.cpp"), "-L", str(LIB_DIR),
           "-lcoldplate_hip", f"-Wl,-rpath,{LIB_DIR}", "-o", str(out)]
    subprocess.run(cmd, check=True, capture_output=True, text=True)


def test_binding_compiles_and_maps_parameters(tmp_path):
    exe = tmp_path / "check_binding"
    compile_check(exe)
    r = subprocess.run([str(exe)], capture_output=True, text=True, env={"HIP_VISIBLE_DEVICES": "-1"}, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "FAIL" not in r.stdout
    assert r.stdout.count("ok ") == 10, r.stdout


def stub_members() -> dict[str, str]:
    text = (STUB / "SIMPLE.h").read_text()
    body = text[text.index("class SIMPLE {"):]
    body = body[:body.index("// what hip_backend.h")]
    out = {}
    for ty, names in re.findall(r"^\s*(float|int|bool)\s+([^;(]+);", body, flags=re.M):
        for n in names.split(","):
            out[n.split("=")[0].strip()] = ty
    return out


@pytest.mark.skipif(not REF_SIMPLE_H.exists(), reason="the reference tree is not mounted here")
def test_stand_in_members_match_reference_header():
    """Every scalar member of the stand-in exists in SIMPLE.h with the same
    type (the binding compiles against the real header by the same names)."""
    ref = REF_SIMPLE_H.read_text()
    members = stub_members()
    assert len(members) >= 55
    for name, ty in members.items():
        m = re.search(rf"^\s*(float|int|bool)\s+[^;]*\b{name}\b\s*(=|;|,)", ref, flags=re.M)
        assert m, name
        assert m.group(1) == ty, (name, ty, m.group(1))


def test_abi_struct_layouts_match_ctypes(tmp_path):
    """sizeof / offsetof of every ABI struct from the C compiler equal the
    ctypes mirrors in cpbind (a field added to one and not the other, or a
    type changed, shifts the rest)."""
    structs = {"cp_params": cb.Params, "cp_step_stats": cb.StepStats, "cp_plane_metrics": cb.PlaneMetrics,
               "cp_slot_info": cb.SlotInfo, "cp_request": cb.Request, "cp_result": cb.Result}
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "coldplate_hip.h"', "int main(void) {"]
    for cname, py in structs.items():
        lines.append(f'printf("{cname} size %zu\\n", sizeof({cname}));')
        for f, _ in py._fields_:
            lines.append(f'printf("{cname} {f} %zu\\n", offsetof({cname}, {f}));')
    lines.append("return 0; }")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", "-Wall", "-I", str(REPO / "include"), str(src), "-o", str(exe)], check=True)
    got = {}
    for line in subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.splitlines():
        s, f, v = line.split()
        got[(s, f)] = int(v)
    for cname, py in structs.items():
        assert got[(cname, "size")] == C.sizeof(py), cname
        for f, _ in py._fields_:
            assert got[(cname, f)] == getattr(py, f).offset, (cname, f)


@pytest.mark.gpu
def test_binding_runs_outer_iterations_on_the_device():
    """The prebuilt binding check (built by coldplate-solver_amd/build.py)
    opens a session from the stand-in SIMPLE, runs three outer iterations
    with calculateStepHip, samples the inlet plane and syncs the fields
    back."""
    assert BUILT.exists(), f"{BUILT} missing: run python coldplate-solver_amd/build.py"
    print("[binding] check_binding --device", flush=True)
    r = subprocess.run([str(BUILT), "--device"], capture_output=True, text=True, timeout=120)
    print(r.stdout, flush=True)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "FAIL" not in r.stdout and "opened a device session" in r.stdout
