"""Builds libcoldplate_hip.so (gfx950) in place, and the CPU oracle.

    python coldplate-solver_amd/build.py           # library + oracle
    python coldplate-solver_amd/build.py --lib     # library only

The library is built in-tree (coldplate-solver_amd/lib/) so that it travels
with the repository snapshot; nothing goes to a JIT cache.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
REPO = PKG.parent
CSRC = PKG / "csrc"
LIB_DIR = PKG / "lib"
OBJ_DIR = PKG / "lib" / "obj"
LIB = LIB_DIR / "libcoldplate_hip.so"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

# Packed FP32 instructions are disabled for all device code.
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-munsafe-fp-atomics", "-Wall", "-Wno-unused-function",
]
HOST_FLAGS = ["-O3", "-std=c++17", "-fPIC", "-fopenmp", "-march=x86-64-v3", "-Wall"]


def run(cmd: list[str]) -> None:
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build_lib() -> Path:
    OBJ_DIR.mkdir(parents=True, exist_ok=True)
    objs = []
    for src in ("cp_kernels.hip", "cp_engine.hip", "cp_thermal.hip"):
        obj = OBJ_DIR / (src + ".o")
        run([HIPCC, *HIP_FLAGS, "-c", str(CSRC / src), "-o", str(obj)])
        objs.append(obj)
    for src in ("cp_geometry.cpp", "cp_io.cpp"):
        obj = OBJ_DIR / (Path(src).stem + ".o")
        run(["g++", *HOST_FLAGS, "-c", str(CSRC / src), "-o", str(obj)])
        objs.append(obj)
    tmp = LIB.with_suffix(".so.tmp")
    run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *map(str, objs), "-lgomp", "-o", str(tmp)])
    os.replace(tmp, LIB)
    return LIB


def build_binding_check() -> Path | None:
    """The reference-side binding (binding/hip_backend.cpp) linked into its
    check program against the SIMPLE stand-in under tests/binding."""
    stub = REPO / "tests" / "binding"
    if not (stub / "check_binding.cpp").exists() or shutil.which("g++") is None:
        return None
    out = stub / "_build" / "check_binding"
    out.parent.mkdir(parents=True, exist_ok=True)
    run(["g++", "-std=c++17", "-O1", "-Wall", "-I", str(stub), "-I", str(REPO / "include"), "-I",
         str(PKG / "binding"), str(stub / "check_binding.cpp"), str(PKG / "binding" / "hip_backend.cpp"),
         "-L", str(LIB_DIR), "-lcoldplate_hip", f"-Wl,-rpath,{LIB_DIR}", "-o", str(out)])
    return out


def build_oracle(with_ref: bool = True) -> None:
    targets = ["all"]
    if with_ref and Path("/root/reference/Utilities/solvers/cpu/pcg_solver.cpp").exists():
        targets.append("ref")
    run(["make", "-C", str(REPO / "oracle"), *targets])


def main(argv: list[str]) -> int:
    if shutil.which(HIPCC) is None and not Path(HIPCC).exists():
        print(f"hipcc not found at {HIPCC}", file=sys.stderr)
        return 1
    build_lib()
    if "--lib" not in argv:
        build_oracle()
        build_binding_check()
    return 0


if __name__ == "__main__":
    sys.exit(main(sys.argv[1:]))
