"""Build libstarflood_hip.so in-tree: C host code with gcc, HIP kernels with
hipcc for gfx950, linked by hipcc.  Compiles every source on each call.

    python -m starflood_amd.build
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import re
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
OBJ = PKG / "_build"
LIB_DIR = PKG / "lib"
LIB = LIB_DIR / "libstarflood_hip.so"
# the same library built with SOLVER_VIS: sim_solv, sim_step and
# solve_gravity_part_tree in the reference's render-buffer prototype form
LIB_SOLVERVIS = LIB_DIR / "libstarflood_hip_solvervis.so"

ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIPCC = shutil.which("hipcc") or f"{ROCM}/bin/hipcc"
CC = os.environ.get("CC", "gcc")

# -ffp-contract=off: the exact kernels keep the reference's unfused FP32
# operation order; fast kernels spell their FMAs out with __builtin_fma[f].
# -packed-fp32-ops: no v_pk_{add,mul,fma}_f32 are generated.
HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-I", str(CSRC),
]
C_FLAGS = [
    "-std=c11", "-O2", "-fPIC", "-ffp-contract=off", "-Wall", "-Wextra", "-Wno-unused-parameter",
    "-D__HIP_PLATFORM_AMD__", f"-I{ROCM}/include", "-I", str(CSRC),
]


INCLUDE = PKG.parent / "include"
_PROTO = re.compile(r"^(?!typedef|#|static|\s)[\w\s\*]*?\b([a-z_][a-z0-9_]*)\s*\(", re.M)


def public_symbols() -> list[str]:
    """Function names declared in include/*.h: the library's whole export list."""
    names = []
    for h in sorted(INCLUDE.glob("*.h")):
        text = re.sub(r"/\*.*?\*/", "", h.read_text(), flags=re.S)
        names += _PROTO.findall(text)
    return sorted(set(names))


def _run(cmd: list[str]) -> None:
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        raise RuntimeError(f"command failed ({proc.returncode}): {' '.join(cmd)}\n{proc.stdout}\n{proc.stderr}")
    noise = [l for l in proc.stderr.splitlines() if "packed-fp32-ops" not in l and l.strip()]
    if noise:
        sys.stderr.write("\n".join(noise) + "\n")


def build(verbose: bool = False) -> Path:
    OBJ.mkdir(parents=True, exist_ok=True)
    (OBJ / "solvervis").mkdir(parents=True, exist_ok=True)
    LIB_DIR.mkdir(parents=True, exist_ok=True)
    dev, host, host_vis = [], [], []
    for src in sorted((CSRC / "device").glob("*.hip")):
        obj = OBJ / (src.stem + ".hip.o")
        dev.append(([HIPCC, *HIP_FLAGS, "-c", str(src), "-o", str(obj)], obj))
    for src in sorted((CSRC / "host").glob("*.c")):
        obj = OBJ / (src.stem + ".c.o")
        host.append(([CC, *C_FLAGS, "-c", str(src), "-o", str(obj)], obj))
        obj_vis = OBJ / "solvervis" / (src.stem + ".c.o")
        host_vis.append(([CC, *C_FLAGS, "-DSOLVER_VIS", "-c", str(src), "-o", str(obj_vis)], obj_vis))
    workers = max(1, min(8, os.cpu_count() or 1))
    with cf.ThreadPoolExecutor(workers) as ex:
        futs = [ex.submit(_run, cmd) for cmd, _ in dev + host + host_vis]
        for f in futs:
            f.result()
    # export exactly what include/*.h declares; bind internal calls locally
    vmap = OBJ / "exports.map"
    vmap.write_text("{\n  global:\n" + "".join(f"    {n};\n" for n in public_symbols()) + "  local: *;\n};\n")
    for lib, objs in ((LIB, dev + host), (LIB_SOLVERVIS, dev + host_vis)):
        tmp = lib.with_suffix(".so.tmp")
        _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", str(tmp), *[str(o) for _, o in objs],
              "-Wl,--no-undefined", "-Wl,-Bsymbolic", f"-Wl,--version-script={vmap}", "-lm"])
        os.replace(tmp, lib)
        if verbose:
            print(f"built {lib}")
    return LIB


if __name__ == "__main__":
    build(verbose=True)
