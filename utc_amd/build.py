"""Build recipe for the gfx950 nbody library and its C++ host programs.

Compiles, in-tree:
  utc_amd/lib/libutcnb.so      HIP kernels + the C ABI of include/utcnb.h
  utc_amd/lib/nbody_main       the UTC nbody driver over the C++ task shim
  utc_amd/lib/nbody_mtask      the multi-task driver
  utc_amd/lib/nbody_mgpu       the multi-GPU driver (one task thread per GPU)
  utc_amd/lib/test_utc_shim    CPU self-test of the task shim
and the CPU oracle (oracle/Makefile).  Every HIP translation unit is built
with the packed-FP32 instructions disabled
(``-Xclang -target-feature -Xclang -packed-fp32-ops``).

Usage: ``python -m utc_amd.build`` (or ``__graft_entry__.build()``).
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent
PKG = ROOT / "utc_amd"
CSRC = PKG / "csrc"
UTC = PKG / "utc"
APPS = PKG / "apps" / "nbody"
OUT = PKG / "lib"
OBJ = PKG / "_build"
INCLUDE = ROOT / "include"

ARCH = "gfx950"
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
HIP_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-fPIC",
    "-std=c++17",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # 64-byte aligned loop headers: the source loop of the force kernels runs
    # ~25% slower on MI355X when its header straddles a 32-byte fetch block.
    "-falign-loops=64",
    "-Wno-unused-command-line-argument",
    f"-I{INCLUDE}",
    f"-I{CSRC}",
]
# Files whose arithmetic must follow the reference's order with no fused
# multiply-add.
NO_CONTRACT = {"ref.hip"}

LIB_SOURCES = [
    "ref.hip",
    "force_fp32.hip",
    "sym.hip",
    "sym_vjp.hip",
    "force_fp64.hip",
    "aux.hip",
    "adjoint.hip",
    "batch.hip",
    "batch_api.cpp",
    "capi.cpp",
    "gen.cpp",
    "autotune.cpp",
    "noalloc.hip",
]


def _run(cmd: list[str], cwd: Path | None = None) -> None:
    proc = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if proc.returncode != 0:
        raise RuntimeError(f"build step failed ({proc.returncode}): {' '.join(cmd)}\n{proc.stdout}")


def _compile(src: Path, obj: Path) -> Path:
    flags = list(HIP_FLAGS)
    if src.name in NO_CONTRACT:
        flags.append("-ffp-contract=off")
    lang = ["-x", "hip"] if src.suffix == ".hip" else []
    _run([HIPCC, *flags, *lang, "-c", str(src), "-o", str(obj)])
    return obj


def build_library(jobs: int | None = None) -> Path:
    OUT.mkdir(parents=True, exist_ok=True)
    OBJ.mkdir(parents=True, exist_ok=True)
    srcs = [CSRC / s for s in LIB_SOURCES]
    missing = [str(s) for s in srcs if not s.exists()]
    if missing:
        raise RuntimeError(f"library sources missing: {missing}")
    jobs = jobs or min(len(srcs), max(1, (os.cpu_count() or 4)))
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(lambda s: _compile(s, OBJ / (s.name + ".o")), srcs))
    lib = OUT / "libutcnb.so"
    tmp = OUT / "libutcnb.so.tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs)])
    os.replace(tmp, lib)
    return lib


def build_host_programs() -> list[Path]:
    """The C++ UTC task shim and the nbody drivers (host code over the C ABI)."""
    if not UTC.exists():
        return []
    shim_src = sorted((UTC / "src").glob("*.cpp"))
    common = [
        "-O2", "-std=c++17", "-pthread", "-D__HIP_PLATFORM_AMD__",
        f"-I{INCLUDE}", f"-I{UTC / 'include'}", f"-I{APPS}", "-I/opt/rocm/include",
    ]
    link = [f"-L{OUT}", "-lutcnb", f"-Wl,-rpath,$ORIGIN", "-L/opt/rocm/lib", "-lamdhip64",
            "-Wl,-rpath,/opt/rocm/lib"]
    outs = []
    targets = {
        "nbody_main": [APPS / "nbody_main.cpp", APPS / "body_task_sgpu.cpp", APPS / "task.cpp"],
        "nbody_mtask": [APPS / "nbody_mtask.cpp", APPS / "body_task_sgpu.cpp", APPS / "task.cpp"],
        "nbody_mgpu": [APPS / "nbody_mgpu.cpp", APPS / "body_task_mgpu.cpp", APPS / "task.cpp"],
        "test_utc_shim": [UTC / "test" / "test_utc_shim.cpp"],
    }

    def one(item):
        name, files = item
        files = [f for f in files if f.exists()]
        if not files:
            return None
        exe = OUT / name
        _run(["g++", *common, "-ffp-contract=off", "-o", str(exe), *map(str, files),
              *map(str, shim_src), *link])
        return exe

    with cf.ThreadPoolExecutor(max_workers=4) as ex:
        for r in ex.map(one, targets.items()):
            if r is not None:
                outs.append(r)
    return outs


def build_oracle() -> None:
    make = shutil.which("make")
    if make is None:
        raise RuntimeError("make not found")
    _run([make, "-s", "-C", str(ROOT / "oracle")])


TEST_DEVICE = ROOT / "tests" / "device"


def build_test_programs() -> list[Path]:
    """Stand-alone gfx950 check programs the GPU tests run (tests/device/*.hip),
    into tests/device/_bin/ (test infrastructure, not part of the library)."""
    outs = []
    out_dir = TEST_DEVICE / "_bin"
    for src in sorted(TEST_DEVICE.glob("*.hip")):
        out_dir.mkdir(parents=True, exist_ok=True)
        exe = out_dir / src.stem
        _run([HIPCC, *HIP_FLAGS, "-o", str(exe), str(src)])
        outs.append(exe)
    return outs


def build_all() -> None:
    with cf.ThreadPoolExecutor(max_workers=2) as ex:
        f_oracle = ex.submit(build_oracle)
        lib = build_library()
        f_oracle.result()
    build_host_programs()
    build_test_programs()
    print(f"built {lib}")


if __name__ == "__main__":
    build_all()
    sys.exit(0)
