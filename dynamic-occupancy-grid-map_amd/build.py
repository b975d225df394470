"""Build recipe for libdogm_hip.so (gfx950) and the C++ drop-in tests.

Run as ``python dynamic-occupancy-grid-map_amd/build.py`` or through ``__graft_entry__.build()``.
Compiles every ``csrc/*.hip`` file with hipcc for gfx950 into objects under ``_build/`` and links
them into ``lib/libdogm_hip.so``.  The numerics are pinned so that device and host arithmetic agree
bit for bit: no FMA contraction, no fast-math, correctly rounded f32 division and square root, no
denormal flushing, and no packed-FP32 instructions.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
REPO = PKG.parent
CSRC = PKG / "csrc"
BUILD = PKG / "_build"
LIB = PKG / "lib"
SONAME = "libdogm_hip.so"

ROCM = Path(os.environ.get("ROCM_PATH", "/opt/rocm"))
HIPCC = str(ROCM / "bin" / "hipcc")

HIP_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-fno-fast-math",
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-fno-gpu-flush-denormals-to-zero",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-result",
]


def _run(cmd: list[str], cwd: Path | None = None) -> None:
    res = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    # hipcc's host pass does not know the packed-fp32 target feature; that warning is expected.
    out = "\n".join(l for l in res.stdout.splitlines() if "packed-fp32-ops" not in l)
    if res.returncode != 0:
        raise RuntimeError(f"command failed ({res.returncode}): {' '.join(cmd)}\n{out}")
    if out.strip():
        print(out)


def build_lib(jobs: int = 8) -> Path:
    BUILD.mkdir(exist_ok=True)
    LIB.mkdir(exist_ok=True)
    sources = sorted(CSRC.glob("*.hip"))
    objs = [BUILD / (s.stem + ".o") for s in sources]

    def compile_one(src_obj):
        src, obj = src_obj
        _run([HIPCC, *HIP_FLAGS, "-c", str(src), "-o", str(obj)])

    with cf.ThreadPoolExecutor(max_workers=max(1, min(jobs, len(sources)))) as ex:
        list(ex.map(compile_one, zip(sources, objs)))
    out = LIB / SONAME
    tmp = LIB / (SONAME + ".tmp")
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs),
          f"-L{ROCM / 'lib'}", "-lrocprofiler-sdk-roctx", f"-Wl,-rpath,{ROCM / 'lib'}"])
    os.replace(tmp, out)
    return out


def build_oracle() -> Path:
    """The CPU checker (test infrastructure; see oracle/Makefile)."""
    _run(["make", "-s", "-C", str(REPO / "oracle"), "-B"])
    return REPO / "oracle" / "liboracle_dogm.so"


def build_cpp_tests() -> list[Path]:
    """C++ drop-in tests: include/dogm/dogm.h linked against libdogm_hip.so with g++."""
    outs = []
    tdir = REPO / "tests" / "cpp"
    for src in sorted(tdir.glob("*.cpp")):
        exe = tdir / src.stem
        _run([
            "g++", "-O2", "-std=c++17", "-Wall", "-Wextra",
            f"-I{REPO / 'include'}", str(src), "-o", str(exe),
            f"-L{LIB}", "-ldogm_hip", f"-Wl,-rpath,{LIB}",
        ])
        outs.append(exe)
    return outs


def build_all() -> None:
    if not shutil.which(HIPCC) and not Path(HIPCC).exists():
        raise RuntimeError(f"hipcc not found at {HIPCC}")
    lib = build_lib()
    print(f"built {lib}")
    print(f"built {build_oracle()}")
    for exe in build_cpp_tests():
        print(f"built {exe}")


if __name__ == "__main__":
    build_all()
    sys.exit(0)
