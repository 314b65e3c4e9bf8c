"""Build recipes for the in-tree native libraries.

* ``umbreon_amd/lib/libumbreon_gpu.so`` -- the gfx950 per-ray backend behind the
  C-ABI in ``include/umbreon_gpu.h`` (HIP kernels + host scene builder).
* ``umbreon_amd/lib/libumbreon.so`` -- the C++ facade (``umbreon::render`` and
  friends, ``include/umbreon/``) layered on the C-ABI.
* ``oracle/liboracle.so`` -- the CPU restatement used only as the checker by
  the tests, ``__graft_entry__.smoke()`` and the CPU-baseline leg of
  ``bench.py``.

Everything is compiled in place so the shared objects travel with the source
tree. ``hipcc`` cross-compiles for gfx950 without a GPU present.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent
PKG = ROOT / "umbreon_amd"
CSRC = PKG / "csrc"
CPP = PKG / "cpp"
LIB = PKG / "lib"
BUILD = PKG / "_build"
ORACLE = ROOT / "oracle"
INCLUDE = ROOT / "include"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

# Exact IEEE fp32: no contraction into FMA and correctly rounded divide/sqrt,
# so the intersectors and shading round the way the written expressions do.
# Packed-fp32 instructions are disabled for all device code.
HIP_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-result",
]
HOST_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-ffp-contract=off", "-Wall"]

GPU_SOURCES = ["ub_kernels.hip", "ub_devbuild.hip", "ub_unet.hip", "ub_static.hip", "ub_pathtrace.hip", "ub_capi.cpp",
               "ub_build.cpp", "ub_edges.cpp"]


def _run(cmd: list[str], cwd: Path | None = None) -> None:
    print("+", " ".join(str(c) for c in cmd), flush=True)
    res = subprocess.run([str(c) for c in cmd], cwd=cwd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout)
        sys.stderr.write(res.stderr)
        raise RuntimeError(f"build step failed ({res.returncode}): {cmd[0]}")
    # surface warnings other than the known host-pass note about the
    # packed-fp32 target feature
    for line in (res.stderr or "").splitlines():
        if "packed-fp32-ops" in line or "not a recognized feature" in line:
            continue
        if "warning" in line or "error" in line:
            sys.stderr.write(line + "\n")


def build_gpu() -> Path:
    """Compile the HIP backend into umbreon_amd/lib/libumbreon_gpu.so."""
    BUILD.mkdir(exist_ok=True)
    LIB.mkdir(exist_ok=True)
    objs = [BUILD / (src + ".o") for src in GPU_SOURCES]

    def compile_one(src: str, obj: Path) -> None:
        lang = ["-x", "hip"] if src.endswith(".hip") else []
        _run([HIPCC, *HIP_FLAGS, *lang, "-I", INCLUDE, "-c", CSRC / src, "-o", obj])

    # the translation units are independent: compile them side by side
    jobs = max(1, min(len(GPU_SOURCES), int(os.environ.get("UBG_BUILD_JOBS", "6"))))
    with ThreadPoolExecutor(jobs) as pool:
        for f in [pool.submit(compile_one, src, obj) for src, obj in zip(GPU_SOURCES, objs)]:
            f.result()
    out = LIB / "libumbreon_gpu.so"
    tmp = LIB / "libumbreon_gpu.so.tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", *objs, "-o", tmp,
          "-Wl,-soname,libumbreon_gpu.so"])
    tmp.replace(out)
    return out


def build_facade() -> Path:
    """Compile the C++ facade (umbreon::render) into libumbreon.so."""
    LIB.mkdir(exist_ok=True)
    srcs = sorted(CPP.glob("*.cpp"))
    out = LIB / "libumbreon.so"
    tmp = LIB / "libumbreon.so.tmp"
    _run(["g++", *HOST_FLAGS, "-shared", "-I", INCLUDE, "-I", INCLUDE / "umbreon", *srcs,
          "-L", LIB, "-lumbreon_gpu", "-Wl,-rpath,$ORIGIN", "-pthread", "-o", tmp])
    tmp.replace(out)
    return out


def build_examples() -> list[Path]:
    """Link the C++ example/test programs against the facade."""
    outs = []
    exdir = PKG / "examples"
    for src in sorted(exdir.glob("*.cpp")):
        exe = LIB / src.stem
        _run(["g++", *HOST_FLAGS, "-I", INCLUDE, "-I", INCLUDE / "umbreon", src, "-L", LIB,
              "-lumbreon", "-lumbreon_gpu", "-Wl,-rpath,$ORIGIN", "-pthread", "-o", exe])
        outs.append(exe)
    return outs


def build_oracle() -> Path:
    """Compile the CPU restatement (checker only) into oracle/liboracle.so."""
    out = ORACLE / "liboracle.so"
    tmp = ORACLE / "liboracle.so.tmp"
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off", "-fno-fast-math",
          "-pthread", "-Wall", ORACLE / "oracle.cpp", "-o", tmp])
    tmp.replace(out)
    return out


def build_all() -> None:
    build_gpu()
    if any(CPP.glob("*.cpp")):
        build_facade()
        if (PKG / "examples").is_dir():
            build_examples()
    build_oracle()


if __name__ == "__main__":
    what = sys.argv[1:] or ["all"]
    for w in what:
        {"all": build_all, "gpu": build_gpu, "facade": build_facade, "oracle": build_oracle,
         "examples": build_examples}[w]()
