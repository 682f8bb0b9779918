# SPDX-License-Identifier: BSD-3-Clause
"""Builds the native libraries of terraingen_amd in-tree.

* ``libtg_hip.so``       - the gfx950 kernels and their C ABI (include/tg_hip.h), by hipcc.
* ``libterraingen_amd.so`` - the C++ host: backend registry with ``hip``, ``RunKernel``, layers,
  crater/boulder placement and the batched tile server, by g++. It opens ``libtg_hip.so`` with
  dlopen on first use, as the reference opens its GPU drivers (src/gpu/CMakeLists.txt:3-6).

Both are built with ``-ffp-contract=off`` as the reference is (its CMakeLists.txt:32-38).
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
INCLUDE = ROOT / "include"

HIP_SO = PKG / "libtg_hip.so"
HOST_SO = PKG / "libterraingen_amd.so"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("TG_OFFLOAD_ARCH", "gfx950")

HIP_FLAGS = [
    "-std=c++17", "-O3", "-fPIC", "-shared",
    f"--offload-arch={ARCH}",
    "-ffp-contract=off",
    # Keep the compiler from forming packed FP32 instructions.
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
]


def _run(cmd: list[str]) -> None:
    print("+", " ".join(str(c) for c in cmd), flush=True)
    subprocess.run([str(c) for c in cmd], check=True)


def build_hip() -> Path:
    """Each source compiled to an object on its own process (they share no device code), then
    linked into the shared library."""
    from concurrent.futures import ThreadPoolExecutor

    sources = [CSRC / "tg_hip.hip", CSRC / "batch.hip", CSRC / "boulders.hip", CSRC / "probe.hip",
               CSRC / "upsample.hip"]
    objdir = PKG / "_build"
    objdir.mkdir(exist_ok=True)
    compile_flags = [f for f in HIP_FLAGS if f != "-shared"]

    def compile_one(src: Path) -> Path:
        obj = objdir / (src.stem + ".o")
        _run([HIPCC, *compile_flags, f"-I{INCLUDE}", f"-I{CSRC}", "-c", src, "-o", obj])
        return obj

    with ThreadPoolExecutor(max_workers=len(sources)) as pool:
        objects = list(pool.map(compile_one, sources))
    tmp = HIP_SO.with_suffix(".so.tmp")
    _run([HIPCC, *HIP_FLAGS, *objects, "-o", tmp])
    os.replace(tmp, HIP_SO)
    return HIP_SO


def host_sources() -> list[Path]:
    return sorted((CSRC / "host").glob("*.cpp"))


def build_host() -> Path | None:
    sources = host_sources()
    if not sources:
        return None
    tmp = HOST_SO.with_suffix(".so.tmp")
    _run(["g++", "-std=c++17", "-O2", "-fPIC", "-shared", "-ffp-contract=off", "-Wall",
          "-Wextra", "-pthread", f"-I{INCLUDE}", f"-I{CSRC}", f"-I{CSRC / 'host'}", *sources,
          "-o", tmp, "-ldl", "-lz"])
    os.replace(tmp, HOST_SO)
    return HOST_SO


def build_oracle(with_reference: bool = True) -> None:
    """The oracle (test infrastructure): liboracle.so, and oracle/_ref from the reference's own
    sources when they are present on this machine."""
    oracle = ROOT / "oracle"
    target = "ref" if with_reference and Path("/root/reference").is_dir() else "all"
    _run(["make", "-C", oracle, "-j8", target])


def build_all() -> None:
    if shutil.which(HIPCC) is None and not Path(HIPCC).exists():
        raise RuntimeError(f"hipcc not found at {HIPCC}")
    build_hip()
    build_host()
    build_oracle()


if __name__ == "__main__":
    build_all()
    sys.exit(0)
