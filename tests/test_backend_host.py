# SPDX-License-Identifier: BSD-3-Clause
"""The backend registry and RunKernel on the host, as the reference tests its own
(tests/unit/gpu/backend_test.cpp): names in any case, the report, the default from the
environment, resolution, and RunKernel leaving a grid alone when it cannot run. Written to hold
with or without a GPU: where hip is available the expectations follow, as the reference's do."""
import ctypes as C
import os
import subprocess
import sys
from pathlib import Path

import numpy as np
import pytest

from terraingen_amd import native as N

ROOT = Path(__file__).resolve().parent.parent
AUTO, CPU, VULKAN, GPU, CUDA, METAL, HIP = range(7)


def hip_available() -> bool:
    return N.backends_report().splitlines()[4].startswith("hip: available on ")


def test_backend_named_reads_each_name_in_any_case():
    assert N.backend_named("auto") == AUTO
    assert N.backend_named("CPU") == CPU
    assert N.backend_named("Vulkan") == VULKAN
    assert N.backend_named("gpu") == GPU
    assert N.backend_named("CUDA") == CUDA
    assert N.backend_named("metal") == METAL
    assert N.backend_named("HiP") == HIP
    assert N.backend_named("opencl") is None
    assert N.backend_named("") is None
    for b in range(7):
        assert N.backend_named(N.backend_name(b)) == b
    assert N.backend_name(99) == "unknown"


def test_backends_report_the_cpu_then_each_gpu_backend_with_its_device_or_why_not():
    lines = N.backends_report().splitlines()
    assert [line.split(":")[0] for line in lines] == ["cpu", "vulkan", "cuda", "metal", "hip",
                                                      "default"]
    assert lines[0] == "cpu: available on host"
    for line, name in zip(lines[1:4], ("vulkan", "cuda", "metal")):
        assert line == f"{name}: unavailable: this build has no {name} device; its GPU " \
                       "backend is hip"
    assert lines[4].startswith("hip: available on ") or lines[4].startswith("hip: unavailable: ")
    assert len(lines[4]) > len("hip: unavailable: ")


def test_defaults_follow_the_environment_until_set_in_code():
    # In a child process: the code default, once set, outlives every test in this one.
    code = r'''
import os
from terraingen_amd import native as N
def default(value):
    if value is None:
        os.environ.pop("TERRAINGEN_BACKEND", None)
    else:
        os.environ["TERRAINGEN_BACKEND"] = value
    try:
        return N.default_backend()
    except ValueError as e:
        return str(e)
assert default(None) == 1, default(None)
assert default("") == 1
assert default("auto") == 0
assert default("cpu") == 1
assert default("VULKAN") == 2
assert default("gpu") == 3
assert default("hip") == 6
msg = default("opencl")
assert msg == 'TERRAINGEN_BACKEND is "opencl"; expected auto, cpu, gpu, vulkan, cuda, metal or hip', msg
N.set_default_backend(1)
assert default("hip") == 1
print("ok")
'''
    out = subprocess.run([sys.executable, "-c", code], cwd=ROOT, capture_output=True, text=True,
                         env={k: v for k, v in os.environ.items() if k != "TERRAINGEN_BACKEND"})
    assert out.returncode == 0, out.stderr
    assert out.stdout.strip() == "ok"


def test_resolve_backend_picks_the_best_gpu_backend_and_otherwise_the_cpu():
    assert N.resolve_backend(CPU) == CPU
    for name, b in (("vulkan", VULKAN), ("cuda", CUDA), ("metal", METAL)):
        with pytest.raises(RuntimeError, match=f"The {name} backend is unavailable"):
            N.resolve_backend(b)
    if hip_available():
        assert N.resolve_backend(HIP) == HIP
        assert N.resolve_backend(GPU) == HIP
        assert N.resolve_backend(AUTO) == HIP
    else:
        with pytest.raises(RuntimeError, match="The hip backend is unavailable"):
            N.resolve_backend(HIP)
        with pytest.raises(RuntimeError, match="No GPU backend is available; hip: "):
            N.resolve_backend(GPU)
        assert N.resolve_backend(AUTO) == CPU


def test_run_kernel_leaves_the_grid_alone_without_a_gpu_form():
    grid = N.Grid(4, 3, 0.0, 3.0, 1.0)
    values = np.ones((3, 4), np.float32)
    kernel = N.noise_kernel([dict(seed=1, wavelength=8.0, amplitude=1.0, octaves=2)])
    for b in (CPU, AUTO, VULKAN, CUDA, METAL, HIP):
        assert not N.run_kernel(b, 0, kernel, grid, values)
    for b in (CPU, AUTO, GPU, VULKAN, CUDA, METAL):
        assert not N.run_kernel(b, N.TG_KERNEL_NOISE_LAYER, kernel, grid, values)
    if not hip_available():
        assert not N.run_kernel(HIP, N.TG_KERNEL_NOISE_LAYER, kernel, grid, values)
    empty = N.Grid(0, 3, 0.0, 3.0, 1.0)
    assert not N.run_kernel(HIP, N.TG_KERNEL_NOISE_LAYER, kernel, empty,
                            np.zeros(0, np.float32))
    assert (values == 1.0).all()


def test_backends_command_prints_the_report():
    out = subprocess.run([sys.executable, "-m", "terraingen_amd", "backends"], cwd=ROOT,
                         capture_output=True, text=True,
                         env={k: v for k, v in os.environ.items() if k != "TERRAINGEN_BACKEND"})
    assert out.returncode == 0, out.stderr
    assert out.stdout.splitlines()[0] == "cpu: available on host"
    assert out.stdout.splitlines()[-1] == "default: cpu"
