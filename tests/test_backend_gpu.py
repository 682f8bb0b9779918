# SPDX-License-Identifier: BSD-3-Clause
"""RunKernel on the hip backend: the reference's strip dispatcher (src/gpu/kernel.cpp:130-153)
over the noise layer and crater kernels, against the oracle and the reference's heights."""
import ctypes as C

import numpy as np
import pytest

from terraingen_amd import native as N

from crater_cases import binding, load

pytestmark = pytest.mark.gpu

HIP, GPU, AUTO = 6, 3, 0


def test_hip_is_the_resolved_gpu_backend():
    line = N.backends_report().splitlines()[4]
    assert line.startswith("hip: available on "), line
    assert N.resolve_backend(GPU) == HIP and N.resolve_backend(AUTO) == HIP


@pytest.mark.parametrize("strip_bytes", [0, 4096])
def test_noise_layer_through_run_kernel_matches_the_oracle(orc, strip_bytes):
    """ApplyCell at each centre measured from the layer's origin, narrowed per column and row,
    bit for bit; a small strip size splits the grid into many strips."""
    grid = N.Grid(97, 61, 1000.0, 2500.0, 0.75)
    origin = (1030.5, 2480.25)
    kernel = N.noise_kernel([dict(seed=7, wavelength=64.0, amplitude=5.0, octaves=6)])
    values = np.full((61, 97), 3.0, np.float32)
    values[::9, ::11] = -9999.0
    assert N.run_kernel(HIP, N.TG_KERNEL_NOISE_LAYER, kernel, grid, values, origin,
                        no_data=-9999.0, max_strip_bytes=strip_bytes)
    cx, cy = N.grid_centers(grid)
    xs = (cx - origin[0]).astype(np.float32)
    ys = (cy - origin[1]).astype(np.float32)
    want = np.full((61, 97), 3.0, np.float32)
    want[::9, ::11] = -9999.0
    params = N.KernelParams(97, 61, N.NODATA_VALUE, -9999.0, 0)
    orc.or_noise_grid(C.byref(kernel), xs.ctypes.data, ys.ctypes.data, 97, 61, C.byref(params),
                      want.ctypes.data, 1)
    assert np.array_equal(values.view(np.uint32), want.view(np.uint32))


def test_craters_through_run_kernel_match_the_reference(hip):
    case = load("c2_powerlaw")
    h, w = case.shape
    g = case.grid
    whole = np.zeros((h, w), np.float32)
    assert N.run_kernel(HIP, N.TG_KERNEL_CRATERS, case.kernel, g, whole, (0.0, 0.0),
                        binding(case))
    strips = np.zeros((h, w), np.float32)
    assert N.run_kernel(HIP, N.TG_KERNEL_CRATERS, case.kernel, g, strips, (0.0, 0.0),
                        binding(case), max_strip_bytes=w * 4 * 37)
    assert np.array_equal(whole.view(np.uint32), strips.view(np.uint32))
    assert np.abs(whole - case.heights).max() <= 2e-5  # FP32 rounding of the fast outline
