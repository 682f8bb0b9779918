# SPDX-License-Identifier: BSD-3-Clause
"""Noise layer on the MI355X against the oracle: bit-exact, since the path is integer hashing and
IEEE FP32 adds and multiplies in the reference's order with contraction off."""
import ctypes as C

import numpy as np
import pytest

from terraingen_amd import native as N

pytestmark = pytest.mark.gpu


def grid_axes(width, height, pixel, center=(0.0, 0.0), origin=(0.0, 0.0)):
    """Column and row offsets of a north-up grid centred on `center`, measured from `origin` in
    FP64 then narrowed (src/gpu/kernel.cpp:55-70; src/sources/flat_source.cpp:23-25)."""
    ox = center[0] - 0.5 * width * pixel
    oy = center[1] + 0.5 * height * pixel
    xs = (ox + (np.arange(width) + 0.5) * pixel - origin[0]).astype(np.float32)
    ys = (oy - (np.arange(height) + 0.5) * pixel - origin[1]).astype(np.float32)
    return xs, ys


def oracle_grid(orc, kernel, xs, ys, heights, params):
    out = heights.copy()
    orc.or_noise_grid(C.byref(kernel), xs.ctypes.data, ys.ctypes.data, xs.size, ys.size,
                      C.byref(params), out.ctypes.data, 0)
    return out


C2_FBM = dict(seed=1, wavelength=512.0, amplitude=20.0, octaves=8, lacunarity=2.0, gain=0.5)


@pytest.mark.parametrize("fractals", [
    [C2_FBM],
    [dict(seed=42)],
    [dict(seed=3, type=1, offset=0.9), dict(seed=4, type=2, blend=1, amplitude=7.0)],
    [dict(seed=5, type=3, basis=1), dict(seed=6, basis=2, blend=2), dict(seed=8, basis=3, blend=3)],
    [dict(seed=9, basis=4, octaves=3), dict(seed=10, basis=5, warp_amplitude=40.0, amplitude=10.0),
     dict(seed=11, basis=6, octaves=2)],
    [dict(seed=12, warp_amplitude=25.0, warp_wavelength=300.0, octaves=12)],
])
def test_noise_strip_bit_exact(hip, orc, fractals):
    kernel = N.noise_kernel(fractals)
    xs, ys = grid_axes(203, 77, 1.75, center=(123.0, -456.0))
    rng = np.random.default_rng(0)
    heights = rng.normal(0, 3, (77, 203)).astype(np.float32)
    params = N.KernelParams(203, 77, N.NODATA_NONE, 0.0, 0)
    want = oracle_grid(orc, kernel, xs, ys, heights, params)
    got = heights.copy()
    hip.run(N.TG_KERNEL_NOISE_LAYER, kernel, got, xs, ys, params)
    print(f"[noise] {len(fractals)} fractals: max |diff| {np.abs(got - want).max():.3g}", flush=True)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))


def test_noise_skips_no_data(hip, orc):
    kernel = N.noise_kernel([C2_FBM])
    xs, ys = grid_axes(64, 33, 0.5)
    heights = np.zeros((33, 64), np.float32)
    heights[::5, ::3] = -9999.0
    heights[1::7, 2::4] = np.nan
    for mode, value in [(N.NODATA_VALUE, -9999.0), (N.NODATA_NAN, 0.0)]:
        params = N.KernelParams(64, 33, mode, value, 0)
        want = oracle_grid(orc, kernel, xs, ys, heights, params)
        got = heights.copy()
        hip.run(N.TG_KERNEL_NOISE_LAYER, kernel, got, xs, ys, params)
        assert np.array_equal(np.isnan(got), np.isnan(want))
        m = ~np.isnan(want)
        assert np.array_equal(got[m].view(np.uint32), want[m].view(np.uint32))
        if mode == N.NODATA_VALUE:
            assert np.all(got[heights == -9999.0] == -9999.0)


def test_strips_join_the_whole_grid(hip):
    """Tiles equal the whole (tests/unit/layers/noise_layer_test.cpp:200-244): running the grid
    as strips of rows gives the same bits as one run."""
    kernel = N.noise_kernel([C2_FBM, dict(seed=2, amplitude=3.0, basis=1)])
    xs, ys = grid_axes(300, 150, 0.5)
    whole = np.zeros((150, 300), np.float32)
    hip.run(N.TG_KERNEL_NOISE_LAYER, kernel, whole, xs, ys, N.KernelParams(300, 150, 0, 0.0, 0))
    parts = np.zeros((150, 300), np.float32)
    for r0 in range(0, 150, 37):
        r1 = min(150, r0 + 37)
        strip = np.ascontiguousarray(parts[r0:r1])
        hip.run(N.TG_KERNEL_NOISE_LAYER, kernel, strip, xs, ys[r0:r1].copy(),
                N.KernelParams(300, r1 - r0, 0, 0.0, r0))
        parts[r0:r1] = strip
    assert np.array_equal(whole.view(np.uint32), parts.view(np.uint32))


def test_full_size_c2_noise_against_fp64(hip, orc):
    """The 4097 x 4097 synthetic source at 0.5 m (origin {0, 0}) through the resident path:
    sampled rows bit-exact to the FP32 oracle, and every sampled cell within 1e-3 m of the FP64
    oracle."""
    kernel = N.noise_kernel([C2_FBM])
    n = 4097
    xs, ys = grid_axes(n, n, 0.5)
    err = C.create_string_buffer(512)
    plan = C.c_void_p()
    assert hip.lib.tg_hip_noise_plan_create(hip.handle, C.byref(kernel), C.byref(plan), err, 512) == 0
    try:
        dh = N.DeviceArray(hip, n * n * 4)
        dh.zero()
        dx = N.DeviceArray.from_host(hip, xs)
        dy = N.DeviceArray.from_host(hip, ys)
        params = N.KernelParams(n, n, 0, 0.0, 0)
        code = hip.lib.tg_hip_noise_launch(hip.handle, plan, dh.ptr, dx.ptr, dy.ptr,
                                           C.byref(params), None, err, 512)
        assert code == 0, err.value
        hip.synchronize()
        got = dh.download(np.empty((n, n), np.float32))
    finally:
        hip.lib.tg_hip_noise_plan_free(plan)
    rows = np.array([0, 1, 777, 2048, 3001, n - 1])
    sub_ys = ys[rows].copy()
    want = np.zeros((rows.size, n), np.float32)
    orc.or_noise_grid(C.byref(kernel), xs.ctypes.data, sub_ys.ctypes.data, n, rows.size,
                      C.byref(N.KernelParams(n, rows.size, 0, 0.0, 0)), want.ctypes.data, 0)
    assert np.array_equal(got[rows].view(np.uint32), want.view(np.uint32))
    want64 = np.zeros((rows.size, n), np.float64)
    xs64, ys64 = xs.astype(np.float64), sub_ys.astype(np.float64)  # kept alive across the call
    orc.or_noise_grid_f64(C.byref(kernel), xs64.ctypes.data, ys64.ctypes.data, n, rows.size,
                          want64.ctypes.data, 0)
    d = np.abs(got[rows] - want64)
    print(f"[noise C2] vs FP64: max {d.max():.3g} m, rms {np.sqrt((d**2).mean()):.3g} m", flush=True)
    assert d.max() <= 2e-4  # FP32 rounding of eight octaves at 20 m (1.8e-5 m measured)
    assert np.isfinite(got).all() and got.std() > 1.0
