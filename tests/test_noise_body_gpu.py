# SPDX-License-Identifier: BSD-3-Clause
"""Noise on a body on the MI355X (tg_hip_run(TG_KERNEL_NOISE_BODY): EvaluateFractals3 at each
cell's body-fixed position, src/noise/surface_noise.cpp:9-20) against the oracle, which is itself
pinned bit for bit to the reference's Fractal3: bit-exact, since the lattice coordinates are
split in FP64 as the reference splits them and the rest is integer hashing and FP32 in the
reference's order with contraction off."""
import ctypes as C

import numpy as np
import pytest

from terraingen_amd import native as N

pytestmark = pytest.mark.gpu

MOON = 1737400.0
C2_FBM = dict(seed=1, wavelength=512.0, amplitude=20.0, octaves=8, lacunarity=2.0, gain=0.5)


def lonlat_positions(lon0, lat0, step, width, height, radius=MOON):
    """Cell centres of a north-up longitude-latitude raster, as Surface::Position places them."""
    lon = lon0 + (np.arange(width) + 0.5) * step
    lat = lat0 - (np.arange(height) + 0.5) * step
    lam, phi = np.meshgrid(np.radians(lon), np.radians(lat))
    pos = radius * np.stack([np.cos(phi) * np.cos(lam), np.cos(phi) * np.sin(lam), np.sin(phi)],
                            axis=-1)
    return np.ascontiguousarray(pos.reshape(-1, 3))


def oracle(orc, kernel, pos, base, params):
    out = base.copy()
    orc.or_body_noise_grid(C.byref(kernel), pos.ctypes.data, pos.shape[0], C.byref(params),
                           out.ctypes.data, 0)
    return out


@pytest.mark.parametrize("fractals", [
    [C2_FBM],
    [dict(seed=7, wavelength=40.0, amplitude=3.0, octaves=12, lacunarity=2.2, gain=0.55),
     dict(seed=3, type=1, offset=0.9, wavelength=900.0, octaves=6, blend=1)],
    [dict(seed=4, type=2, wavelength=300.0, octaves=5, amplitude=7.0),
     dict(seed=5, type=3, offset=0.7, wavelength=2000.0, octaves=9, blend=3)],
])
def test_body_noise_bit_exact(hip, orc, fractals):
    kernel = N.noise_kernel(fractals)
    w, h = 131, 57
    pos = lonlat_positions(23.4, -41.2, 0.0007, w, h)
    base = np.full((h, w), 1.5, np.float32)
    params = N.KernelParams(w, h, N.NODATA_NONE, 0.0, 0)
    got = base.copy()
    hip.run(N.TG_KERNEL_NOISE_BODY, kernel, got, np.zeros(w, np.float32),
            np.zeros(h, np.float32), params, [(pos, 1)])
    want = oracle(orc, kernel, pos, base, params)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))


def test_body_noise_skips_no_data_and_joins_strips(hip, orc):
    kernel = N.noise_kernel([C2_FBM])
    w, h = 97, 77
    pos = lonlat_positions(-120.0, 61.0, 0.002, w, h)
    base = np.zeros((h, w), np.float32)
    base[::6, ::5] = -9999.0
    params = N.KernelParams(w, h, N.NODATA_VALUE, -9999.0, 0)
    want = oracle(orc, kernel, pos, base, params)
    grid = N.Grid(w, h, -120.0, 61.0, 0.002)
    got = base.copy()
    assert N.run_kernel(N.backend_named("hip"), N.TG_KERNEL_NOISE_BODY, kernel, grid, got,
                        (0.0, 0.0), [(pos, 1)], no_data=-9999.0, max_strip_bytes=w * 4 * 13)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
    assert np.all(got[::6, ::5] == -9999.0)


def test_body_noise_refuses_warped_fractals(hip):
    kernel = N.noise_kernel([dict(C2_FBM, warp_amplitude=30.0)])
    w, h = 8, 8
    pos = lonlat_positions(0.0, 0.0, 0.001, w, h)
    with pytest.raises(N.HipError, match="unwarped Perlin"):
        hip.run(N.TG_KERNEL_NOISE_BODY, kernel, np.zeros((h, w), np.float32),
                np.zeros(w, np.float32), np.zeros(h, np.float32),
                N.KernelParams(w, h, N.NODATA_NONE, 0.0, 0), [(pos, 1)])
