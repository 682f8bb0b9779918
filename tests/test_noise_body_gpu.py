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


MIXED_BASES = [
    dict(C2_FBM, warp_amplitude=40.0, warp_wavelength=800.0),
    dict(seed=11, basis=1, wavelength=150.0, octaves=6, amplitude=4.0),
    dict(seed=12, basis=3, wavelength=60.0, octaves=3, amplitude=2.0, blend=1,
         warp_amplitude=10.0, warp_wavelength=90.0),
    dict(seed=13, basis=2, type=2, wavelength=250.0, octaves=4, blend=3),
    dict(seed=14, basis=6, wavelength=30.0, octaves=2, amplitude=1.5, blend=2),
]


@pytest.mark.parametrize("fractals", [
    MIXED_BASES,
    [dict(seed=21, basis=4, type=1, offset=0.8, wavelength=70.0, octaves=3, amplitude=5.0),
     dict(seed=22, basis=5, type=3, offset=0.6, wavelength=45.0, octaves=3, amplitude=2.0,
          warp_amplitude=25.0, warp_wavelength=300.0)],
])
def test_body_noise_every_basis_and_warp_bit_exact(hip, orc, fractals):
    """Simplex, value, Worley F1/F2/F2-F1 and Voronoi bases and 3D domain warp (DomainWarp3,
    include/terraingen/noise/fractals.h:221-243) on a lunar-radius body, bit for bit."""
    kernel = N.noise_kernel(fractals)
    w, h = 61, 43
    pos = lonlat_positions(-77.3, 12.9, 0.0004, w, h)
    base = np.full((h, w), -2.0, np.float32)
    params = N.KernelParams(w, h, N.NODATA_NONE, 0.0, 0)
    got = base.copy()
    hip.run(N.TG_KERNEL_NOISE_BODY, kernel, got, np.zeros(w, np.float32),
            np.zeros(h, np.float32), params, [(pos, 1)])
    want = oracle(orc, kernel, pos, base, params)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
    assert np.ptp(want) > 1.0
