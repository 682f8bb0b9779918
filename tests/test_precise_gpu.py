# SPDX-License-Identifier: BSD-3-Clause
"""Kilometre-scale craters: pairs with craters of 1 km and more are evaluated in FP64 (the FP32
arithmetic of the reference rounds to about 1.5e-7 D of height, 2e-3 m at 20 km). Tiles on the
Moon around 5-20 km craters, and plane tiles among 1-20 km craters, are bit-equal to the plain
per-tile path and within 1e-3 m of the FP64 oracle, closer to it than the FP32 oracle is."""
import numpy as np
import pytest

from terraingen_amd import native as N
from terraingen_amd.pipeline import CraterLayerSpec, SphereTileBatch, TileBatch, c4_tiles

from batch_oracle import plane_tile, sphere_tile

pytestmark = pytest.mark.gpu

TOL_FP64 = 1e-3
C2_FRACTAL = dict(seed=1, wavelength=512.0, amplitude=20.0, octaves=8, lacunarity=2.0, gain=0.5)
MOON = 1737400.0


def run_batch(hip, b, use_graph=True):
    out = N.DeviceArray(hip, b.cells * 4)
    b.run(out, use_graph)
    b.wait()
    return out.download(np.empty(b.cells, np.float32))


def check(hip, orc, b, oracle_of, label, closer_than_fp32=True):
    got = run_batch(hip, b)
    worst64 = worst_ref = 0.0
    for i in range(b.tiles.size):
        f32, f64, inp = oracle_of(orc, b, i)
        mine = b.tile_heights(got, i)
        alone = b.plain_run(inp)
        assert np.array_equal(mine.view(np.uint32), alone.view(np.uint32)), i
        worst64 = max(worst64, float(np.abs(mine.astype(np.float64) - f64).max()))
        worst_ref = max(worst_ref, float(np.abs(f32.astype(np.float64) - f64).max()))
    print(f"[{label}] bit-equal to the plain path; vs FP64 oracle max {worst64:.3g} m (the "
          f"reference's FP32 arithmetic: {worst_ref:.3g} m)", flush=True)
    assert worst64 <= TOL_FP64
    if closer_than_fp32:
        assert worst64 < worst_ref


def moon_craters():
    """Twelve craters of 5-20 km at mid latitudes, MPD with freshness and irregular outlines."""
    rng = np.random.default_rng(14)
    c = np.zeros(12, N.CRATER_DTYPE)
    c["x"] = rng.uniform(10.0, 14.0, c.size)
    c["y"] = rng.uniform(-3.0, 3.0, c.size)
    c["diameter"] = np.exp(rng.uniform(np.log(5000.0), np.log(20000.0), c.size))
    c["age_gyr"] = rng.uniform(0.0, 1.0, c.size)
    c["id"] = np.arange(1, c.size + 1, dtype=np.uint64)
    return c


def test_km_craters_on_the_moon(hip, orc):
    craters = moon_craters()
    b = SphereTileBatch(hip, [C2_FRACTAL], CraterLayerSpec(seed=4), region=None, radius=MOON,
                        craters_given=craters)
    b.build()
    unit = MOON * np.pi / 180.0
    rng = np.random.default_rng(3)
    tiles = []
    # Tiles astride each crater's rim and one on its floor, 0.5-8 m a cell.
    for c in craters[:6]:
        r = 0.5 * float(c["diameter"])
        for radii in (0.45, 0.97, 1.6):
            ang = rng.uniform(0.0, 2.0 * np.pi)
            n = int(2 ** rng.integers(5, 8) + 1)
            step = float(np.exp(rng.uniform(np.log(0.5), np.log(8.0)))) / unit
            lon = float(c["x"]) + radii * r * np.cos(ang) / unit / np.cos(np.radians(float(c["y"])))
            lat = float(c["y"]) + radii * r * np.sin(ang) / unit
            tiles.append((lon - 0.5 * n * step, lat + 0.5 * n * step, step, n, n))
    arr = np.zeros(len(tiles), N.TILE_DTYPE)
    for j, t in enumerate(tiles):
        arr[j] = t
    b.stage(arr)
    try:
        check(hip, orc, b, sphere_tile, "km craters on the Moon")
    finally:
        b.close()


def test_km_craters_on_a_plane(hip, orc):
    """Craters only: 60 km from the origin the plane noise's FP32 positions (the synthetic
    source's origin is {0, 0}, synthetic_source.cpp:76-80) round to 4 mm, which the noise keeps
    bit for bit from the reference."""
    spec = CraterLayerSpec(min_diameter=1000.0, max_diameter=20000.0, seed=4)
    b = TileBatch(hip, [], spec, region=(0.0, 0.0, 60000.0, 60000.0), finest_pixel=1.0)
    b.build()
    assert b.records.size > 50 and (b.records["diameter"] >= 5000.0).any()
    b.stage(c4_tiles(16, 60000.0, seed=6, min_k=5, max_k=7, pixel_range=(1.0, 8.0)))
    try:
        check(hip, orc, b, plane_tile, "km craters on a plane")
    finally:
        b.close()


def test_craters_large_against_a_small_body(hip, orc):
    """Craters of 4-7 km on a body of 10 km radius: cells up to 2.4 crater radii out lie more than
    half the body's radius from the centre along the chord, where the FP64 offset's arcsine ratio
    leaves its series and halves the angle instead. Bit-equal to the plain path and within the
    bound of the FP64 oracle (at this size the FP32 arithmetic is as close, so only the bound is
    asserted)."""
    radius = 10000.0
    rng = np.random.default_rng(21)
    c = np.zeros(6, N.CRATER_DTYPE)
    c["x"] = np.arange(6) * 55.0 + rng.uniform(0.0, 10.0, 6)
    c["y"] = rng.uniform(-20.0, 20.0, 6)
    c["diameter"] = rng.uniform(4000.0, 7000.0, 6)
    c["age_gyr"] = rng.uniform(0.0, 1.0, 6)
    c["id"] = np.arange(1, 7, dtype=np.uint64)
    b = SphereTileBatch(hip, [C2_FRACTAL], CraterLayerSpec(seed=4), region=None, radius=radius,
                        craters_given=c)
    b.build()
    unit = radius * np.pi / 180.0
    tiles = []
    for k in c:
        r = 0.5 * float(k["diameter"])
        for radii in (0.97, 1.6, 2.4):
            ang = rng.uniform(0.0, 2.0 * np.pi)
            n = 33
            step = 4.0 / unit
            lon = float(k["x"]) + radii * r * np.cos(ang) / unit / np.cos(np.radians(float(k["y"])))
            lat = float(k["y"]) + radii * r * np.sin(ang) / unit
            tiles.append((lon - 0.5 * n * step, lat + 0.5 * n * step, step, n, n))
    arr = np.zeros(len(tiles), N.TILE_DTYPE)
    for j, t in enumerate(tiles):
        arr[j] = t
    b.stage(arr)
    try:
        check(hip, orc, b, sphere_tile, "craters large against a 10 km body",
              closer_than_fp32=False)
    finally:
        b.close()
