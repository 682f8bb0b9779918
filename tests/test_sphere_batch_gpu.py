# SPDX-License-Identifier: BSD-3-Clause
"""Batched tiles on a sphere (tg_hip_batch_surface): one persistent launch over longitude-latitude
tiles of mixed size and cell size gives, tile for tile, the bits of the plain per-tile path on a
body (tg_hip_run of the body noise kernel, then the crater kernel's body branch) and, within the
stated tolerance, the FP64 oracle and the reference's own heights on a sphere."""
import ctypes as C

import numpy as np
import pytest

from terraingen_amd import native as N
from terraingen_amd.pipeline import (CraterLayerSpec, SphereTileBatch, global_craters,
                                     sphere_tiles)

from crater_cases import SPHERE_NAMES, load_sphere, sphere_oracle_heights_f64

pytestmark = pytest.mark.gpu

TOL_FP64 = 1e-3  # max |dh| against the FP64 oracle, meters
C2_FRACTAL = dict(seed=1, wavelength=512.0, amplitude=20.0, octaves=8, lacunarity=2.0, gain=0.5)
BODY = 20000.0  # a 20 km body: craters of 110 m to 2 km cover tiles of a few meters to a km


def run_batch(hip, b, use_graph):
    out = N.DeviceArray(hip, b.cells * 4)
    b.run(out, use_graph)
    b.wait()
    return out.download(np.empty(b.cells, np.float32))


@pytest.fixture(scope="module")
def batch(hip):
    craters = global_craters(20000, BODY, max_diameter=2000.0, seed=6)
    b = SphereTileBatch(hip, [C2_FRACTAL], CraterLayerSpec(seed=4), region=None, radius=BODY,
                        craters_given=craters)
    b.build()
    tiles = sphere_tiles(24, BODY, seed=9, min_k=4, max_k=7, pixel_range=(0.25, 50.0))
    # One tile across the antimeridian and one at the north pole's edge.
    unit = BODY * np.pi / 180.0
    tiles[0] = (179.9, 10.0, 2.0 / unit, 65, 65)
    tiles[1] = (-30.0, 90.0, 5.0 / unit, 33, 33)
    b.stage(tiles)
    print(f"[sphere batch] {tiles.size} tiles, {b.cells} cells, {b.records.size} craters, "
          f"{b.entries.size} entries", flush=True)
    yield b
    b.close()


def test_sphere_batch_equals_each_tile_alone(hip, batch):
    got = run_batch(hip, batch, use_graph=False)
    shaped = 0
    for i in range(batch.tiles.size):
        inp = batch.plain_inputs(i)
        alone = batch.plain_run(inp)
        mine = batch.tile_heights(got, i)
        assert np.array_equal(mine.view(np.uint32), alone.view(np.uint32)), i
        shaped += inp["craters"] is not None
    assert shaped >= batch.tiles.size // 2, shaped
    print(f"[sphere batch] bit-equal to the plain body path, {shaped} tiles with craters",
          flush=True)


def test_sphere_graph_replays_match_eager(hip, batch):
    eager = run_batch(hip, batch, use_graph=False)
    for _ in range(3):
        again = run_batch(hip, batch, use_graph=True)
        assert np.array_equal(again.view(np.uint32), eager.view(np.uint32))
        hip.synchronize()


def test_sphere_hoisted_path_is_bit_identical(hip, batch):
    batch.tune(hoisted=False)
    full = run_batch(hip, batch, use_graph=False)
    batch.tune(hoisted=True)
    hoisted = run_batch(hip, batch, use_graph=False)
    assert np.array_equal(full.view(np.uint32), hoisted.view(np.uint32))


@pytest.mark.parametrize("fractals", [
    [C2_FRACTAL],
    [dict(seed=7, wavelength=40.0, amplitude=3.0, octaves=12, lacunarity=2.2, gain=0.55),
     dict(seed=3, type=1, offset=0.9, wavelength=900.0, octaves=6, blend=1)],
    [dict(C2_FRACTAL, warp_amplitude=40.0, warp_wavelength=800.0),
     dict(seed=11, basis=1, wavelength=150.0, octaves=6, amplitude=4.0),
     dict(seed=12, basis=3, wavelength=60.0, octaves=3, amplitude=2.0, blend=1,
          warp_amplitude=10.0, warp_wavelength=90.0),
     dict(seed=13, basis=2, type=2, wavelength=250.0, octaves=4, blend=3),
     dict(seed=14, basis=6, wavelength=30.0, octaves=2, amplitude=1.5, blend=2)],
])
def test_sphere_noise_bit_exact_across_cell_sizes(hip, orc, fractals):
    """Tiles on the Moon from 3 cm to 60 m a cell, noise only, so an octave's lattice cells per
    8 x 8 square range from far under one (corner hashes shared across the wave) to many (hashed
    per lane); each tile's heights equal the oracle's bit for bit, for Perlin stacks and for a
    warped stack of every other basis."""
    moon = 1737400.0
    b = SphereTileBatch(hip, fractals, CraterLayerSpec(seed=4), region=None, radius=moon,
                        craters_given=global_craters(1, moon, max_diameter=2000.0, seed=1))
    b.build()
    tiles = sphere_tiles(16, moon, seed=5, min_k=3, max_k=7, pixel_range=(0.03, 60.0))
    b.stage(tiles)
    try:
        got = run_batch(hip, b, use_graph=False)
        for i in range(tiles.size):
            inp = b.plain_inputs(i)
            g = inp["grid"]
            want = np.zeros((g.height, g.width), np.float32)
            params = N.KernelParams(g.width, g.height, 0, 0.0, 0)
            pos = inp["positions"]
            orc.or_body_noise_grid(C.byref(b.noise_kernel), pos.ctypes.data, pos.shape[0],
                                   C.byref(params), want.ctypes.data, 0)
            if inp["craters"] is not None:
                continue  # the one crater reaches this tile
            mine = b.tile_heights(got, i)
            assert np.array_equal(mine.view(np.uint32), want.view(np.uint32)), (
                i, float(tiles["pixel"][i]) * moon * np.pi / 180.0, int(g.width))
    finally:
        b.close()


def frames64(lon, lat, radius):
    """MakeCraterFrame in FP64 (centre, east, north, radius; 10 values per crater)."""
    lam, phi = np.radians(lon), np.radians(lat)
    centre = radius * np.stack([np.cos(phi) * np.cos(lam), np.cos(phi) * np.sin(lam),
                                np.sin(phi)], axis=1)
    east = np.stack([-np.sin(lam), np.cos(lam), np.zeros_like(lam)], axis=1)
    north = np.stack([-np.sin(phi) * np.cos(lam), -np.sin(phi) * np.sin(lam), np.cos(phi)],
                     axis=1)
    return np.ascontiguousarray(np.concatenate([centre, east, north,
                                                np.full((lam.size, 1), radius)], axis=1))


def test_sphere_batch_matches_oracles(hip, orc, batch):
    """Each tile against the FP32 oracle (the reference's CPU arithmetic) and the FP64 oracle
    (positions and frames in FP64, FP64 noise and crater shaping)."""
    got = run_batch(hip, batch, use_graph=True)
    worst64 = worst32 = 0.0
    for i in range(batch.tiles.size):
        inp = batch.plain_inputs(i)
        g = inp["grid"]
        params = N.KernelParams(g.width, g.height, 0, 0.0, 0)
        f32 = np.zeros((g.height, g.width), np.float32)
        f64 = np.zeros((g.height, g.width), np.float64)
        pos = inp["positions"]
        orc.or_body_noise_grid(C.byref(batch.noise_kernel), pos.ctypes.data, pos.shape[0],
                               C.byref(params), f32.ctypes.data, 0)
        orc.or_body_noise_grid_f64(C.byref(batch.noise_kernel), pos.ctypes.data, pos.shape[0],
                                   f64.ctypes.data, 0)
        if inp["craters"] is not None:
            kernel, records, data, tiles, entries, frames, points = inp["craters"]
            orc.or_shape_grid_body(C.byref(kernel), records.ctypes.data, data.ctypes.data,
                                   tiles.ctypes.data, entries.ctypes.data, frames.ctypes.data,
                                   points.ctypes.data, g.width, g.height, C.byref(params),
                                   f32.ctypes.data, 0)
            foot = inp["foot"]
            fr = frames64(foot[:, 0], foot[:, 1], batch.radius)
            orc.or_shape_grid_body_f64(C.byref(kernel), records.ctypes.data, data.ctypes.data,
                                       tiles.ctypes.data, entries.ctypes.data, fr.ctypes.data,
                                       pos.ctypes.data, g.width, g.height, 0, f64.ctypes.data, 0)
        mine = batch.tile_heights(got, i)
        worst32 = max(worst32, float(np.abs(mine - f32).max()))
        worst64 = max(worst64, float(np.abs(mine.astype(np.float64) - f64).max()))
    print(f"[sphere batch] vs FP32 oracle max {worst32:.3g} m, vs FP64 oracle max "
          f"{worst64:.3g} m", flush=True)
    assert worst64 <= TOL_FP64
    assert worst32 <= TOL_FP64


@pytest.mark.parametrize("name", SPHERE_NAMES)
def test_sphere_batch_matches_reference_heights(hip, orc, name):
    """The reference's crater layers on a sphere (records, frames and heights from its own CPU
    pipeline) as a one-tile batch with every crater in every bin: within the tolerance of the
    reference's heights and of the FP64 oracle. Cell positions here are the library's own, an ulp
    from the reference's."""
    case = load_sphere(name)
    n = case.records.size
    diam = np.ascontiguousarray(case.records["diameter"].astype(np.float64))
    cat = C.c_void_p()
    hip.call("tg_hip_catalogue_create", hip.handle, C.byref(case.kernel),
             case.records.ctypes.data, n, case.data.ctypes.data, case.data.size,
             diam.ctypes.data, C.byref(cat))
    b = C.c_void_p()
    try:
        hip.call("tg_hip_catalogue_frames", cat, case.frames.ctypes.data, n)
        across = (case.side + 15) // 16
        descs = np.zeros(1, N.TILE_DESC_DTYPE)
        descs[0] = (case.lon, case.lat, case.step, case.side, case.side, 0, 0, across)
        bins = np.zeros(across * across, N.BIN_DTYPE)
        bins["count"] = n
        entries = np.arange(n, dtype=np.uint32)
        hip.call("tg_hip_batch_create", hip.handle, 1, bins.size, n, C.byref(b))
        hip.call("tg_hip_batch_surface", b, case.radius)
        hip.call("tg_hip_batch_stage", b, descs.ctypes.data, 1, bins.ctypes.data, bins.size,
                 entries.ctypes.data, n)
        out = N.DeviceArray(hip, case.side * case.side * 4)
        hip.call("tg_hip_batch_run", b, None, 0.0, cat, out.ptr, case.side * case.side, 0)
        hip.call("tg_hip_batch_wait", b)
        got = out.download(np.empty(case.shape, np.float32))
    finally:
        if b.value:
            hip.lib.tg_hip_batch_free(b)
        hip.lib.tg_hip_catalogue_free(cat)
    d_ref = np.abs(got - case.heights)
    d64 = np.abs(got.astype(np.float64) - sphere_oracle_heights_f64(orc, case))
    print(f"[sphere batch {name}] vs reference max {d_ref.max():.3g} m, vs FP64 oracle max "
          f"{d64.max():.3g} m", flush=True)
    assert d_ref.max() <= TOL_FP64 and d64.max() <= TOL_FP64
    assert np.abs(case.heights).max() > 1.0  # the fixture is shaped


def test_sphere_batch_refuses_what_it_cannot_run(hip, batch):
    out = N.DeviceArray(hip, batch.cells * 4)
    # A catalogue without frames cannot shape tiles on a sphere.
    cat = C.c_void_p()
    diam = np.ascontiguousarray(batch.foot[:, 3])
    hip.call("tg_hip_catalogue_create", hip.handle, C.byref(batch.kernel),
             batch.records.ctypes.data, batch.records.size, batch.data.ctypes.data,
             batch.data.size, diam.ctypes.data, C.byref(cat))
    try:
        with pytest.raises(N.HipError, match="frames"):
            hip.call("tg_hip_batch_run", batch.batch, None, 0.0, cat, out.ptr, batch.cells, 0)
        with pytest.raises(N.HipError, match="one frame per crater"):
            hip.call("tg_hip_catalogue_frames", cat, batch.frames.ctypes.data, 3)
    finally:
        hip.lib.tg_hip_catalogue_free(cat)
    with pytest.raises(N.HipError, match="already has its frames"):
        hip.call("tg_hip_catalogue_frames", batch.catalogue, batch.frames.ctypes.data,
                 batch.frames.size)


def test_graph_recaptures_after_frames_are_attached(hip, batch):
    """A graph captured on a batch whose bins list no crater, before the catalogue had frames, is
    not replayed once frames are attached and a batch with craters is staged: the run re-captures
    and equals the eager path."""
    cat = C.c_void_p()
    diam = np.ascontiguousarray(batch.foot[:, 3])
    hip.call("tg_hip_catalogue_create", hip.handle, C.byref(batch.kernel),
             batch.records.ctypes.data, batch.records.size, batch.data.ctypes.data,
             batch.data.size, diam.ctypes.data, C.byref(cat))
    b = C.c_void_p()
    try:
        near = batch.tiles[:6].copy()
        descs, bins, entries = N.sphere_bin(batch.foot, batch.radius, near, 1)
        assert entries.size > 0
        empty = bins.copy()
        empty["first"] = 0
        empty["count"] = 0
        hip.call("tg_hip_batch_create", hip.handle, descs.size, bins.size, entries.size,
                 C.byref(b))
        hip.call("tg_hip_batch_surface", b, batch.radius)
        cells = int(sum(int(d["width"]) * int(d["height"]) for d in descs))
        out = N.DeviceArray(hip, cells * 4)
        one = np.zeros(1, np.uint32)
        hip.call("tg_hip_batch_stage", b, descs.ctypes.data, descs.size, empty.ctypes.data,
                 empty.size, one.ctypes.data, 0)
        hip.call("tg_hip_batch_run", b, batch.plan, 0.0, cat, out.ptr, cells, 1)
        hip.call("tg_hip_batch_wait", b)
        hip.call("tg_hip_catalogue_frames", cat, batch.frames.ctypes.data, batch.frames.size)
        hip.call("tg_hip_batch_stage", b, descs.ctypes.data, descs.size, bins.ctypes.data,
                 bins.size, entries.ctypes.data, entries.size)
        hip.call("tg_hip_batch_run", b, batch.plan, 0.0, cat, out.ptr, cells, 0)
        hip.call("tg_hip_batch_wait", b)
        eager = out.download(np.empty(cells, np.float32))
        for _ in range(2):
            hip.call("tg_hip_batch_run", b, batch.plan, 0.0, cat, out.ptr, cells, 1)
            hip.call("tg_hip_batch_wait", b)
            graphed = out.download(np.empty(cells, np.float32))
            assert np.array_equal(graphed.view(np.uint32), eager.view(np.uint32))
    finally:
        if b.value:
            hip.lib.tg_hip_batch_free(b)
        hip.lib.tg_hip_catalogue_free(cat)
