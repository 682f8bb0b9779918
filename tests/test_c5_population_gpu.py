# SPDX-License-Identifier: BSD-3-Clause
"""The served stream's crater population (the dense field's: N(>D) = 0.0156 D^-2.2 per m^2, D
0.5-300 m, complex from 50 m with central peaks, wall terraces from 150 m, full at 300 m) through
the batched kernels and the tile server, on a plane and on the Moon. Complex craters send the
batches to the general kernel variant (ApplyCrater's full evaluation, craters.h:1186-1211, with
KernelPeakRelief and KernelTerracedRadii, craters.h:979-1024,1128-1155). Each tile is bit-equal
to the plain per-tile path and within 1e-3 m of the FP64 oracle; each served tile is the same
alone, in a mix and through the batched path."""
import threading

import numpy as np
import pytest

from terraingen_amd import native as N
from terraingen_amd.pipeline import (CraterLayerSpec, SphereTileBatch, TileBatch, TileServer,
                                     c4_tiles)
from terraingen_amd.serving import GreatCircleStream

from batch_oracle import features_in_reach, plane_tile, sphere_tile

pytestmark = pytest.mark.gpu

TOL_FP64 = 1e-3  # max |dh| against the FP64 oracle, meters
C2_FRACTAL = dict(seed=1, wavelength=512.0, amplitude=20.0, octaves=8, lacunarity=2.0, gain=0.5)
C3_CRATERS = CraterLayerSpec(sfd_coefficient=0.0156, sfd_exponent=2.2, min_diameter=0.5,
                             max_diameter=300.0, complex_=1, transition_diameter=50.0, peaks=1,
                             terraces=1, terrace_onset=150.0, terrace_full=300.0, seed=5)
MOON = 1737400.0
SIDE = 1500.0  # m, the plane catalogue's square: with seed 5 it holds a 287 m terraced crater


def run_batch(hip, b):
    out = N.DeviceArray(hip, b.cells * 4)
    b.run(out, use_graph=True)
    b.wait()
    return out.download(np.empty(b.cells, np.float32))


def on_biggest(craters, n, pixel, x_of, y_of):
    """Tiles of n x n cells of `pixel` centred on the largest complex craters (the ones with peaks
    and terraces)."""
    order = np.argsort(craters["diameter"])[::-1][:3]
    out = []
    for k in order:
        x, y = x_of(craters[k]), y_of(craters[k])
        out.append((x - 0.5 * n * pixel, y + 0.5 * n * pixel, pixel, n, n))
    return out


@pytest.fixture(scope="module")
def plane(hip):
    b = TileBatch(hip, [C2_FRACTAL], C3_CRATERS, region=(0.0, 0.0, SIDE, SIDE),
                  finest_pixel=0.125)
    b.build()
    tiles = c4_tiles(14, SIDE, seed=21, min_k=4, max_k=7, pixel_range=(0.25, 2.0))
    extra = on_biggest(b.crater_list, 129, 2.0, lambda c: float(c["x"]),
                       lambda c: float(c["y"]))
    arr = np.zeros(tiles.size + len(extra), N.TILE_DTYPE)
    arr[:tiles.size] = tiles
    for j, t in enumerate(extra):
        arr[tiles.size + j] = t
    b.stage(arr)
    print(f"[c5 plane] {arr.size} tiles, {b.cells} cells, {b.records.size} craters "
          f"({int((b.records['peak_count'] > 0).sum())} with peaks, "
          f"{int((b.records['terrace_count'] > 0).sum())} terraced)", flush=True)
    yield b
    b.close()


def swath_stream():
    return GreatCircleStream(path_length=600.0, radius=MOON)


@pytest.fixture(scope="module")
def sphere(hip):
    stream = swath_stream()
    craters = stream.swath_craters(300.0, coefficient=0.0156, exponent=2.2, min_diameter=0.5,
                                   max_diameter=300.0, margin=300.0, seed=7)
    # A 600 m swath seldom draws a crater of 150 m or more: three terraced ones are added.
    big = np.zeros(3, N.CRATER_DTYPE)
    big["x"], big["y"] = stream.track(np.array([150.0, 300.0, 450.0]),
                                      np.array([-100.0, 60.0, 120.0]))
    big["diameter"] = (170.0, 240.0, 290.0)
    big["age_gyr"] = (0.2, 0.05, 0.6)
    big["id"] = craters["id"].max() + np.arange(1, 4, dtype=np.uint64)
    craters = np.concatenate([craters, big])
    b = SphereTileBatch(hip, [C2_FRACTAL], C3_CRATERS, region=None, finest_pixel=0.125,
                        radius=MOON, craters_given=craters)
    b.build()
    unit = MOON * np.pi / 180.0
    rng = np.random.default_rng(8)
    n = 14
    lon, lat = stream.track(rng.uniform(0.0, 600.0, n), rng.uniform(-250.0, 250.0, n))
    edge = 2 ** rng.integers(4, 8, n) + 1
    step = np.exp(rng.uniform(np.log(0.25), np.log(2.0), n)) / unit
    tiles = [(float(lon[k]), float(lat[k]), float(step[k]), int(edge[k]), int(edge[k]))
             for k in range(n)]
    tiles += on_biggest(craters, 129, 2.0 / unit, lambda c: float(c["x"]),
                        lambda c: float(c["y"]))
    arr = np.zeros(len(tiles), N.TILE_DTYPE)
    for j, t in enumerate(tiles):
        arr[j] = t
    b.stage(arr)
    print(f"[c5 sphere] {arr.size} tiles, {b.cells} cells, {b.records.size} craters", flush=True)
    yield b
    b.close()


def check_batch(hip, orc, b, oracle_of, label):
    got = run_batch(hip, b)
    worst64 = worst32 = 0.0
    seen = {"peaks": 0, "terraces": 0, "craters": 0}
    for i in range(b.tiles.size):
        f32, f64, inp = oracle_of(orc, b, i)
        mine = b.tile_heights(got, i)
        alone = b.plain_run(inp)
        assert np.array_equal(mine.view(np.uint32), alone.view(np.uint32)), i
        worst32 = max(worst32, float(np.abs(mine - f32).max()))
        worst64 = max(worst64, float(np.abs(mine.astype(np.float64) - f64).max()))
        for k, v in features_in_reach(inp).items():
            seen[k] += v
    print(f"[{label}] bit-equal to the plain path; vs FP32 oracle max {worst32:.3g} m, vs FP64 "
          f"oracle max {worst64:.3g} m; craters in reach {seen}", flush=True)
    assert seen["peaks"] > 0 and seen["terraces"] > 0, seen
    assert worst64 <= TOL_FP64
    assert worst32 <= TOL_FP64


def test_c3_population_plane_batch(hip, orc, plane):
    check_batch(hip, orc, plane, plane_tile, "c5 plane")


def test_c3_population_sphere_batch(hip, orc, sphere):
    check_batch(hip, orc, sphere, sphere_tile, "c5 sphere")


def serve_alone_and_mixed(hip, b):
    tiles = [(float(t["origin_x"]), float(t["origin_y"]), float(t["pixel"]), int(t["width"]),
              int(t["height"])) for t in b.tiles]
    server = TileServer(b, slots=3, max_batch_tiles=8, max_batch_cells=1 << 20,
                        pool_bytes=1 << 26)
    try:
        alone = [server.wait(server.submit(*t))[0] for t in tiles]
        tickets = [None] * len(tiles)

        def submitter(k):
            for i in range(k, len(tiles), 3):
                tickets[i] = server.submit(*tiles[i])

        threads = [threading.Thread(target=submitter, args=(k,)) for k in range(3)]
        for th in threads:
            th.start()
        for th in threads:
            th.join()
        sizes = []
        for i, ticket in enumerate(tickets):
            h, info = server.wait(ticket)
            sizes.append(info["batch_tiles"])
            assert np.array_equal(h.view(np.uint32), alone[i].view(np.uint32)), i
        stats = server.stats()
    finally:
        server.close()
    assert stats["failed"] == 0 and max(sizes) > 1
    assert stats["max_in_flight"] <= 3
    got = run_batch(hip, b)
    for i in range(len(tiles)):
        assert np.array_equal(b.tile_heights(got, i).view(np.uint32),
                              alone[i].view(np.uint32)), i


def test_c3_population_plane_server(hip, plane):
    serve_alone_and_mixed(hip, plane)


def test_c3_population_sphere_server(hip, sphere):
    serve_alone_and_mixed(hip, sphere)
