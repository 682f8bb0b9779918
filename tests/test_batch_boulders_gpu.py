# SPDX-License-Identifier: BSD-3-Clause
"""The boulder layer fused into the batched kernels (tg_hip_boulder_catalogue_create,
tg_hip_batch_stage_boulders): on a plane each tile is bit-equal to its noise, craters and then
boulders through the plain kernels (the boulders prepared at the tile's spacing, as the reference
prepares them per grid); on a plane and on a sphere the boulder stage is held to the FP64 oracle
(boulder_relief.cpp:44-57,244-262 with Surface::SphereOffset on the sphere) over the heights the
craters leave. The boulders are placed as the reference's BoulderLayer places them (C3's sizes,
N(>D) = 3.8e-3 D^-3 per m^2, D 0.2-3 m)."""
import ctypes as C

import numpy as np
import pytest

from terraingen_amd import native as N
from terraingen_amd.pipeline import CraterLayerSpec, SphereTileBatch, TileBatch, global_craters

pytestmark = pytest.mark.gpu

# The dome in FP32 against FP64: about ten times the 3.6e-7 m measured on the plane.
TOL_FP64 = 5e-6
C2_FRACTAL = dict(seed=1, wavelength=512.0, amplitude=20.0, octaves=8, lacunarity=2.0, gain=0.5)


def run(hip, b, use_graph=False):
    out = N.DeviceArray(hip, b.cells * 4)
    b.run(out, use_graph)
    b.wait()
    return out.download(np.empty(b.cells, np.float32))


@pytest.fixture(scope="module")
def plane(hip):
    region = (0.0, 0.0, 120.0, 120.0)
    placed = N.place_boulders(region, 0.25, 3.8e-3, 3.0, 0.2, 3.0, seed=2)
    b = TileBatch(hip, [C2_FRACTAL], CraterLayerSpec(seed=4), region=region,
                  boulders=N.relief_boulders(placed))
    b.build()
    tiles = np.zeros(6, N.TILE_DTYPE)
    tiles["width"] = (65, 33, 129, 40, 17, 97)
    tiles["height"] = (33, 65, 97, 40, 50, 81)
    tiles["pixel"] = (0.25, 0.3, 0.4, 0.7, 1.1, 0.5)
    tiles["origin_x"] = (10.0, 40.0, 60.0, 5.0, 90.0, 20.0)
    tiles["origin_y"] = (110.0, 100.0, 70.0, 40.0, 115.0, 60.0)
    b.stage(tiles)
    print(f"[plane boulders] {tiles.size} tiles, {b.cells} cells, {b.records.size} craters, "
          f"{b.boulder_records.size} boulders, {b.boulder_entries.size} boulder entries",
          flush=True)
    yield b
    b.close()


def test_plane_batch_with_boulders_equals_the_plain_layers(hip, plane):
    got = run(hip, plane)
    for i in range(plane.tiles.size):
        alone = plane.plain_run(plane.plain_inputs(i))
        mine = plane.tile_heights(got, i)
        assert np.array_equal(mine.view(np.uint32), alone.view(np.uint32)), i


def test_plane_boulder_stage_within_fp64_oracle(hip, orc, plane):
    got = run(hip, plane)
    worst, drawn = 0.0, 0
    for i in range(plane.tiles.size):
        inp = plane.plain_inputs(i)
        base = plane.plain_run(inp, with_boulders=False)
        want = base.astype(np.float64)
        g = inp["grid"]
        orc.or_boulder_grid_f64(plane.boulders.ctypes.data, plane.boulders.size,
                                C.byref(N.Grid(g.width, g.height, g.origin_x, g.origin_y,
                                               g.pixel)), 1, want.ctypes.data)
        mine = plane.tile_heights(got, i).astype(np.float64)
        worst = max(worst, float(np.abs(mine - want).max()))
        drawn += int((mine > base).sum())
    print(f"[plane boulders] {drawn} cells raised; max |dh| vs FP64 oracle {worst:.3g} m",
          flush=True)
    assert drawn > 150
    assert worst <= TOL_FP64


def test_plane_boulder_graph_replays_and_restage(hip, plane):
    eager = run(hip, plane)
    for _ in range(3):
        again = run(hip, plane, use_graph=True)
        assert np.array_equal(again.view(np.uint32), eager.view(np.uint32))
        hip.synchronize()
    # Staging the tiles again without their boulder lists drops the boulders, and the graph
    # captured with them is not replayed.
    plane.stage(plane.tiles, with_boulders=False)
    bare = run(hip, plane, use_graph=True)
    for i in range(plane.tiles.size):
        alone = plane.plain_run(plane.plain_inputs(i), with_boulders=False)
        assert np.array_equal(plane.tile_heights(bare, i).view(np.uint32),
                              alone.view(np.uint32)), i
    plane.stage(plane.tiles)
    assert np.array_equal(run(hip, plane, use_graph=True).view(np.uint32),
                          eager.view(np.uint32))


def test_boulder_lists_are_checked(hip, plane):
    bins = plane.boulder_bins.copy()
    entries = plane.boulder_entries.copy()
    bad = entries.copy()
    bad[0] = plane.boulder_records.size  # names a boulder the catalogue does not hold
    with pytest.raises(RuntimeError, match="boulder"):
        hip.call("tg_hip_batch_stage_boulders", plane.batch, plane.boulder_catalogue,
                 bins.ctypes.data, bins.size, bad.ctypes.data, bad.size, 0)
    with pytest.raises(RuntimeError, match="one bin per staged bin"):
        hip.call("tg_hip_batch_stage_boulders", plane.batch, plane.boulder_catalogue,
                 bins.ctypes.data, bins.size - 1, entries.ctypes.data, entries.size, 0)
    plane.stage(plane.tiles)


BODY = 20000.0  # a 20 km body, so a tile of a few meters spans a visible slice of a degree


@pytest.fixture(scope="module")
def sphere(hip):
    unit = BODY * np.pi / 180.0
    region = (30.0, -10.0 - 80.0 / unit, 30.0 + 80.0 / unit, -10.0)
    placed = N.place_boulders(region, 0.25, 3.8e-3, 3.0, 0.2, 3.0, seed=3, sphere_radius=BODY)
    craters = global_craters(20000, BODY, max_diameter=2000.0, seed=6)
    b = SphereTileBatch(hip, [C2_FRACTAL], CraterLayerSpec(seed=4), region=None, radius=BODY,
                        craters_given=craters, boulders=N.relief_boulders(placed))
    b.build()
    tiles = np.zeros(5, N.TILE_DTYPE)
    tiles["width"] = (65, 129, 33, 81, 49)
    tiles["height"] = (65, 97, 49, 81, 33)
    tiles["pixel"] = np.array([0.25, 0.3, 0.45, 0.6, 1.0]) / unit
    tiles["origin_x"] = 30.0 + np.array([5.0, 20.0, 45.0, 2.0, 30.0]) / unit
    tiles["origin_y"] = -10.0 - np.array([4.0, 30.0, 10.0, 50.0, 20.0]) / unit
    b.stage(tiles, with_boulders=False)
    base = run(hip, b)
    b.stage(tiles)
    print(f"[sphere boulders] {tiles.size} tiles, {b.cells} cells, {b.boulder_records.size} "
          f"boulders, {b.boulder_entries.size} boulder entries", flush=True)
    yield b, base
    b.close()


def test_sphere_boulder_stage_within_fp64_oracle(hip, orc, sphere):
    b, base = sphere
    got = run(hip, b)
    worst, drawn = 0.0, 0
    for i in range(b.tiles.size):
        t = b.tiles[i]
        grid = N.Grid(int(t["width"]), int(t["height"]), float(t["origin_x"]),
                      float(t["origin_y"]), float(t["pixel"]))
        under = b.tile_heights(base, i)
        want = under.astype(np.float64)
        orc.or_boulder_body_f64(b.boulders.ctypes.data, b.boulders.size, C.byref(grid), BODY,
                                1, want.ctypes.data)
        mine = b.tile_heights(got, i)
        worst = max(worst, float(np.abs(mine.astype(np.float64) - want).max()))
        drawn += int((mine > under).sum())
    print(f"[sphere boulders] {drawn} cells raised; max |dh| vs FP64 oracle {worst:.3g} m",
          flush=True)
    assert drawn > 150
    assert worst <= TOL_FP64


def test_sphere_boulder_graph_replays(hip, sphere):
    b, _ = sphere
    eager = run(hip, b)
    for _ in range(3):
        again = run(hip, b, use_graph=True)
        assert np.array_equal(again.view(np.uint32), eager.view(np.uint32))
        hip.synchronize()


def test_server_draws_the_boulders(hip, plane, sphere):
    """The tile server with the batch's boulder layer (tgx_server_boulders): each served tile,
    alone or in a batch of several, equals the batched path's heights bit for bit."""
    from terraingen_amd.pipeline import TileServer

    for b in (plane, sphere[0]):
        want = run(hip, b)
        server = TileServer(b, slots=2, max_batch_tiles=4, max_batch_cells=1 << 18,
                            pool_bytes=1 << 24)
        try:
            tiles = [(float(t["origin_x"]), float(t["origin_y"]), float(t["pixel"]),
                      int(t["width"]), int(t["height"])) for t in b.tiles]
            tickets = [server.submit(*t) for t in tiles]
            for i, ticket in enumerate(tickets):
                h, _ = server.wait(ticket)
                assert np.array_equal(h.view(np.uint32), b.tile_heights(want, i).view(np.uint32)), i
        finally:
            server.close()


def test_strip_pass_with_boulders_equals_two_passes(hip):
    """A grid's craters and boulders in one strip pass (tg_hip_crater_boulder_launch) give the
    heights of the crater pass then the boulder pass (C3's populations on a 513 x 385 grid of
    0.25 m: complex craters with peaks and terraces, and the reference's boulder placement)."""
    from terraingen_amd.pipeline import PlaneTerrain

    grid = N.north_up_grid(513, 385, 0.25, (12.0, -30.0))
    half_w, half_h = 0.5 * 513 * 0.25, 0.5 * 385 * 0.25
    region = (12.0 - half_w - 2.0, -30.0 - half_h - 2.0, 12.0 + half_w + 2.0, -30.0 + half_h + 2.0)
    placed = N.place_boulders(region, 0.25, 3.8e-3, 3.0, 0.2, 3.0, seed=2)
    spec = CraterLayerSpec(sfd_coefficient=0.0156, sfd_exponent=2.2, min_diameter=0.5,
                           max_diameter=300.0, seed=2, complex_=1, transition_diameter=50.0,
                           peaks=1, terraces=1, terrace_onset=150.0, terrace_full=300.0)
    terrain = PlaneTerrain(hip, grid, [C2_FRACTAL], 0.0, spec,
                           boulders=N.relief_boulders(placed))
    terrain.prepare()
    try:
        out = N.DeviceArray(hip, grid.width * grid.height * 4)
        terrain.launch(out, fused=False)
        hip.synchronize()
        two = out.download(np.empty((grid.height, grid.width), np.float32))
        terrain.launch(out, fused=True)
        hip.synchronize()
        one = out.download(np.empty((grid.height, grid.width), np.float32))
        out2 = N.DeviceArray(hip, grid.width * grid.height * 4)
        terrain.launch_noise(out2)
        terrain.launch_craters(out2)
        hip.synchronize()
        bare = out2.download(np.empty((grid.height, grid.width), np.float32))
    finally:
        terrain.close()
    print(f"[strip boulders] {terrain.crater_count} craters, {terrain.boulder_count} boulders, "
          f"{int((two > bare).sum())} cells raised", flush=True)
    assert (two > bare).sum() > 1000
    assert np.array_equal(one.view(np.uint32), two.view(np.uint32))
    # The same grid through the batched kernel as one tile: the same crater list (generated
    # over the grid's box at its spacing) and boulders, the same bits.
    box = (grid.origin_x, grid.origin_y - grid.height * grid.pixel,
           grid.origin_x + grid.width * grid.pixel, grid.origin_y)
    batch = TileBatch(hip, [C2_FRACTAL], spec, region=box, finest_pixel=grid.pixel,
                      boulders=N.relief_boulders(placed))
    batch.build()
    try:
        tile = np.zeros(1, N.TILE_DTYPE)
        tile[0] = (grid.origin_x, grid.origin_y, grid.pixel, grid.width, grid.height)
        batch.stage(tile)
        assert batch.records.size == terrain.crater_count
        got = run(hip, batch).reshape(grid.height, grid.width)
    finally:
        batch.close()
    assert np.array_equal(got.view(np.uint32), two.view(np.uint32))
