# SPDX-License-Identifier: BSD-3-Clause
"""Batched tiles: one persistent launch over tiles of mixed size and pixel size gives, tile for
tile, the same heights as evaluating each tile alone through the plain path, and the oracle's."""
import ctypes as C

import numpy as np
import pytest

from terraingen_amd import native as N
from terraingen_amd.pipeline import CraterLayerSpec, TileBatch, c4_tiles

pytestmark = pytest.mark.gpu

C2_FRACTAL = dict(seed=1, wavelength=512.0, amplitude=20.0, octaves=8, lacunarity=2.0, gain=0.5)


@pytest.fixture(scope="module")
def batch(hip):
    b = TileBatch(hip, [C2_FRACTAL], CraterLayerSpec(seed=4), region=(0.0, 0.0, 300.0, 300.0))
    b.build()
    tiles = c4_tiles(24, 300.0, seed=9, min_k=4, max_k=8, pixel_range=(0.25, 2.0))
    b.stage(tiles)
    yield b
    b.close()


def run_batch(hip, b, use_graph):
    out = N.DeviceArray(hip, b.cells * 4)
    b.run(out, use_graph)
    b.wait()
    return out.download(np.empty(b.cells, np.float32))


def test_batched_equals_each_tile_alone(hip, batch):
    got = run_batch(hip, batch, use_graph=False)
    worst = 0.0
    for i in range(batch.tiles.size):
        alone = batch.plain_run(batch.plain_inputs(i))
        mine = batch.tile_heights(got, i)
        worst = max(worst, float(np.abs(mine - alone).max()))
        assert np.array_equal(mine.view(np.uint32), alone.view(np.uint32)), i
    print(f"[batch] {batch.tiles.size} tiles, {batch.cells} cells: bit-equal to the plain path",
          flush=True)


def test_graph_replays_match_eager(hip, batch):
    """The captured graph, replayed several times with eager work in between, gives the eager
    result every time."""
    eager = run_batch(hip, batch, use_graph=False)
    for _ in range(3):
        again = run_batch(hip, batch, use_graph=True)
        assert np.array_equal(again.view(np.uint32), eager.view(np.uint32))
        hip.synchronize()


def test_batched_tile_matches_oracle(hip, orc, batch):
    got = run_batch(hip, batch, use_graph=True)
    from crater_cases import oracle_heights_f64  # noqa: F401  (FP64 oracle on the plane)

    for i in (0, 5, 11):
        inp = batch.plain_inputs(i)
        g = inp["grid"]
        want = np.zeros((g.height, g.width), np.float32)
        params = N.KernelParams(g.width, g.height, 0, 0.0, 0)
        orc.or_noise_grid(C.byref(batch.noise_kernel), inp["xs"].ctypes.data,
                          inp["ys"].ctypes.data, g.width, g.height, C.byref(params),
                          want.ctypes.data, 1)
        if inp["craters"] is not None:
            kernel, records, data, tiles, entries = inp["craters"]
            orc.or_shape_grid(C.byref(kernel), records.ctypes.data, data.ctypes.data,
                              tiles.ctypes.data, entries.ctypes.data, inp["xs"].ctypes.data,
                              inp["ys"].ctypes.data, inp["lows"].ctypes.data, g.width, g.height,
                              C.byref(params), want.ctypes.data, 1)
        d = np.abs(batch.tile_heights(got, i) - want)
        assert d.max() <= 1e-4, (i, d.max())


@pytest.mark.parametrize("fractals", [
    [C2_FRACTAL],
    [dict(seed=12, wavelength=40.0, octaves=12, amplitude=3.0),
     dict(seed=3, type=1, offset=0.9, blend=1)],
    [dict(seed=5, wavelength=90.0, octaves=6), dict(seed=6, warp_amplitude=25.0,
                                                    warp_wavelength=300.0, octaves=7)],
])
def test_batched_noise_bit_exact_across_pixel_sizes(hip, orc, fractals):
    """Tiles from 3 cm to 60 m a pixel, so the noise's octaves span from far under one lattice
    cell per 8 x 8 square (corner hashes shared across the wave) to many (hashed per lane), on
    every sign of the octave rotation; each tile's heights equal the oracle's bit for bit."""
    b = TileBatch(hip, fractals, CraterLayerSpec(seed=4), region=(0.0, 0.0, 60.0, 60.0))
    b.build()
    tiles = c4_tiles(20, 3000.0, seed=3, min_k=3, max_k=7, pixel_range=(0.03, 60.0))
    tiles["origin_x"] += 2.0e4  # away from the crater catalogue: noise only
    tiles["origin_y"] -= 7.5e3
    b.stage(tiles)
    try:
        got = run_batch(hip, b, use_graph=False)
        for i in range(tiles.size):
            inp = b.plain_inputs(i)
            assert inp["craters"] is None
            g = inp["grid"]
            want = np.zeros((g.height, g.width), np.float32)
            params = N.KernelParams(g.width, g.height, 0, 0.0, 0)
            orc.or_noise_grid(C.byref(b.noise_kernel), inp["xs"].ctypes.data,
                              inp["ys"].ctypes.data, g.width, g.height, C.byref(params),
                              want.ctypes.data, 1)
            mine = b.tile_heights(got, i)
            assert np.array_equal(mine.view(np.uint32), want.view(np.uint32)), (
                i, float(tiles["pixel"][i]), int(g.width))
    finally:
        b.close()


def test_async_staging_matches_staging(hip, batch):
    """tg_hip_batch_stage_async queues the same copies on the batch's stream: the next run gives
    the heights of a synchronous staging."""
    want = run_batch(hip, batch, use_graph=False)
    descs, bins, entries = (np.ascontiguousarray(a) for a in (batch.descs, batch.bins,
                                                              batch.entries))
    hip.call("tg_hip_batch_stage_async", batch.batch, descs.ctypes.data, descs.size,
             bins.ctypes.data, bins.size, entries.ctypes.data, entries.size)
    got = run_batch(hip, batch, use_graph=False)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))


def test_stage_rejects_malformed_batches(hip, batch):
    bins = batch.bins.copy()
    bins["count"][0] = 10 ** 8  # runs past the entries
    with pytest.raises(N.HipError):
        hip.call("tg_hip_batch_stage", batch.batch, batch.descs.ctypes.data, batch.descs.size,
                 bins.ctypes.data, bins.size, batch.entries.ctypes.data, batch.entries.size)
    descs = batch.descs.copy()
    descs["bins_across"][1] += 1
    with pytest.raises(N.HipError):
        hip.call("tg_hip_batch_stage", batch.batch, descs.ctypes.data, descs.size,
                 batch.bins.ctypes.data, batch.bins.size, batch.entries.ctypes.data,
                 batch.entries.size)
    # Restage the good batch so later tests see it.
    hip.call("tg_hip_batch_stage", batch.batch, batch.descs.ctypes.data, batch.descs.size,
             batch.bins.ctypes.data, batch.bins.size, batch.entries.ctypes.data,
             batch.entries.size)


def test_crater_walks_agree(hip, batch):
    """Walking the bin lists together, per lane or culled 64 entries at a time applies the same
    craters in the same order."""
    batch.tune(N.TG_WALK_TOGETHER, 8)
    together = run_batch(hip, batch, use_graph=False)
    batch.tune(N.TG_WALK_PER_LANE, 4)
    per_lane = run_batch(hip, batch, use_graph=True)
    assert np.array_equal(together.view(np.uint32), per_lane.view(np.uint32))
    batch.tune(N.TG_WALK_CULLED, 8)
    culled = run_batch(hip, batch, use_graph=False)
    assert np.array_equal(together.view(np.uint32), culled.view(np.uint32))
    batch.tune()


def test_hoisted_mpd_path_is_bit_identical(hip, batch):
    """The per-crater constants of the MPD profile path give the bits of the full evaluation."""
    batch.tune(hoisted=False)
    full = run_batch(hip, batch, use_graph=False)
    batch.tune(hoisted=True)
    hoisted = run_batch(hip, batch, use_graph=False)
    assert np.array_equal(full.view(np.uint32), hoisted.view(np.uint32))


@pytest.mark.parametrize("chunk,groups", [(1, 1), (3, 8), (256, 2)])
def test_queue_chunks_cover_every_bin_once(hip, batch, chunk, groups):
    """However the bins are taken (one per atomic or many, few workgroups or many), every cell of
    every tile is written once with the same heights."""
    batch.tune(N.TG_WALK_TOGETHER, 8, True, 8)
    want = run_batch(hip, batch, use_graph=False)
    batch.tune(N.TG_WALK_TOGETHER, groups, True, chunk)
    out = N.DeviceArray(hip, batch.cells * 4)
    poison = np.full(batch.cells, np.nan, np.float32)
    out.upload(poison)
    batch.run(out, use_graph=False)
    batch.wait()
    got = out.download(np.empty(batch.cells, np.float32))
    batch.tune()
    assert np.isfinite(got).all()
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
