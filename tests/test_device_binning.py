# SPDX-License-Identifier: BSD-3-Clause
"""Binning on the device (tgx_index_windows, tg_hip_batch_stage_windows): each tile's crater
windows, listed into bins, give the host's bins and entries (tgx_index_bin) exactly, on a plane
and on a sphere; the device's ballot-compacted lists are those bins bit for bit; and the batch's
heights are the same for every bin side (8, 16, 32 cells) and every culled-walk depth (32, 64,
128 entries), since each choice visits the same craters in the same order."""
import numpy as np
import pytest

from terraingen_amd import native as N
from terraingen_amd.pipeline import (CraterLayerSpec, SphereTileBatch, TileBatch, c4_tiles,
                                     global_craters, sphere_tiles)

C2_FRACTAL = dict(seed=1, wavelength=512.0, amplitude=20.0, octaves=8, lacunarity=2.0, gain=0.5)
BODY = 20000.0


def bins_from_windows(descs, tile_windows, windows, side):
    """The bins the windows give: every window's crater in each bin its cells meet, in window
    order (restated in numpy from the rule, not from the library's code)."""
    bins, entries = [], []
    for t, d in enumerate(descs):
        first, count = tile_windows[2 * t], tile_windows[2 * t + 1]
        w = windows[first:first + count]
        across = int(d["bins_across"])
        down = (int(d["height"]) + side - 1) // side
        for by in range(down):
            for bx in range(across):
                hit = ((w["col0"] // side <= bx) & (w["col1"] // side >= bx) &
                       (w["row0"] // side <= by) & (w["row1"] // side >= by))
                bins.append((t, len(entries), int(hit.sum()), 0))
                entries += w["crater"][hit].tolist()
    return np.array(bins, N.BIN_DTYPE), np.array(entries, np.uint32)


def plane_index():
    rng = np.random.default_rng(3)
    n = 4000
    foot = np.zeros((n, 4))
    foot[:, 0] = rng.uniform(0.0, 300.0, n)
    foot[:, 1] = rng.uniform(0.0, 300.0, n)
    foot[:, 3] = 10.0 ** rng.uniform(0.0, 1.8, n)
    foot[:, 2] = 1.5 * foot[:, 3]
    return N.CatalogueIndex(foot), c4_tiles(30, 300.0, seed=9, min_k=4, max_k=8,
                                             pixel_range=(0.25, 2.0))


def sphere_index():
    craters = global_craters(20000, BODY, max_diameter=2000.0, seed=6)
    foot = np.zeros((craters.size, 4))
    foot[:, 0], foot[:, 1] = craters["x"], craters["y"]
    foot[:, 3] = craters["diameter"]
    foot[:, 2] = 1.6 * craters["diameter"]
    tiles = sphere_tiles(24, BODY, seed=9, min_k=4, max_k=7, pixel_range=(0.25, 50.0))
    unit = BODY * np.pi / 180.0
    tiles[0] = (179.9, 10.0, 2.0 / unit, 65, 65)
    tiles[1] = (-30.0, 90.0, 5.0 / unit, 33, 33)
    return N.CatalogueIndex(foot, radius=BODY), tiles


@pytest.mark.parametrize("make", [plane_index, sphere_index], ids=["plane", "sphere"])
def test_windows_give_the_host_bins(make):
    index, tiles = make()
    descs, bins, entries = index.bin(tiles)
    wd, tw, windows = index.windows(tiles, 16)
    assert np.array_equal(wd, descs)
    got_bins, got_entries = bins_from_windows(wd, tw, windows, 16)
    assert np.array_equal(got_bins, bins) and np.array_equal(got_entries, entries)
    assert entries.size > 1000


@pytest.mark.parametrize("side", [8, 32])
def test_windows_descriptors_for_other_bin_sides(side):
    index, tiles = plane_index()
    descs, tw, windows = index.windows(tiles, side)
    assert (descs["bins_across"] == (tiles["width"] + side - 1) // side).all()
    per = descs["bins_across"].astype(np.int64) * ((tiles["height"] + side - 1) // side)
    assert np.array_equal(descs["bin_first"], np.concatenate([[0], np.cumsum(per)[:-1]]))
    assert (windows["col1"] < np.repeat(tiles["width"], tw[1::2])).all()
    with pytest.raises(N.HipError, match="bin side"):
        index.windows(tiles, 12)


def heights(hip, b):
    out = N.DeviceArray(hip, b.cells * 4)
    b.run(out, use_graph=False)
    b.wait()
    return out.download(np.empty(b.cells, np.float32))


@pytest.mark.gpu
@pytest.mark.parametrize("kind", ["plane", "sphere"])
def test_device_bins_equal_the_host_bins_and_heights(hip, kind):
    if kind == "plane":
        b = TileBatch(hip, [C2_FRACTAL], CraterLayerSpec(seed=4), region=(0.0, 0.0, 300.0, 300.0))
        tiles = c4_tiles(24, 300.0, seed=9, min_k=4, max_k=8, pixel_range=(0.25, 2.0))
    else:
        craters = global_craters(20000, BODY, max_diameter=2000.0, seed=6)
        b = SphereTileBatch(hip, [C2_FRACTAL], CraterLayerSpec(seed=4), region=None,
                            radius=BODY, craters_given=craters)
        tiles = sphere_tiles(24, BODY, seed=9, min_k=4, max_k=7, pixel_range=(0.25, 50.0))
    b.build()
    try:
        b.stage(tiles)
        host_bins, host_entries = b.bins.copy(), b.entries.copy()
        want = heights(hip, b)
        b.stage(tiles, device_bins=True, bin_side=16)
        b.read_bins()
        assert np.array_equal(b.bins, host_bins)
        assert np.array_equal(b.entries, host_entries)
        assert host_entries.size > 500
        got = heights(hip, b)
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
        for side in (8, 32):
            b.stage(tiles, device_bins=True, bin_side=side)
            for depth in (32, 64, 128):
                hip.call("tg_hip_batch_cull_depth", b.batch, depth)
                got = heights(hip, b)
                assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), (side, depth)
            hip.call("tg_hip_batch_cull_depth", b.batch, 64)
        print(f"[device bins] {kind}: {host_entries.size} entries equal to the host's; heights "
              f"equal for bin sides 8/16/32 and depths 32/64/128", flush=True)
    finally:
        b.close()


@pytest.mark.gpu
def test_device_bins_of_a_row_wider_than_one_pass(hip):
    """A tile more than 512 bins across is binned in several passes along each row: its bins and
    entries are still the host's, and its heights the same at bin sides 8 and 32."""
    b = TileBatch(hip, [C2_FRACTAL], CraterLayerSpec(seed=4), region=(0.0, 0.0, 300.0, 300.0))
    tiles = np.zeros(2, N.TILE_DTYPE)
    tiles[0] = (-2200.0, 160.0, 0.25, 8300, 24)  # 519 bins of 16 across, craters past the 512th
    tiles[1] = (40.0, 250.0, 0.25, 65, 65)
    b.build()
    try:
        b.stage(tiles)
        host_bins, host_entries = b.bins.copy(), b.entries.copy()
        want = heights(hip, b)
        b.stage(tiles, device_bins=True, bin_side=16)
        b.read_bins()
        assert np.array_equal(b.bins, host_bins)
        assert np.array_equal(b.entries, host_entries)
        assert host_entries.size > 100
        for side in (8, 32):
            b.stage(tiles, device_bins=True, bin_side=side)
            got = heights(hip, b)
            assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), side
    finally:
        b.close()
