# SPDX-License-Identifier: BSD-3-Clause
"""tg_hip_batch_run_static, the batched launch that allocates nothing: its tiles, bins, entries and
queue counters live in the library's static device scratch. Its heights are those of
tg_hip_batch_stage + tg_hip_batch_run, bit for bit, on two calls in a row, the second with a
smaller batch (what the first left in the scratch must not leak into the second), on a plane and
on a sphere; batches beyond the scratch are refused before anything runs."""
import ctypes as C

import numpy as np
import pytest

from terraingen_amd import native as N
from terraingen_amd.pipeline import (CraterLayerSpec, SphereTileBatch, TileBatch, c4_tiles,
                                     global_craters, sphere_tiles)

pytestmark = pytest.mark.gpu

C2_FRACTAL = dict(seed=1, wavelength=512.0, amplitude=20.0, octaves=8, lacunarity=2.0, gain=0.5)
BODY = 20000.0


def staged_run(hip, b, tiles):
    """The existing entry points: stage, run, wait; the heights on the host."""
    b.stage(tiles)
    out = N.DeviceArray(hip, b.cells * 4)
    b.run(out, use_graph=False)
    b.wait()
    return out.download(np.empty(b.cells, np.float32))


def static_run(hip, b, radius):
    """The static entry point on the batch `b` last staged, on the null stream."""
    out = N.DeviceArray(hip, b.cells * 4)
    # A value the kernel overwrites everywhere it writes, so a missed tile shows.
    hip.call("tg_hip_fill", hip.handle, out.ptr, C.c_float(-12345.0), b.cells, None)
    hip.call("tg_hip_batch_run_static", hip.handle, b.descs.ctypes.data, b.descs.size,
             b.bins.ctypes.data, b.bins.size, b.entries.ctypes.data, b.entries.size, b.plan,
             C.c_float(b.elevation), b.catalogue, radius, out.ptr, b.cells, None)
    hip.synchronize()
    return out.download(np.empty(b.cells, np.float32))


def check_two_calls(hip, b, tiles, radius, smaller):
    for n, some in ((tiles.size, tiles), (smaller, tiles[:smaller])):
        print(f"[static] radius {radius}, {n} tiles", flush=True)
        want = staged_run(hip, b, some)
        got = static_run(hip, b, radius)
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), n
        assert b.entries.size > 0


def test_static_run_equals_staged_run_on_a_plane(hip):
    b = TileBatch(hip, [C2_FRACTAL], CraterLayerSpec(seed=4), region=(0.0, 0.0, 300.0, 300.0))
    b.build()
    try:
        tiles = c4_tiles(24, 300.0, seed=9, min_k=4, max_k=8, pixel_range=(0.25, 2.0))
        check_two_calls(hip, b, tiles, 0.0, smaller=7)
    finally:
        b.close()


def test_static_run_equals_staged_run_on_a_sphere(hip):
    craters = global_craters(20000, BODY, max_diameter=2000.0, seed=6)
    b = SphereTileBatch(hip, [C2_FRACTAL], CraterLayerSpec(seed=4), region=None, radius=BODY,
                        craters_given=craters)
    b.build()
    try:
        tiles = sphere_tiles(16, BODY, seed=9, min_k=4, max_k=7, pixel_range=(0.25, 50.0))
        check_two_calls(hip, b, tiles, BODY, smaller=5)
    finally:
        b.close()


def test_static_run_refuses_batches_beyond_its_scratch(hip):
    """Checked on the host before anything is copied or launched."""
    tiles = np.zeros(4097, N.TILE_DESC_DTYPE)
    bins = np.zeros(1, N.BIN_DTYPE)
    entries = np.zeros(1, np.uint32)
    out = N.DeviceArray(hip, 4)
    with pytest.raises(N.HipError, match="static scratch"):
        hip.call("tg_hip_batch_run_static", hip.handle, tiles.ctypes.data, tiles.size,
                 bins.ctypes.data, bins.size, entries.ctypes.data, entries.size, None,
                 C.c_float(0.0), None, 0.0, out.ptr, 1, None)
