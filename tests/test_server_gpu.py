# SPDX-License-Identifier: BSD-3-Clause
"""The tile server: each tile gives the same heights alone and in a mix of concurrent requests
(bit for bit), the same as the batched path, and admission stays inside the HBM pool."""
import threading

import numpy as np
import pytest

from terraingen_amd.pipeline import CraterLayerSpec, TileBatch, TileServer, c4_tiles

pytestmark = pytest.mark.gpu

C2_FRACTAL = dict(seed=1, wavelength=512.0, amplitude=20.0, octaves=8, lacunarity=2.0, gain=0.5)


@pytest.fixture(scope="module")
def batch(hip):
    b = TileBatch(hip, [C2_FRACTAL], CraterLayerSpec(seed=4), region=(0.0, 0.0, 400.0, 400.0))
    b.build()
    yield b
    b.close()


def tile_list(n, seed):
    t = c4_tiles(n, 400.0, seed=seed, min_k=4, max_k=8, pixel_range=(0.25, 2.0))
    return [(float(x["origin_x"]), float(x["origin_y"]), float(x["pixel"]), int(x["width"]),
             int(x["height"])) for x in t]


def test_alone_equals_mix_equals_batched(hip, batch):
    tiles = tile_list(40, 11)
    server = TileServer(batch, slots=3, max_batch_tiles=8, max_batch_cells=1 << 20,
                        pool_bytes=1 << 26)
    alone = []
    for t in tiles:
        h, info = server.wait(server.submit(*t))
        assert info["batch_tiles"] == 1
        alone.append(h)
    # The mix: three threads submit at once, so batches hold several tiles and overlap.
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
        assert info["submitted"] <= info["admitted"] <= info["launched"] <= info["done"]
        assert np.array_equal(h.view(np.uint32), alone[i].view(np.uint32)), i
    stats = server.stats()
    server.close()
    print(f"[server] {len(tiles)} tiles; batch sizes in the mix {sorted(set(sizes))}; {stats}",
          flush=True)
    assert max(sizes) > 1 and stats["failed"] == 0
    # The batched path over the same tiles.
    arr = c4_tiles(len(tiles), 400.0, seed=11, min_k=4, max_k=8, pixel_range=(0.25, 2.0))
    batch.stage(arr)
    from terraingen_amd import native as N

    out = N.DeviceArray(hip, batch.cells * 4)
    batch.run(out, use_graph=False)
    batch.wait()
    got = out.download(np.empty(batch.cells, np.float32))
    for i in range(len(tiles)):
        assert np.array_equal(batch.tile_heights(got, i).view(np.uint32),
                              alone[i].view(np.uint32)), i


def test_admission_stays_inside_the_pool(hip, batch):
    tiles = tile_list(24, 12)
    biggest = max(w * h for *_, w, h in tiles)
    pool = ((biggest * 4 + 255) // 256) * 256 * 2  # room for about two tiles at a time
    server = TileServer(batch, slots=4, max_batch_tiles=16, max_batch_cells=biggest,
                        pool_bytes=pool)
    tickets = [server.submit(*t) for t in tiles]
    for ticket in tickets:
        server.wait(ticket)
    stats = server.stats()
    server.close()
    assert stats["tiles"] == len(tiles) and stats["failed"] == 0
    assert 0 < stats["pool_peak_bytes"] <= pool


def test_submit_rejects_oversized_and_malformed_tiles(hip, batch):
    server = TileServer(batch, slots=1, max_batch_tiles=4, max_batch_cells=4096,
                        pool_bytes=1 << 20)
    from terraingen_amd import native as N

    with pytest.raises(N.HipError, match="more cells"):
        server.submit(0.0, 0.0, 1.0, 65, 65)
    with pytest.raises(N.HipError, match="malformed"):
        server.submit(0.0, 0.0, 0.0, 16, 16)
    h, _ = server.wait(server.submit(10.0, 300.0, 0.5, 33, 17))
    assert h.shape == (17, 33) and np.isfinite(h).all()
    server.close()


def test_sphere_server_alone_equals_mix_equals_batched(hip):
    """A server over a sphere's index runs longitude-latitude tiles on that sphere: each tile's
    heights are the same alone, in a mix, and through the batched path."""
    from terraingen_amd import native as N
    from terraingen_amd.pipeline import SphereTileBatch, global_craters, sphere_tiles

    body = 20000.0
    b = SphereTileBatch(hip, [C2_FRACTAL], CraterLayerSpec(seed=4), region=None, radius=body,
                        craters_given=global_craters(8000, body, max_diameter=2000.0, seed=3))
    b.build()
    arr = sphere_tiles(24, body, seed=5, min_k=4, max_k=7, pixel_range=(0.25, 8.0))
    tiles = [(float(x["origin_x"]), float(x["origin_y"]), float(x["pixel"]), int(x["width"]),
              int(x["height"])) for x in arr]
    server = TileServer(b, slots=3, max_batch_tiles=8, max_batch_cells=1 << 20,
                        pool_bytes=1 << 26)
    try:
        alone = [server.wait(server.submit(*t))[0] for t in tiles]
        tickets = [server.submit(*t) for t in tiles]
        sizes = []
        for i, ticket in enumerate(tickets):
            h, info = server.wait(ticket)
            sizes.append(info["batch_tiles"])
            assert np.array_equal(h.view(np.uint32), alone[i].view(np.uint32)), i
        assert server.stats()["failed"] == 0 and max(sizes) > 1
    finally:
        server.close()
    b.stage(arr)
    out = N.DeviceArray(hip, b.cells * 4)
    b.run(out, use_graph=False)
    b.wait()
    got = out.download(np.empty(b.cells, np.float32))
    shaped = 0
    for i in range(len(tiles)):
        assert np.array_equal(b.tile_heights(got, i).view(np.uint32),
                              alone[i].view(np.uint32)), i
        shaped += b.plain_inputs(i)["craters"] is not None
    b.close()
    assert shaped > 0
