# SPDX-License-Identifier: BSD-3-Clause
"""A plane terrain evaluated on the hip backend: a source (flat, or synthetic fractal noise),
optional noise layers and a crater layer, composed from the host library's placement,
preparation and binning and the kernel library's resident launches.

It follows the reference's evaluation order (src/pipeline/terrain.cpp:289-336): the source fills
the grid, then each layer applies in turn. The crater layer's craters are those generated over the
grid (CraterLayer::ApplySamples, src/layers/craters/crater_layer.cpp:337-361), prepared and binned
on the host, shaped on the device.
"""
from __future__ import annotations

import ctypes as C
import time
from dataclasses import dataclass, field

import numpy as np

from . import native as N


@dataclass
class CraterLayerSpec:
    """A crater layer with a power-law density N(>=D) = sfd_coefficient D^-sfd_exponent per m^2, or
    the reference's default models."""
    sfd_coefficient: float = 0.05
    sfd_exponent: float = 2.0
    min_diameter: float = 1.0
    max_diameter: float = 200.0
    lifetime_gyr: float = 1.0
    profile: int = 1            # TG_PROFILE_MPD; 0 = Fassett-Thomson
    complex_: int = 0
    transition_diameter: float = 15000.0
    irregular: int = 1
    peaks: int = 0
    terraces: int = 0
    terrace_onset: float = 0.0   # m; 0 with terrace_full 0 keeps the defaults (20 km, 65 km)
    terrace_full: float = 0.0
    seed: int = 1
    rim_sigma: float = 0.0
    noise_amplitude: float = 0.0
    noise_rim_amplitude: float = 0.0
    wall_streaks: float = 0.0
    ejecta: float = 0.0
    # 1: the reference's DefaultCraterModels (lunar equilibrium density, steady-state ages,
    # diffused MPD profile, matched diffusion); the SFD, profile and lifetime fields are then
    # unused, and the diameters are the layer's defaults unless set.
    model: int = 0

    def models(self) -> N.CraterModels:
        return N.CraterModels(self.profile, self.complex_, self.transition_diameter,
                              self.lifetime_gyr, self.model)

    def apply(self) -> N.Apply:
        return N.Apply(self.irregular, self.peaks, self.terraces, self.rim_sigma,
                       self.noise_amplitude, self.noise_rim_amplitude, self.wall_streaks,
                       self.ejecta, self.terrace_onset, self.terrace_full)


@dataclass
class PlaneTerrain:
    """Grid, source and layers; `prepare` does the host work, `launch` the device work."""
    dev: N.HipDevice
    grid: N.Grid
    source_fractals: list = field(default_factory=list)   # SyntheticSource fractals (origin 0)
    elevation: float = 0.0
    craters: CraterLayerSpec | None = None
    tile_side: int = 16
    timings: dict = field(default_factory=dict)
    # A boulder layer after the craters: a list in BOULDER_DTYPE (the reference's own, or
    # scatter_boulders), drawn as dome relief at the grid's spacing.
    boulders: np.ndarray | None = None

    def prepare(self) -> None:
        g = self.grid
        cx64, cy64 = N.grid_centers(g)
        self.xs, xlo = N.split(cx64)
        self.ys, ylo = N.split(cy64)
        self.lows = np.concatenate([xlo, ylo])
        self.d_xs = N.DeviceArray.from_host(self.dev, self.xs)
        self.d_ys = N.DeviceArray.from_host(self.dev, self.ys)
        self.params = N.KernelParams(g.width, g.height, N.NODATA_NONE, 0.0, 0)
        err = C.create_string_buffer(512)
        self.plan = None
        if self.source_fractals:
            kernel = N.noise_kernel(self.source_fractals)
            plan = C.c_void_p()
            code = self.dev.lib.tg_hip_noise_plan_create(self.dev.handle, C.byref(kernel),
                                                         C.byref(plan), err, 512)
            if code:
                raise N.HipError(err.value.decode())
            self.plan = plan
        self.crater_set = None
        self.crater_count = 0
        if self.craters is not None:
            self._prepare_craters()
        self.boulder_set = None
        self.boulder_count = 0
        if self.boulders is not None:
            self._prepare_boulders()

    def _prepare_boulders(self) -> None:
        g = self.grid
        t0 = time.perf_counter()
        kernel, records, foot = N.prepare_boulders(self.boulders, g.pixel)
        tiles, entries = N.bin_plane(g, foot[:, 0], foot[:, 1], foot[:, 2], self.tile_side)
        t1 = time.perf_counter()
        out = C.c_void_p()
        self.dev.call("tg_hip_boulder_set_create", self.dev.handle, C.byref(kernel),
                      records.ctypes.data, records.size, tiles.ctypes.data, tiles.size,
                      entries.ctypes.data, entries.size, self.lows.ctypes.data, self.lows.size,
                      C.byref(out))
        self.boulder_set = out
        self.boulder_kernel, self.boulder_records, self.boulder_foot = kernel, records, foot
        self.boulder_tiles, self.boulder_entries = tiles, entries
        self.boulder_count = records.size
        self.timings.update(boulder_prepare_ms=(t1 - t0) * 1e3,
                            boulder_upload_ms=(time.perf_counter() - t1) * 1e3)

    def _prepare_craters(self) -> None:
        spec = self.craters
        g = self.grid
        t0 = time.perf_counter()
        x1 = g.origin_x + g.width * g.pixel
        y0 = g.origin_y - g.height * g.pixel
        gen = N.Generation(min_x=g.origin_x, min_y=y0, max_x=x1, max_y=g.origin_y,
                           sfd_coefficient=spec.sfd_coefficient, sfd_exponent=spec.sfd_exponent,
                           min_diameter=spec.min_diameter, max_diameter=spec.max_diameter,
                           margin_radii=N.margin_radii(spec.models(), spec.apply()),
                           cell_size_factor=16.0, min_resolved_diameter=2.0 * g.pixel,
                           lifetime_gyr=spec.lifetime_gyr, seed=spec.seed, model=spec.model)
        self.crater_list = N.generate_craters(gen)
        t1 = time.perf_counter()
        kernel, records, data, foot = N.prepare_craters(self.crater_list, spec.models(),
                                                        spec.apply(), spec.seed, g.pixel)
        t2 = time.perf_counter()
        tiles, entries = N.bin_plane(g, foot[:, 0], foot[:, 1], foot[:, 2], self.tile_side)
        t3 = time.perf_counter()
        self.kernel, self.records, self.data, self.foot = kernel, records, data, foot
        self.tiles, self.entries = tiles, entries
        frames = np.zeros(1, N.FRAME_DTYPE)
        out = C.c_void_p()
        err = C.create_string_buffer(512)
        code = self.dev.lib.tg_hip_crater_set_create(
            self.dev.handle, C.byref(kernel), records.ctypes.data, records.size,
            data.ctypes.data, data.size, tiles.ctypes.data, tiles.size, entries.ctypes.data,
            entries.size, frames.ctypes.data, 1, self.lows.ctypes.data, self.lows.size,
            C.byref(out), err, 512)
        if code:
            raise N.HipError(err.value.decode())
        t4 = time.perf_counter()
        self.crater_set = out
        self.crater_count = records.size
        self.timings.update(generate_ms=(t1 - t0) * 1e3, prepare_ms=(t2 - t1) * 1e3,
                            bin_ms=(t3 - t2) * 1e3, upload_ms=(t4 - t3) * 1e3)

    def launch_noise(self, heights: N.DeviceArray, stream=None) -> None:
        err = C.create_string_buffer(512)
        if self.plan is not None:
            code = self.dev.lib.tg_hip_noise_fill(self.dev.handle, self.plan, self.elevation,
                                                  heights.ptr, self.d_xs.ptr, self.d_ys.ptr,
                                                  C.byref(self.params), stream, err, 512)
            if code:
                raise N.HipError(err.value.decode())
        else:  # a flat source at `elevation`
            self.dev.call("tg_hip_fill", self.dev.handle, heights.ptr, self.elevation,
                          self.grid.width * self.grid.height, stream)

    def launch_craters(self, heights: N.DeviceArray, stream=None) -> None:
        if self.crater_set is None:
            return
        err = C.create_string_buffer(512)
        code = self.dev.lib.tg_hip_crater_launch(self.dev.handle, self.crater_set, heights.ptr,
                                                 self.d_xs.ptr, self.d_ys.ptr,
                                                 C.byref(self.params), stream, err, 512)
        if code:
            raise N.HipError(err.value.decode())

    def launch_boulders(self, heights: N.DeviceArray, stream=None) -> None:
        if self.boulder_set is None:
            return
        self.dev.call("tg_hip_boulder_launch", self.dev.handle, self.boulder_set, heights.ptr,
                      self.d_xs.ptr, self.d_ys.ptr, C.byref(self.params), stream)

    def launch_craters_and_boulders(self, heights: N.DeviceArray, stream=None) -> None:
        """Both layers in one pass over the grid (tg_hip_crater_boulder_launch)."""
        self.dev.call("tg_hip_crater_boulder_launch", self.dev.handle, self.crater_set,
                      self.boulder_set, heights.ptr, self.d_xs.ptr, self.d_ys.ptr,
                      C.byref(self.params), stream)

    def launch(self, heights: N.DeviceArray, stream=None, fused: bool = True) -> None:
        """The source fill, then the craters and the boulders: in one pass when the terrain has
        both (and `fused`), else one launch per layer."""
        self.launch_noise(heights, stream)
        if fused and self.crater_set is not None and self.boulder_set is not None:
            self.launch_craters_and_boulders(heights, stream)
            return
        self.launch_craters(heights, stream)
        self.launch_boulders(heights, stream)

    def evaluate(self) -> np.ndarray:
        g = self.grid
        heights = N.DeviceArray(self.dev, g.width * g.height * 4)
        self.launch(heights)
        self.dev.synchronize()
        return heights.download(np.empty((g.height, g.width), np.float32))

    def close(self) -> None:
        if getattr(self, "plan", None) is not None:
            self.dev.lib.tg_hip_noise_plan_free(self.plan)
            self.plan = None
        if getattr(self, "crater_set", None) is not None:
            self.dev.lib.tg_hip_crater_set_free(self.crater_set)
            self.crater_set = None
        if getattr(self, "boulder_set", None) is not None:
            self.dev.lib.tg_hip_boulder_set_free(self.boulder_set)
            self.boulder_set = None


def apply_order(craters: np.ndarray) -> np.ndarray:
    """CraterApplyOrder, oldest first (src/layers/craters/shaping/apply_options.cpp:96-128):
    decreasing age, then decreasing diameter, then increasing ID, then list index."""
    idx = np.arange(craters.size)
    return np.lexsort((idx, craters["id"], -craters["diameter"], -craters["age_gyr"]))


@dataclass
class TileBatch:
    """Many plane tiles evaluated in one persistent launch against a resident crater catalogue
    (tg_hip_catalogue_*, tg_hip_batch_*), and, for comparison, one at a time through the plain
    synchronous path (tg_hip_run per layer, host memory in and out)."""
    dev: N.HipDevice
    source_fractals: list
    craters: CraterLayerSpec
    region: tuple  # (min_x, min_y, max_x, max_y) of the catalogue, meters
    finest_pixel: float = 0.25
    elevation: float = 0.0
    timings: dict = field(default_factory=dict)
    # A boulder layer after the craters: a list in BOULDER_DTYPE (x, y in meters on a plane,
    # longitude and latitude on a sphere), drawn in each tile at the tile's spacing.
    boulders: np.ndarray | None = None

    def build(self) -> None:
        spec = self.craters
        x0, y0, x1, y1 = self.region
        t0 = time.perf_counter()
        gen = N.Generation(min_x=x0, min_y=y0, max_x=x1, max_y=y1,
                           sfd_coefficient=spec.sfd_coefficient, sfd_exponent=spec.sfd_exponent,
                           min_diameter=spec.min_diameter, max_diameter=spec.max_diameter,
                           margin_radii=N.margin_radii(spec.models(), spec.apply()),
                           cell_size_factor=16.0, min_resolved_diameter=2.0 * self.finest_pixel,
                           lifetime_gyr=spec.lifetime_gyr, seed=spec.seed)
        self.crater_list = N.generate_craters(gen)
        t1 = time.perf_counter()
        self.kernel, self.records, self.data, self.foot = N.prepare_craters(
            self.crater_list, spec.models(), spec.apply(), spec.seed, self.finest_pixel)
        # The catalogue rows in apply order, for per-tile preparation on the plain path; no crater
        # is dropped at the finest pixel size, so record i is row i.
        if self.records.size != self.crater_list.size:
            raise ValueError("the catalogue's finest pixel size drops craters: lower min_diameter "
                             "or finest_pixel")
        self.ordered = self.crater_list[apply_order(self.crater_list)]
        self.index = N.CatalogueIndex(self.foot)
        t2 = time.perf_counter()
        diam = np.ascontiguousarray(self.foot[:, 3])
        cat = C.c_void_p()
        self.dev.call("tg_hip_catalogue_create", self.dev.handle, C.byref(self.kernel),
                      self.records.ctypes.data, self.records.size, self.data.ctypes.data,
                      self.data.size, diam.ctypes.data, C.byref(cat))
        self.catalogue = cat
        t3 = time.perf_counter()
        self.plan = None
        if self.source_fractals:
            kernel = N.noise_kernel(self.source_fractals)
            plan = C.c_void_p()
            self.dev.call("tg_hip_noise_plan_create", self.dev.handle, C.byref(kernel),
                          C.byref(plan))
            self.plan = plan
            self.noise_kernel = kernel
        self.batch = None
        self.timings.update(generate_ms=(t1 - t0) * 1e3, prepare_ms=(t2 - t1) * 1e3,
                            catalogue_upload_ms=(t3 - t2) * 1e3)
        self._build_boulders()

    def _build_boulders(self) -> None:
        """The boulder catalogue on the device (tg_hip_boulder_catalogue_create): every boulder
        resolved at the finest spacing, prepared there for its geometry, with its diameter and,
        on a sphere, its frame, and an index over relief boxes of 0.75 D (half of D plus the
        widest band a tile that resolves the boulder draws it with, D / 2)."""
        self.boulder_catalogue = None
        if self.boulders is None:
            return
        t0 = time.perf_counter()
        radius = getattr(self, "radius", 0.0)
        _, records, foot = N.prepare_boulders(self.boulders, self.finest_pixel)
        self.boulder_records, self.boulder_foot = records, foot
        diam = np.ascontiguousarray(foot[:, 3])
        frames = N.sphere_frames(foot[:, 0], foot[:, 1], radius) if radius > 0.0 else None
        box = np.column_stack([foot[:, 0], foot[:, 1], 0.75 * foot[:, 3], foot[:, 3]])
        self.boulder_index = N.CatalogueIndex(box, radius=radius)
        out = C.c_void_p()
        self.dev.call("tg_hip_boulder_catalogue_create", self.dev.handle, records.ctypes.data,
                      records.size, diam.ctypes.data,
                      frames.ctypes.data if frames is not None else None, C.byref(out))
        self.boulder_catalogue = out
        self.timings.update(boulder_prepare_ms=(time.perf_counter() - t0) * 1e3)

    def _stage_boulders(self, threads: int) -> None:
        """Bins the boulders against the staged tiles on the host (tgx_index_bin) and stages
        their lists (tg_hip_batch_stage_boulders)."""
        descs, bins, entries = self.boulder_index.bin(self.tiles, threads)
        if not np.array_equal(descs["bin_first"], self.descs["bin_first"]):
            raise RuntimeError("the boulder bins do not follow the crater bins")
        self.boulder_bins, self.boulder_entries = bins, entries
        self.dev.call("tg_hip_batch_stage_boulders", self.batch, self.boulder_catalogue,
                      bins.ctypes.data, bins.size, entries.ctypes.data, entries.size, 0)

    def stage(self, tiles: np.ndarray, threads: int = 0, device_bins: bool = False,
              bin_side: int = 16, with_boulders: bool = True) -> None:
        """Bins `tiles` against the catalogue and stages them on the batch: on the host into
        16-cell bins (tgx_index_bin, tg_hip_batch_stage), or with `device_bins` the host finds
        each tile's crater windows (tgx_index_windows) and the device bins them into bins of
        `bin_side` cells (tg_hip_batch_stage_windows). The boulders, when the batch has them
        (and `with_boulders`), are binned on the host and staged after the craters."""
        t0 = time.perf_counter()
        self.tiles = np.ascontiguousarray(tiles, N.TILE_DTYPE)
        if device_bins:
            self.descs, tile_windows, windows = self.index.windows(self.tiles, bin_side, threads)
            s = bin_side
            w = self.tiles["width"].astype(np.int64)
            h = self.tiles["height"].astype(np.int64)
            bins = int(((w + s - 1) // s * ((h + s - 1) // s)).sum())
            across = windows["col1"].astype(np.int64) // s - windows["col0"] // s + 1
            down = windows["row1"].astype(np.int64) // s - windows["row0"] // s + 1
            entries = int((across * down).sum())
            self.bins = self.entries = None
        else:
            self.descs, self.bins, self.entries = self.index.bin(self.tiles, threads)
            bins, entries, bin_side = self.bins.size, self.entries.size, 16
        t1 = time.perf_counter()
        need = (self.tiles.size, bins, entries)
        if self.batch is None or any(c < n for c, n in zip(self._capacity, need)):
            if self.batch is not None:
                self.dev.lib.tg_hip_batch_free(self.batch)
            cap = (int(self.tiles.size * 1.25) + 1, int(bins * 1.25) + 1,
                   int(entries * 1.25) + 1)
            b = C.c_void_p()
            self.dev.call("tg_hip_batch_create", self.dev.handle, *cap, C.byref(b))
            self.batch, self._capacity = b, cap
            if getattr(self, "radius", 0.0) > 0.0:
                self.dev.call("tg_hip_batch_surface", self.batch, self.radius)
        if device_bins:
            staged = N.SZ()
            self.dev.call("tg_hip_batch_stage_windows", self.batch, self.descs.ctypes.data,
                          self.descs.size, bin_side, tile_windows.ctypes.data,
                          windows.ctypes.data, windows.size, C.byref(staged))
            if staged.value != entries:
                raise RuntimeError(f"the device listed {staged.value} entries, the windows give "
                                   f"{entries}")
        else:
            self.dev.call("tg_hip_batch_stage", self.batch, self.descs.ctypes.data,
                          self.descs.size, self.bins.ctypes.data, self.bins.size,
                          self.entries.ctypes.data, self.entries.size)
        self.bin_side = bin_side
        if self.boulder_catalogue is not None and with_boulders:
            if bin_side != 16:
                raise ValueError("boulders are binned into 16-cell bins")
            self._stage_boulders(threads)
        t2 = time.perf_counter()
        self.bin_count = bins
        self.cells = int((self.tiles["width"].astype(np.int64) *
                          self.tiles["height"].astype(np.int64)).sum())
        self.timings.update(bin_ms=(t1 - t0) * 1e3, stage_ms=(t2 - t1) * 1e3)

    def read_bins(self) -> None:
        """The staged bins and entries in host memory (tg_hip_batch_read_bins), for a batch the
        device binned."""
        nb, ne = N.SZ(), N.SZ()
        self.dev.call("tg_hip_batch_read_bins", self.batch, None, 0, None, 0, C.byref(nb),
                      C.byref(ne))
        bins = np.zeros(nb.value, N.BIN_DTYPE)
        entries = np.zeros(max(ne.value, 1), np.uint32)
        self.dev.call("tg_hip_batch_read_bins", self.batch, bins.ctypes.data, bins.size,
                      entries.ctypes.data, entries.size, C.byref(nb), C.byref(ne))
        self.bins, self.entries = bins, entries[:ne.value]

    @property
    def stream(self):
        return self.dev.lib.tg_hip_batch_stream(self.batch)

    def tune(self, walk: int = N.TG_WALK_CULLED, groups_per_cu: int = 8,
             hoisted: bool = True, chunk: int = 8) -> None:
        """The staged batch's launch choices (tg_hip_batch_tune)."""
        self.dev.call("tg_hip_batch_tune", self.batch, walk, groups_per_cu, int(hoisted), chunk)

    def autotune(self, out: N.DeviceArray,
                 walks=(N.TG_WALK_TOGETHER, N.TG_WALK_PER_LANE, N.TG_WALK_CULLED),
                 groups=(4, 8), chunks=(2, 8, 32), depths=(64,), bin_sides=None,
                 repeats: int = 2) -> dict:
        """Times the launch choices on the staged tiles with HIP events on the batch's stream,
        keeps the fastest and returns {(bin side, walk, groups per CU, chunk, depth): ms}: the
        crater walk, workgroups per CU and bins per queue take first (at the staged bin side and
        a depth of 64), then, at the best of those, each culled-walk depth and, with
        `bin_sides`, each bin side (restaging the tiles with device binning). Every choice gives
        the same heights, so it changes only the time."""
        timings = {}

        def time_one(key):
            self.run(out, use_graph=False)  # warm: code objects, caches
            a, b = N.Event(self.dev), N.Event(self.dev)
            a.record(self.stream)
            for _ in range(repeats):
                self.run(out, use_graph=False)
            b.record(self.stream)
            self.wait()
            timings[key] = a.elapsed_ms(b) / repeats

        side = getattr(self, "bin_side", 16)
        for walk in walks:
            for g in groups:
                for chunk in chunks:
                    self.tune(walk, g, True, chunk)
                    time_one((side, walk, g, chunk, 64))
        best = min(timings, key=timings.get)
        _, walk, g, chunk, _ = best
        for s in (bin_sides or (side,)):
            if s != getattr(self, "bin_side", 16):
                self.stage(self.tiles, device_bins=True, bin_side=s)
            self.tune(walk, g, True, chunk)
            for depth in depths:
                if (s, walk, g, chunk, depth) in timings:
                    continue
                self.dev.call("tg_hip_batch_cull_depth", self.batch, depth)
                time_one((s, walk, g, chunk, depth))
        best = min(timings, key=timings.get)
        if best[0] != getattr(self, "bin_side", 16):
            self.stage(self.tiles, device_bins=True, bin_side=best[0])
        self.tune(best[1], best[2], True, best[3])
        self.dev.call("tg_hip_batch_cull_depth", self.batch, best[4])
        self.tuned = best
        return timings

    def run(self, out: N.DeviceArray, use_graph: bool = True) -> None:
        self.dev.call("tg_hip_batch_run", self.batch, self.plan, self.elevation, self.catalogue,
                      out.ptr, out.nbytes // 4, 1 if use_graph else 0)

    def wait(self) -> None:
        self.dev.call("tg_hip_batch_wait", self.batch)

    def tile_heights(self, out_host: np.ndarray, i: int) -> np.ndarray:
        d = self.descs[i]
        n = int(d["width"]) * int(d["height"])
        o = int(d["out_offset"])
        return out_host[o:o + n].reshape(int(d["height"]), int(d["width"]))

    # ---- the plain per-tile path ---------------------------------------------------------
    def tile_entries(self, i: int) -> np.ndarray:
        """Catalogue rows (apply order) that tile i's bins list, each once, in order."""
        if self.bins is None:
            self.read_bins()
        d = self.descs[i]
        side = getattr(self, "bin_side", 16)
        nb = int(d["bins_across"]) * ((int(d["height"]) + side - 1) // side)
        first = self.bins[int(d["bin_first"])]
        last = self.bins[int(d["bin_first"]) + nb - 1]
        # A tile's bins list their entries one after another.
        return np.unique(self.entries[int(first["first"]):int(last["first"]) +
                                      int(last["count"])])

    def largest_in_reach(self) -> np.ndarray:
        """Per tile, the largest diameter among the craters its bins list (0 when none)."""
        out = np.zeros(self.descs.size)
        for i in range(self.descs.size):
            idx = self.tile_entries(i)
            if idx.size:
                out[i] = float(self.foot[idx, 3].max())
        return out

    def tile_spacing(self, i: int) -> float:
        """The cell spacing tile i's craters are prepared at (meters)."""
        return float(self.tiles["pixel"][i])

    def tile_craters(self, i: int):
        """Tile i's craters as the reference's per-grid path prepares them, at its spacing
        (resolution baked into the records): (kernel, records, data, footprint) or None."""
        idx = self.tile_entries(i)
        if idx.size == 0:
            return None
        spec = self.craters
        return N.prepare_craters(self.ordered[idx], spec.models(), spec.apply(), spec.seed,
                                 self.tile_spacing(i))

    def plain_inputs(self, i: int) -> dict:
        """Tile i's inputs as the reference's per-grid path builds them: its craters prepared at
        its pixel size (resolution baked into the records), binned into its 16 x 16 tiles."""
        t = self.tiles[i]
        grid = N.Grid(int(t["width"]), int(t["height"]), float(t["origin_x"]),
                      float(t["origin_y"]), float(t["pixel"]))
        cx64, cy64 = N.grid_centers(grid)
        xs, xlo = N.split(cx64)
        ys, ylo = N.split(cy64)
        inputs = dict(grid=grid, xs=xs, ys=ys, lows=np.concatenate([xlo, ylo]), craters=None)
        prepared = self.tile_craters(i)
        if prepared is not None:
            kernel, records, data, foot = prepared
            tiles, entries = N.bin_plane(grid, foot[:, 0], foot[:, 1], foot[:, 2], 16)
            inputs["craters"] = (kernel, records, data, tiles, entries)
            inputs["foot"] = foot
        return inputs

    def plain_run(self, inputs: dict, with_boulders: bool = True) -> np.ndarray:
        """The tile through tg_hip_run: the noise layer kernel, the crater kernel, then the
        boulder kernel when the batch has boulders (and `with_boulders`), each a synchronous call
        with host buffers, as RunKernel drives a device per layer."""
        g = inputs["grid"]
        h = np.full((g.height, g.width), self.elevation, np.float32)
        params = N.KernelParams(g.width, g.height, N.NODATA_NONE, 0.0, 0)
        if self.plan is not None:
            self.dev.run(N.TG_KERNEL_NOISE_LAYER, self.noise_kernel, h, inputs["xs"],
                         inputs["ys"], params)
        if inputs["craters"] is not None:
            kernel, records, data, tiles, entries = inputs["craters"]
            extra = [(records, 0), (data, 0), (tiles, 0), (entries, 0),
                     (np.zeros(1, N.FRAME_DTYPE), 0), (np.zeros(1, N.POINT_DTYPE), 0),
                     (inputs["lows"], 0)]
            self.dev.run(N.TG_KERNEL_CRATERS, kernel, h, inputs["xs"], inputs["ys"], params,
                         extra)
        if self.boulders is not None and with_boulders:
            # The boulder layer at the tile's spacing, as the reference prepares it per grid.
            g = inputs["grid"]
            kernel, records, foot = N.prepare_boulders(self.boulders, g.pixel)
            if records.size:
                tiles, entries = N.bin_plane(g, foot[:, 0], foot[:, 1], foot[:, 2], 16)
                self.dev.run(N.TG_KERNEL_BOULDERS, kernel, h, inputs["xs"], inputs["ys"], params,
                             [(records, 0), (tiles, 0), (entries, 0), (inputs["lows"], 0)])
        return h

    def close(self) -> None:
        for name, fn in (("batch", "tg_hip_batch_free"), ("catalogue", "tg_hip_catalogue_free"),
                         ("boulder_catalogue", "tg_hip_boulder_catalogue_free"),
                         ("plan", "tg_hip_noise_plan_free")):
            h = getattr(self, name, None)
            if h is not None:
                getattr(self.dev.lib, fn)(h)
                setattr(self, name, None)


MOON_RADIUS = 1737400.0  # meters


def body_tile_lists(descs: np.ndarray, bins: np.ndarray, entries: np.ndarray):
    """A one-tile batch's bins as the crater kernel's tile lists on a body (``[tilesX, side,
    onBody = 1, (first, count)...]``, craters.h:1341-1344) and entries."""
    d = descs[0]
    heads = [int(d["bins_across"]), 16, 1]
    for b in bins:
        heads += [int(b["first"]), int(b["count"])]
    return (np.asarray(heads, np.uint32),
            np.ascontiguousarray(entries if entries.size else np.zeros(1, np.uint32), np.uint32))


@dataclass
class SphereTileBatch(TileBatch):
    """TileBatch for longitude-latitude tiles on a sphere of `radius` meters (a tile's origin and
    pixel in degrees; tg_hip_batch_surface): each cell's noise is the stack's at its body-fixed
    position (EvaluateFractals3, src/noise/surface_noise.cpp:9-20) and the craters are placed by
    their frames (MakeCraterFrame, crater_kernel.cpp:285-301). The catalogue is a given crater
    list with longitude and latitude in its x and y columns (``global_craters`` makes one). The
    plain path per tile is tg_hip_run of TG_KERNEL_NOISE_BODY then TG_KERNEL_CRATERS on a body,
    with the tile's positions, split points and frames worked out on the host."""
    radius: float = MOON_RADIUS
    craters_given: np.ndarray | None = None

    @property
    def unit_length(self) -> float:
        """Surface::UnitLength of the sphere: meters per degree of arc (surface.cpp:49)."""
        return self.radius * (np.pi / 180.0)

    def build(self) -> None:
        spec = self.craters
        t0 = time.perf_counter()
        if self.craters_given is None:
            raise ValueError("a sphere batch takes its catalogue as craters_given")
        self.crater_list = np.ascontiguousarray(self.craters_given, N.CRATER_DTYPE)
        self.kernel, self.records, self.data, self.foot = N.prepare_craters(
            self.crater_list, spec.models(), spec.apply(), spec.seed, self.finest_pixel)
        if self.records.size != self.crater_list.size:
            raise ValueError("the catalogue's finest spacing drops craters")
        self.ordered = self.crater_list[apply_order(self.crater_list)]
        self.frames = N.sphere_frames(self.foot[:, 0], self.foot[:, 1], self.radius)
        t1 = time.perf_counter()
        diam = np.ascontiguousarray(self.foot[:, 3])
        cat = C.c_void_p()
        self.dev.call("tg_hip_catalogue_create", self.dev.handle, C.byref(self.kernel),
                      self.records.ctypes.data, self.records.size, self.data.ctypes.data,
                      self.data.size, diam.ctypes.data, C.byref(cat))
        self.catalogue = cat
        self.dev.call("tg_hip_catalogue_frames", cat, self.frames.ctypes.data, self.frames.size)
        self.index = N.CatalogueIndex(self.foot, radius=self.radius)
        t2 = time.perf_counter()
        self.plan = None
        if self.source_fractals:
            kernel = N.noise_kernel(self.source_fractals)
            plan = C.c_void_p()
            self.dev.call("tg_hip_noise_plan_create", self.dev.handle, C.byref(kernel),
                          C.byref(plan))
            self.plan = plan
            self.noise_kernel = kernel
        self.batch = None
        self.timings.update(prepare_ms=(t1 - t0) * 1e3, catalogue_upload_ms=(t2 - t1) * 1e3)
        self._build_boulders()

    def tile_spacing(self, i: int) -> float:
        """RasterGrid::Spacing of tile i: its pixel (degrees) times the unit length."""
        return float(self.tiles["pixel"][i]) * self.unit_length

    def band_lists(self, foot: np.ndarray, origin_x: float, origin_y: float, pixel: float,
                   width: int, height: int):
        """The 16 x 16 tile lists on a body (``[tilesX, 16, 1, (first, count)...]``) and entries
        of a grid of `width` x `height` cells from (origin_x, origin_y), degrees, for craters of
        footprint `foot`."""
        one = np.zeros(1, N.TILE_DTYPE)
        one[0] = (origin_x, origin_y, pixel, width, height)
        descs, bins, entries = N.sphere_bin(foot, self.radius, one, 1)
        return body_tile_lists(descs, bins, entries)

    def plain_inputs(self, i: int) -> dict:
        """Tile i's inputs as the reference's per-grid path on a body builds them: its craters
        prepared at its spacing (RasterGrid::Spacing: pixel times the unit length), binned into
        its 16 x 16 tiles, their frames, and each cell's position and split point."""
        t = self.tiles[i]
        grid = N.Grid(int(t["width"]), int(t["height"]), float(t["origin_x"]),
                      float(t["origin_y"]), float(t["pixel"]))
        inputs = dict(grid=grid, positions=N.sphere_positions(grid, self.radius), craters=None)
        prepared = self.tile_craters(i)
        if prepared is not None:
            kernel, records, data, foot = prepared
            tiles, entries = self.band_lists(foot, grid.origin_x, grid.origin_y, grid.pixel,
                                             grid.width, grid.height)
            frames = N.sphere_frames(foot[:, 0], foot[:, 1], self.radius)
            inputs["craters"] = (kernel, records, data, tiles, entries, frames,
                                 N.sphere_split_points(grid, self.radius))
            inputs["foot"] = foot
        return inputs

    def plain_run(self, inputs: dict) -> np.ndarray:
        """The tile through tg_hip_run: the body noise kernel then the crater kernel's body
        branch, each a synchronous call with host buffers."""
        g = inputs["grid"]
        h = np.full((g.height, g.width), self.elevation, np.float32)
        params = N.KernelParams(g.width, g.height, N.NODATA_NONE, 0.0, 0)
        xs, ys = np.zeros(g.width, np.float32), np.zeros(g.height, np.float32)
        if self.plan is not None:
            self.dev.run(N.TG_KERNEL_NOISE_BODY, self.noise_kernel, h, xs, ys, params,
                         [(inputs["positions"], 1)])
        if inputs["craters"] is not None:
            kernel, records, data, tiles, entries, frames, points = inputs["craters"]
            extra = [(records, 0), (data, 0), (tiles, 0), (entries, 0), (frames, 0),
                     (points, 1), (np.zeros(1, np.float32), 0)]
            self.dev.run(N.TG_KERNEL_CRATERS, kernel, h, xs, ys, params, extra)
        return h


def sphere_min_diameter(count: int, radius: float, coefficient: float, exponent: float,
                        max_diameter: float) -> float:
    """The smallest diameter for which the cumulative power law N(>=D) = coefficient D^-exponent
    per m^2, truncated at `max_diameter`, puts `count` craters on a sphere of `radius` meters."""
    area = 4.0 * np.pi * radius * radius
    return float((count / (coefficient * area) + max_diameter ** -exponent) ** (-1.0 / exponent))


def global_craters(count: int, radius: float = MOON_RADIUS, coefficient: float = 0.05,
                   exponent: float = 2.0, max_diameter: float = 20000.0,
                   seed: int = 4) -> np.ndarray:
    """A synthetic global catalogue (CRATER_DTYPE, longitude and latitude in degrees in x and y):
    `count` craters with centres uniform on the sphere, diameters from the cumulative power law
    N(>=D) = coefficient D^-exponent per m^2 between ``sphere_min_diameter`` and `max_diameter`
    (inverse transform), ages uniform over 1 Gyr, IDs 1..count."""
    rng = np.random.default_rng(seed)
    d_min = sphere_min_diameter(count, radius, coefficient, exponent, max_diameter)
    out = np.zeros(count, N.CRATER_DTYPE)
    out["x"] = rng.uniform(-180.0, 180.0, count)
    out["y"] = np.degrees(np.arcsin(rng.uniform(-1.0, 1.0, count)))
    lo, hi = d_min ** -exponent, max_diameter ** -exponent
    out["diameter"] = (lo - rng.uniform(0.0, 1.0, count) * (lo - hi)) ** (-1.0 / exponent)
    out["age_gyr"] = rng.uniform(0.0, 1.0, count)
    out["id"] = np.arange(1, count + 1, dtype=np.uint64)
    return out


def placed_global_craters(count: int, radius: float = MOON_RADIUS, coefficient: float = 0.05,
                          exponent: float = 2.0, max_diameter: float = 20000.0, seed: int = 4,
                          lifetime_gyr: float = 1.0, threads: int = 0) -> np.ndarray:
    """A global catalogue as the reference places it: GenerateCraters over the whole sphere in
    cube-sphere cells (src/layers/craters/generation/population.cpp:381-424), diameters from the
    cumulative power law N(>=D) = coefficient D^-exponent per m^2 between ``sphere_min_diameter``
    (about `count` craters expected) and `max_diameter`, ages uniform over `lifetime_gyr`;
    longitude and latitude in x and y, in the reference's order."""
    d_min = sphere_min_diameter(count, radius, coefficient, exponent, max_diameter)
    return N.generate_sphere_craters(radius, (-180.0, -90.0, 180.0, 90.0), coefficient, exponent,
                                     d_min, max_diameter, lifetime_gyr=lifetime_gyr, seed=seed,
                                     threads=threads)


def sphere_tiles(count: int, radius: float = MOON_RADIUS, seed: int = 4, min_k: int = 6,
                 max_k: int = 11, pixel_range=(0.25, 8.0)) -> np.ndarray:
    """Mixed longitude-latitude tiles on a sphere: edges 2^k + 1 with k uniform in [min_k,
    max_k], cell sizes log-uniform over `pixel_range` meters (in degrees, over the unit length),
    north-west corners uniform on the sphere (moved south of the pole where a tile would
    cross it)."""
    rng = np.random.default_rng(seed)
    tiles = np.zeros(count, N.TILE_DTYPE)
    edge = 2 ** rng.integers(min_k, max_k + 1, count) + 1
    meters = np.exp(rng.uniform(np.log(pixel_range[0]), np.log(pixel_range[1]), count))
    step = meters / (radius * (np.pi / 180.0))
    tiles["width"] = edge
    tiles["height"] = edge
    tiles["pixel"] = step
    tiles["origin_x"] = rng.uniform(-180.0, 180.0, count)
    north = np.degrees(np.arcsin(rng.uniform(-1.0, 1.0, count)))
    tiles["origin_y"] = np.clip(north, -90.0 + edge * step, 90.0)
    return tiles


def c4_tiles(count: int, side: float, seed: int = 4, min_k: int = 6, max_k: int = 11,
             pixel_range=(0.25, 8.0)) -> np.ndarray:
    """Mixed tiles: edges 2^k + 1 with k uniform in [min_k, max_k], pixel sizes log-uniform over
    `pixel_range`, west/north edges uniform over a square of `side` meters."""
    rng = np.random.default_rng(seed)
    tiles = np.zeros(count, N.TILE_DTYPE)
    edge = 2 ** rng.integers(min_k, max_k + 1, count) + 1
    tiles["width"] = edge
    tiles["height"] = edge
    tiles["pixel"] = np.exp(rng.uniform(np.log(pixel_range[0]), np.log(pixel_range[1]), count))
    tiles["origin_x"] = rng.uniform(0.0, side, count)
    tiles["origin_y"] = rng.uniform(0.0, side, count)
    return tiles


def plane_batch_view(grid: N.Grid, tiles: np.ndarray, entries: np.ndarray):
    """One grid's 16 x 16 tile lists (``tgx_bin_plane``'s ``[tilesX, side, onBody, (first,
    count)...]``) as a one-tile batch: (tile descriptors, bins, entries)."""
    across, side = int(tiles[0]), int(tiles[1])
    if side != 16:
        raise ValueError("batched bins are 16 x 16 cells")
    down = (grid.height + side - 1) // side
    descs = np.zeros(1, N.TILE_DESC_DTYPE)
    descs["origin_x"], descs["origin_y"], descs["pixel"] = (grid.origin_x, grid.origin_y,
                                                            grid.pixel)
    descs["width"], descs["height"], descs["bins_across"] = grid.width, grid.height, across
    bins = np.zeros(across * down, N.BIN_DTYPE)
    bins["first"] = tiles[3:3 + 2 * bins.size:2]
    bins["count"] = tiles[4:4 + 2 * bins.size:2]
    return descs, bins, np.ascontiguousarray(entries, np.uint32)


def crater_pair_cost(records: np.ndarray, ops_pair: float, ops_harmonic: float,
                     ops_ellipse: float) -> np.ndarray:
    """Per-crater algorithmic cost of one (cell, crater) pair inside the reach: the base pair cost,
    a term per outline harmonic, and one for an elongated outline."""
    return (ops_pair + ops_harmonic * records["harmonic_count"].astype(np.float64) +
            ops_ellipse * (records["ellipticity"] != 1.0))


class TileServer:
    """The native tile server (tgx_server_*) over a built TileBatch's resident noise plan,
    catalogue and index: continuous batching of tile requests onto the batched kernel."""

    def __init__(self, batch: TileBatch, slots: int = 3, max_batch_tiles: int = 64,
                 max_batch_cells: int = 1 << 24, pool_bytes: int = 1 << 31,
                 bin_threads: int = 0, poll_microseconds: int = 20):
        self.batch = batch
        self.lib = N.host_library()
        self.config = N.ServerConfig(slots, max_batch_tiles, max_batch_cells, pool_bytes,
                                     bin_threads, poll_microseconds)
        h = C.c_void_p()
        err = C.create_string_buffer(512)
        code = self.lib.tgx_server_create(batch.dev.handle, batch.plan, batch.elevation,
                                          batch.catalogue, batch.index.handle,
                                          C.byref(self.config), C.byref(h), err, 512)
        if code:
            raise N.HipError(err.value.decode())
        self.handle = h.value
        self._dst = {}
        if getattr(batch, "boulder_catalogue", None) is not None:
            # The batch's boulder layer, drawn after the craters in every served tile.
            if self.lib.tgx_server_boulders(self.handle, batch.boulder_catalogue,
                                            batch.boulder_index.handle, err, 512):
                raise N.HipError(err.value.decode())

    def submit(self, origin_x: float, origin_y: float, pixel: float, width: int,
               height: int) -> int:
        """Queues a tile; returns its ticket. Heights land in a new array owned by the server
        object until `wait` returns it."""
        dst = np.empty((height, width), np.float32)
        tile = N.TileDesc(origin_x, origin_y, pixel, width, height)
        ticket = C.c_uint64()
        err = C.create_string_buffer(512)
        code = self.lib.tgx_server_submit(self.handle, C.byref(tile), dst.ctypes.data,
                                          C.byref(ticket), err, 512)
        if code:
            raise N.HipError(err.value.decode())
        self._dst[ticket.value] = dst
        return ticket.value

    def wait(self, ticket: int, timeout: float = -1.0):
        """(heights, info dict) of a finished ticket; blocks until it finishes."""
        info = N.TicketInfo()
        err = C.create_string_buffer(512)
        code = self.lib.tgx_server_wait(self.handle, ticket, timeout, C.byref(info), err, 512)
        if code == N.TG_NOT_READY:
            raise TimeoutError(f"ticket {ticket} not finished after {timeout} s")
        dst = self._dst.pop(ticket)
        if code:
            raise N.HipError(err.value.decode())
        return dst, {name: getattr(info, name) for name, _ in N.TicketInfo._fields_}

    def stats(self) -> dict:
        st = N.ServerStats()
        self.lib.tgx_server_stats_get(self.handle, C.byref(st))
        return {name: getattr(st, name) for name, _ in N.ServerStats._fields_ if name != "pad"}

    def close(self) -> None:
        if getattr(self, "handle", None):
            self.lib.tgx_server_free(self.handle)
            self.handle = None

    def __del__(self):
        self.close()


def scatter_boulders(region: tuple, coefficient: float, exponent: float, min_d: float,
                     max_d: float, seed: int, rim_share: float = 0.0, craters=None,
                     min_crater_d: float = 10.0) -> np.ndarray:
    """A synthetic boulder field in BOULDER_DTYPE, for measurement: Poisson count from the
    cumulative power law N(>D) = coefficient D^-exponent per m^2 between min_d and max_d over
    `region` (min_x, min_y, max_x, max_y), diameters by its inverse CDF, uniform centres, heights
    and burial drawn as the reference's BoulderPose defaults draw them (height 0.4-0.6 D, buried
    0.1-0.4 of it), no facets. A `rim_share` of them sit around the rims of `craters` (a
    CRATER_DTYPE list) at least `min_crater_d` across, picked in proportion to rim length, 1-1.5
    radii from the centre. Not the reference's placement (BoulderLayer thins candidates by its
    abundance model); the same sizes, density and layout in distribution only."""
    rng = np.random.default_rng(seed)
    x0, y0, x1, y1 = region
    area = (x1 - x0) * (y1 - y0)
    lo, hi = min_d ** -exponent, max_d ** -exponent
    n = int(rng.poisson(coefficient * (lo - hi) * area))
    out = np.zeros(n, N.BOULDER_DTYPE)
    u = rng.random(n)
    out["diameter"] = (lo - u * (lo - hi)) ** (-1.0 / exponent)
    out["x"] = rng.uniform(x0, x1, n)
    out["y"] = rng.uniform(y0, y1, n)
    ratio = rng.uniform(0.4, 0.6, n)
    burial = rng.uniform(0.1, 0.4, n)
    out["exposed"] = out["diameter"] * ratio * (1.0 - burial)
    if rim_share > 0.0 and craters is not None and len(craters):
        big = craters[craters["diameter"] >= min_crater_d]
        if big.size:
            near = rng.random(n) < rim_share
            k = int(near.sum())
            pick = rng.choice(big.size, k, p=big["diameter"] / big["diameter"].sum())
            radius = 0.5 * big["diameter"][pick] * rng.uniform(1.0, 1.5, k)
            angle = rng.uniform(0.0, 2.0 * np.pi, k)
            out["x"][near] = big["x"][pick] + radius * np.cos(angle)
            out["y"][near] = big["y"][pick] + radius * np.sin(angle)
    return out
