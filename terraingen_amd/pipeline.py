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
    """A crater layer with a power-law density N(>=D) = sfd_coefficient D^-sfd_exponent per m^2."""
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
    seed: int = 1
    rim_sigma: float = 0.0
    noise_amplitude: float = 0.0
    noise_rim_amplitude: float = 0.0
    wall_streaks: float = 0.0
    ejecta: float = 0.0

    def models(self) -> N.CraterModels:
        return N.CraterModels(self.profile, self.complex_, self.transition_diameter,
                              self.lifetime_gyr)

    def apply(self) -> N.Apply:
        return N.Apply(self.irregular, self.peaks, 0, self.rim_sigma, self.noise_amplitude,
                       self.noise_rim_amplitude, self.wall_streaks, self.ejecta)


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
                           lifetime_gyr=spec.lifetime_gyr, seed=spec.seed)
        self.crater_list = N.generate_craters(gen)
        t1 = time.perf_counter()
        kernel, records, data, foot = N.prepare_craters(self.crater_list, spec.models(),
                                                        spec.apply(), spec.seed, g.pixel)
        t2 = time.perf_counter()
        tiles, entries = N.bin_plane(g, foot[:, 0], foot[:, 1], foot[:, 2], self.tile_side)
        t3 = time.perf_counter()
        self.kernel, self.records, self.data = kernel, records, data
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
        else:
            heights.zero()  # a flat source at 0 m

    def launch_craters(self, heights: N.DeviceArray, stream=None) -> None:
        if self.crater_set is None:
            return
        err = C.create_string_buffer(512)
        code = self.dev.lib.tg_hip_crater_launch(self.dev.handle, self.crater_set, heights.ptr,
                                                 self.d_xs.ptr, self.d_ys.ptr,
                                                 C.byref(self.params), stream, err, 512)
        if code:
            raise N.HipError(err.value.decode())

    def launch(self, heights: N.DeviceArray, stream=None) -> None:
        self.launch_noise(heights, stream)
        self.launch_craters(heights, stream)

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
