# SPDX-License-Identifier: BSD-3-Clause
"""Boulder fixtures from the reference (tests/golden/boulders_*.npz) as the arrays the boulder
kernel binds: the grid, the boulder list, its prepared records and tile lists, the cell positions
and their remainders, the base heights the layer starts from and the reference's result."""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from pathlib import Path

import numpy as np

from oracle import RefBoulderCase
from terraingen_amd import native as N

GOLDEN = Path(__file__).resolve().parent / "golden"
NAMES = ["default_flat", "c3_ventifacts", "c3_fade"]
C2_FRACTAL = dict(seed=1, wavelength=512.0, amplitude=20.0, octaves=8, lacunarity=2.0, gain=0.5)


@dataclass
class BoulderCase:
    name: str
    spec: object
    grid: N.Grid
    boulders: np.ndarray
    heights: np.ndarray       # the reference's FP32 heights: source, then the boulder layer
    base: np.ndarray          # the source alone (the layer's base)
    kernel: object
    records: np.ndarray
    tiles: np.ndarray
    entries: np.ndarray
    xs: np.ndarray
    ys: np.ndarray
    lows: np.ndarray

    @property
    def shape(self):
        return self.heights.shape


def source_heights(orc, spec, grid) -> np.ndarray:
    """The case's source alone, FP32, as the reference fills it (flat, or SyntheticSource noise
    at origin {0, 0} on the oracle)."""
    h = np.full((grid.height, grid.width), spec.elevation, np.float32)
    if spec.source_kind == 1:
        kernel = N.noise_kernel([C2_FRACTAL])
        cx, cy = N.grid_centers(grid)
        xs, ys = cx.astype(np.float32), cy.astype(np.float32)
        params = N.KernelParams(grid.width, grid.height, 0, 0.0, 0)
        orc.or_noise_grid(C.byref(kernel), xs.ctypes.data, ys.ctypes.data, grid.width,
                          grid.height, C.byref(params), h.ctypes.data, 1)
    return h


def load(name: str, orc) -> BoulderCase:
    z = np.load(GOLDEN / f"boulders_{name}.npz")
    spec = RefBoulderCase.from_buffer_copy(z["case"].tobytes())
    grid = N.north_up_grid(spec.width, spec.height, spec.pixel, (spec.center_x, spec.center_y))
    boulders = np.ascontiguousarray(z["boulders"]).view(N.BOULDER_DTYPE)
    kernel, records, foot = N.prepare_boulders(boulders, spec.pixel)
    tiles, entries = N.bin_plane(grid, foot[:, 0], foot[:, 1], foot[:, 2], 16)
    cx, cy = N.grid_centers(grid)
    xs, xlo = N.split(cx)
    ys, ylo = N.split(cy)
    return BoulderCase(name, spec, grid, boulders, np.ascontiguousarray(z["heights"]),
                       source_heights(orc, spec, grid), kernel, records, tiles, entries, xs, ys,
                       np.concatenate([xlo, ylo]))


def tgx_grid(g: N.Grid) -> N.Grid:
    return N.Grid(g.width, g.height, g.origin_x, g.origin_y, g.pixel)


def oracle_heights(orc, case: BoulderCase) -> np.ndarray:
    h = case.base.copy()
    orc.or_boulder_grid(case.boulders.ctypes.data, case.boulders.size,
                        C.byref(tgx_grid(case.grid)), 1, h.ctypes.data)
    return h


def oracle_heights_f64(orc, case: BoulderCase) -> np.ndarray:
    h = case.base.astype(np.float64)
    orc.or_boulder_grid_f64(case.boulders.ctypes.data, case.boulders.size,
                            C.byref(tgx_grid(case.grid)), 1, h.ctypes.data)
    return h


def binding(case: BoulderCase) -> list[tuple[np.ndarray, int]]:
    """The boulder kernel's extra buffers in binding order: boulders, tiles, entries, lows."""
    return [(case.records, 0), (case.tiles, 0), (case.entries, 0), (case.lows, 0)]


SPHERE_NAMES = ["midlat", "polar_fade"]


@dataclass
class SphereBoulderCase:
    name: str
    spec: object
    grid: N.Grid         # longitude-latitude, degrees
    boulders: np.ndarray
    heights: np.ndarray  # the reference's FP32 heights over a raster of zeros

    @property
    def spacing(self) -> float:
        """RasterGrid::Spacing on the sphere: the cell's height in degrees times the unit
        length."""
        return self.grid.pixel * self.spec.radius * np.pi / 180.0


def load_sphere(name: str) -> SphereBoulderCase:
    from oracle import RefBoulderSphereCase

    z = np.load(GOLDEN / f"boulders_sphere_{name}.npz")
    spec = RefBoulderSphereCase.from_buffer_copy(z["case"].tobytes())
    grid = N.Grid(spec.side, spec.side, spec.lon, spec.lat, spec.step)
    boulders = np.ascontiguousarray(z["boulders"]).view(N.BOULDER_DTYPE)
    return SphereBoulderCase(name, spec, grid, boulders, np.ascontiguousarray(z["heights"]))


def sphere_oracle_f64(orc, case: SphereBoulderCase, base: np.ndarray | None = None) -> np.ndarray:
    """or_boulder_body_f64 over `base` (zeros by default), FP64."""
    h = (np.zeros(case.heights.shape) if base is None else base).astype(np.float64)
    orc.or_boulder_body_f64(case.boulders.ctypes.data, case.boulders.size,
                            C.byref(tgx_grid(case.grid)), case.spec.radius, 1, h.ctypes.data)
    return h
