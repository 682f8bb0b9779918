# SPDX-License-Identifier: BSD-3-Clause
"""Crater fixtures from the reference (tests/golden/craters_*.npz) as the arrays the crater kernel
binds: kernel struct, records, data, tiles, entries, column/row positions and their remainders."""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from pathlib import Path

import numpy as np

from terraingen_amd import native as N

GOLDEN = Path(__file__).resolve().parent / "golden"
NAMES = ["c1_default", "c2_powerlaw", "c3_complex", "features_ft"]


@dataclass
class CraterCase:
    name: str
    kernel: N.CraterShape
    records: np.ndarray
    data: np.ndarray
    grid: N.Grid
    heights: np.ndarray       # the reference's FP32 heights (flat source at 0 m plus craters)
    crater_list: dict
    tiles: np.ndarray
    entries: np.ndarray
    xs: np.ndarray            # float32 column x (origin 0, as ShapeTilesOnGpu binds it)
    ys: np.ndarray
    lows: np.ndarray          # x remainders per column, then y remainders per row
    cx64: np.ndarray          # FP64 cell centres
    cy64: np.ndarray

    @property
    def shape(self):
        return self.heights.shape


def load(name: str, side: int = 16) -> CraterCase:
    z = np.load(GOLDEN / f"craters_{name}.npz")
    kernel = N.CraterShape.from_buffer_copy(z["kernel"].tobytes())
    records = np.ascontiguousarray(z["records"]).view(N.RECORD_DTYPE)
    data = np.ascontiguousarray(z["data"], np.float32)
    spec = z["case"].tobytes()
    # ref_case starts with width, height (u32) then pixel, center_x, center_y (f64).
    width, height = np.frombuffer(spec[:8], np.uint32)
    pixel, cx, cy = np.frombuffer(spec[8:32], np.float64)
    grid = N.north_up_grid(int(width), int(height), float(pixel), (float(cx), float(cy)))
    cx64, cy64 = N.grid_centers(grid)
    xs, xlo = N.split(cx64)
    ys, ylo = N.split(cy64)
    centers_x = records["x"].astype(np.float64) + records["x_low"]
    centers_y = records["y"].astype(np.float64) + records["y_low"]
    reach = N.record_reach(records, data, kernel.reach_radii)
    tiles, entries = N.bin_plane(grid, centers_x, centers_y, reach, side)
    return CraterCase(name, kernel, records, data, grid, np.ascontiguousarray(z["heights"]),
                      {k: z[k] for k in ("x", "y", "d", "age", "id")}, tiles, entries, xs, ys,
                      np.concatenate([xlo, ylo]), cx64, cy64)


def oracle_heights(orc, case: CraterCase, base: np.ndarray | None = None) -> np.ndarray:
    h, w = case.shape
    out = np.zeros((h, w), np.float32) if base is None else base.copy()
    params = N.KernelParams(w, h, 0, 0.0, 0)
    orc.or_shape_grid(C.byref(case.kernel), case.records.ctypes.data, case.data.ctypes.data,
                      case.tiles.ctypes.data, case.entries.ctypes.data, case.xs.ctypes.data,
                      case.ys.ctypes.data, case.lows.ctypes.data, w, h, C.byref(params),
                      out.ctypes.data, 0)
    return out


def oracle_heights_f64(orc, case: CraterCase) -> np.ndarray:
    h, w = case.shape
    out = np.zeros((h, w), np.float64)
    orc.or_shape_grid_f64(C.byref(case.kernel), case.records.ctypes.data, case.data.ctypes.data,
                          case.tiles.ctypes.data, case.entries.ctypes.data, case.cx64.ctypes.data,
                          case.cy64.ctypes.data, w, h, 0, out.ctypes.data, 0)
    return out


def binding(case: CraterCase) -> list[tuple[np.ndarray, int]]:
    """The extra buffers of craters.slang in binding order (apply.cpp:181-188): records, data,
    tiles, entries, frames (a placeholder on a plane), points (one placeholder), lows."""
    frames = np.zeros(1, N.FRAME_DTYPE)
    points = np.zeros(1, N.POINT_DTYPE)
    return [(case.records, 0), (case.data, 0), (case.tiles, 0), (case.entries, 0), (frames, 0),
            (points, 0), (case.lows, 0)]
