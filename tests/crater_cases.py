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
NAMES = ["c1_default", "c2_powerlaw", "c3_complex", "features_ft", "terraces_mpd", "terraces_ft"]


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


# Crater layers on a sphere (tests/golden/craters_sphere_*.npz, make_sphere_fixtures.py).
SPHERE_NAMES = ["ft_midlat", "mpd_antimeridian", "ft_dense_south"]


@dataclass
class SphereCase:
    name: str
    kernel: N.CraterShape
    records: np.ndarray
    data: np.ndarray
    frames: np.ndarray        # tg_crater_frame per crater (MakeCraterFrame)
    points: np.ndarray        # tg_split_point per cell, row-major (MakeSplitPoint)
    heights: np.ndarray       # the reference's FP32 heights: zeros shaped by the crater layer
    side: int
    lon: float                # north-west corner, degrees
    lat: float
    step: float               # cell side, degrees
    radius: float             # sphere radius, meters
    tiles: np.ndarray         # 16 x 16-cell tiles on a body, every tile listing every crater
    entries: np.ndarray

    @property
    def shape(self):
        return self.heights.shape

    def cell_lonlat(self) -> tuple[np.ndarray, np.ndarray]:
        """FP64 cell-centre longitude and latitude (degrees), each of shape (side, side)."""
        k = np.arange(self.side, dtype=np.float64) + 0.5
        lon = self.lon + k * self.step
        lat = self.lat - k * self.step
        return np.meshgrid(lon, lat)


def load_sphere(name: str) -> SphereCase:
    z = np.load(GOLDEN / f"craters_sphere_{name}.npz")
    kernel = N.CraterShape.from_buffer_copy(z["kernel"].tobytes())
    records = np.ascontiguousarray(z["records"]).view(N.RECORD_DTYPE)
    spec = z["case"].tobytes()
    # ref_sphere_case: side (u32, padded to 8), then lon, lat, step, radius (f64).
    side = int(np.frombuffer(spec[:4], np.uint32)[0])
    lon, lat, step, radius = np.frombuffer(spec[8:40], np.float64)
    n = records.size
    across = (side + 15) // 16
    heads = [across, 16, 1]
    for _ in range(across * across):
        heads += [0, n]
    return SphereCase(name, kernel, records, np.ascontiguousarray(z["data"], np.float32),
                      np.ascontiguousarray(z["frames"]).view(N.FRAME_DTYPE),
                      np.ascontiguousarray(z["points"]).view(N.POINT_DTYPE),
                      np.ascontiguousarray(z["heights"]), side, float(lon), float(lat),
                      float(step), float(radius), np.asarray(heads, np.uint32),
                      np.arange(n, dtype=np.uint32))


def sphere_position(lon: np.ndarray, lat: np.ndarray, radius: float) -> np.ndarray:
    """Body-fixed position (x, y, z) in FP64 of longitude/latitude in degrees on a sphere."""
    lam, phi = np.radians(lon), np.radians(lat)
    return radius * np.stack([np.cos(phi) * np.cos(lam), np.cos(phi) * np.sin(lam),
                              np.sin(phi)], axis=-1)


def sphere_frames64(case: SphereCase) -> np.ndarray:
    """Each crater's frame in FP64 from its centre's longitude and latitude (read off the
    fixture's split centre): centre, local east, local north, sphere radius; 10 values each."""
    f = case.frames
    x = f["x"].astype(np.float64) + f["x_low"]
    y = f["y"].astype(np.float64) + f["y_low"]
    zc = f["z"].astype(np.float64) + f["z_low"]
    lam = np.arctan2(y, x)
    phi = np.arcsin(np.clip(zc / case.radius, -1.0, 1.0))
    centre = sphere_position(np.degrees(lam), np.degrees(phi), case.radius)
    east = np.stack([-np.sin(lam), np.cos(lam), np.zeros_like(lam)], axis=-1)
    north = np.stack([-np.sin(phi) * np.cos(lam), -np.sin(phi) * np.sin(lam), np.cos(phi)],
                     axis=-1)
    out = np.concatenate([centre, east, north, np.full((lam.size, 1), case.radius)], axis=1)
    return np.ascontiguousarray(out)


def sphere_oracle_heights(orc, case: SphereCase) -> np.ndarray:
    """The FP32 oracle on a body (or_shape_grid_body) over the fixture's own points and frames."""
    out = np.zeros(case.shape, np.float32)
    params = N.KernelParams(case.side, case.side, 0, 0.0, 0)
    orc.or_shape_grid_body(C.byref(case.kernel), case.records.ctypes.data, case.data.ctypes.data,
                           case.tiles.ctypes.data, case.entries.ctypes.data,
                           case.frames.ctypes.data, case.points.ctypes.data, case.side,
                           case.side, C.byref(params), out.ctypes.data, 0)
    return out


def sphere_oracle_heights_f64(orc, case: SphereCase) -> np.ndarray:
    """The FP64 oracle on a body: cell positions and frames computed in FP64 from longitude and
    latitude, independent of the fixture's FP32 splits."""
    lon, lat = case.cell_lonlat()
    pos = np.ascontiguousarray(sphere_position(lon, lat, case.radius).reshape(-1, 3))
    frames = sphere_frames64(case)
    out = np.zeros(case.shape, np.float64)
    orc.or_shape_grid_body_f64(C.byref(case.kernel), case.records.ctypes.data,
                               case.data.ctypes.data, case.tiles.ctypes.data,
                               case.entries.ctypes.data, frames.ctypes.data, pos.ctypes.data,
                               case.side, case.side, 0, out.ctypes.data, 0)
    return out


def sphere_binding(case: SphereCase, rows: slice | None = None) -> list[tuple[np.ndarray, int]]:
    """craters.slang's buffers on a body: points are per cell (one per cell of the strip), lows
    a placeholder."""
    points = case.points if rows is None else np.ascontiguousarray(case.points[rows])
    return [(case.records, 0), (case.data, 0), (case.tiles, 0), (case.entries, 0),
            (case.frames, 0), (np.ascontiguousarray(points.reshape(-1)), 1),
            (np.zeros(1, np.float32), 0)]
