# SPDX-License-Identifier: BSD-3-Clause
"""Sources with posts on the device: heights held on a lattice of posts and upsampled onto any
grid with a kernel (tg_hip_upsample), as the reference's Upsampler::Fill serves them
(src/sources/upsampling.cpp:297-335).

* ``SyntheticPosts``: a SyntheticSource with a pixel size (src/sources/synthetic_source.cpp:
  20-53,87-95): its noise evaluated at the posts of the lattice NorthUp(-p/2, p/2, p) (post (col,
  row) at (col p, -row p)), on the device, then upsampled.
* ``RasterPosts``: the posts of a raster (a DEM read from a GeoTIFF, for instance) held on the
  device, as the reference's RasterPosts / GeoTiffPosts read them (upsampling.cpp:267-283): reads
  beyond the raster repeat its edge, columns of a raster around a body wrap.

Both fill a grid at a sample spacing as Upsampler::Fill does: at the posts' spacing or coarser a
cell centred on a post takes the post, and every other cell the kernel's interpolation.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import native as N

# KernelRadius (src/raster/resample.cpp:87-99).
RADIUS = {N.TG_RESAMPLE_NEAREST: 1, N.TG_RESAMPLE_BILINEAR: 1, N.TG_RESAMPLE_CUBIC: 2,
          N.TG_RESAMPLE_LANCZOS3: 3, N.TG_RESAMPLE_CUBIC_BSPLINE: 2}
# The kernel upsamplers' names (src/sources/upsampling.cpp:214-222).
METHODS = {"nearest": N.TG_RESAMPLE_NEAREST, "bilinear": N.TG_RESAMPLE_BILINEAR,
           "cubic": N.TG_RESAMPLE_CUBIC, "catmull_rom": N.TG_RESAMPLE_CUBIC,
           "lanczos3": N.TG_RESAMPLE_LANCZOS3, "cubic_bspline": N.TG_RESAMPLE_CUBIC_BSPLINE}
# How far below the post spacing a spacing still counts as it (upsampling.cpp:24).
SPACING_TOLERANCE = 1e-9


def lattice_positions(grid: N.Grid, lattice_x: float, lattice_y: float, post: float,
                      post_height: float = 0.0):
    """The lattice position (in posts, from post (0, 0)) of each column's and each row's cell
    centres: GeoTransform::WorldToPixel of the lattice less half a post, in the kernel's FP64
    operations."""
    zero = 0.0
    ph = -(post_height if post_height > 0.0 else post)
    cols = np.arange(grid.width, dtype=np.float64)
    rows = np.arange(grid.height, dtype=np.float64)
    x = grid.origin_x + (cols + 0.5) * grid.pixel + 0.5 * zero
    y = grid.origin_y + 0.5 * zero + (rows + 0.5) * -grid.pixel
    det = post * ph - zero * zero
    u = ((x - lattice_x) * ph - zero) / det - 0.5
    v = ((y - lattice_y) * post - zero) / det - 0.5
    return u, v


def tap_window(grid: N.Grid, lattice_x: float, lattice_y: float, post: float, method: int,
               post_height: float = 0.0):
    """The posts a fill of `grid` reads: (col0, row0, width, height) of the lattice, one post to
    spare each side of the kernel's taps."""
    u, v = lattice_positions(grid, lattice_x, lattice_y, post, post_height)
    r = RADIUS[method]
    c0 = int(np.floor(u.min())) - r - 1
    c1 = int(np.floor(u.max())) + r + 1
    r0 = int(np.floor(v.min())) - r - 1
    r1 = int(np.floor(v.max())) + r + 1
    return c0, r0, c1 - c0 + 1, r1 - r0 + 1


def at_post_spacing(post: float, spacing: float, post_height: float = 0.0) -> bool:
    """AtPostSpacing (upsampling.cpp:289-294) of a north-up lattice."""
    ph = -(post_height if post_height > 0.0 else post)
    return spacing >= abs(post * ph - 0.0 * 0.0) ** 0.5 * (1.0 - SPACING_TOLERANCE)


def upsample_spec(grid: N.Grid, lattice_x: float, lattice_y: float, post: float, window,
                  method: int, spacing: float, wrap_columns: int = 0, post_height: float = 0.0,
                  covered=None, out_pitch: int = 0) -> N.Upsample:
    """The tg_upsample of a fill of `grid` from the posts `window` (col0, row0, width, height)
    of a lattice; `covered` (col0, row0, col1, row1) marks the lattice cells with heights."""
    c0, r0, w, h = window
    cover = covered if covered is not None else (0, 0, 0, 0)
    return N.Upsample(lattice_x, lattice_y, post, post_height, c0, r0, w, h,
                      1 if wrap_columns else 0, wrap_columns, grid.origin_x, grid.origin_y,
                      grid.pixel, grid.width, grid.height, method,
                      1 if at_post_spacing(post, spacing, post_height) else 0, out_pitch,
                      1 if covered is not None else 0, *cover)


@dataclass
class SyntheticPosts:
    """A SyntheticSource whose noise lies on posts `post` meters apart, upsampled with `method`
    (cubic by default, as the reference's DefaultUpsampler)."""
    dev: N.HipDevice
    fractals: list
    elevation: float
    post: float
    method: int = N.TG_RESAMPLE_CUBIC

    def __post_init__(self):
        if not self.post > 0.0:
            raise ValueError("a synthetic source's pixel size must be positive")
        self.kernel = N.noise_kernel(self.fractals)
        plan = C.c_void_p()
        self.dev.call("tg_hip_noise_plan_create", self.dev.handle, C.byref(self.kernel),
                      C.byref(plan))
        self.plan = plan
        # The lattice NoisePosts samples: NorthUp(-p/2, p/2, p).
        self.lattice_x, self.lattice_y = -0.5 * self.post, 0.5 * self.post

    def post_coordinates(self, window):
        """The FP32 coordinates the noise takes at the window's posts: each post's world point
        (GeoTransform::PixelToWorld of its cell centre) narrowed as FractalStackRows narrows it."""
        c0, r0, w, h = window
        zero = 0.0
        cols = np.arange(c0, c0 + w, dtype=np.float64)
        rows = np.arange(r0, r0 + h, dtype=np.float64)
        x = self.lattice_x + (cols + 0.5) * self.post + 0.5 * zero
        y = self.lattice_y + 0.5 * zero + (rows + 0.5) * -self.post
        return x.astype(np.float32), y.astype(np.float32)

    def posts(self, window, stream=None) -> N.DeviceArray:
        """The noise at the window's posts on the device (tg_hip_noise_fill), row by row."""
        fill = self.prepare_window(window)
        fill.launch_posts(stream)
        self.dev.synchronize(stream)
        return fill.d_posts

    def prepare_window(self, window) -> "PreparedFill":
        c0, r0, w, h = window
        xs, ys = self.post_coordinates(window)
        return PreparedFill(self, window, N.DeviceArray(self.dev, w * h * 4),
                            N.DeviceArray.from_host(self.dev, xs),
                            N.DeviceArray.from_host(self.dev, ys), None)

    def prepare(self, grid: N.Grid, spacing: float = 0.0) -> "PreparedFill":
        """The device buffers and the spec of a fill of `grid` at `spacing` (0: the grid's, as
        ResolvedSpacing gives it), so the fill can be launched again allocating nothing."""
        spacing = max(spacing, grid.pixel)
        window = tap_window(grid, self.lattice_x, self.lattice_y, self.post, self.method)
        fill = self.prepare_window(window)
        fill.spec = upsample_spec(grid, self.lattice_x, self.lattice_y, self.post, window,
                                  self.method, spacing)
        return fill

    def fill(self, grid: N.Grid, out: N.DeviceArray, spacing: float = 0.0, stream=None):
        """SyntheticSource::FillSamples of `grid` at `spacing` (0: the grid's) into `out`;
        returns the spec and the posts."""
        fill = self.prepare(grid, spacing)
        fill.launch(out, stream)
        self.dev.synchronize(stream)
        return fill.spec, fill.d_posts

    def close(self):
        if getattr(self, "plan", None) is not None:
            self.dev.lib.tg_hip_noise_plan_free(self.plan)
            self.plan = None


@dataclass
class PreparedFill:
    """A SyntheticPosts fill with its device buffers made: the posts' window, its FP32
    coordinates, the posts and the upsampling spec. Launches queue on `stream` and do not wait."""
    source: SyntheticPosts
    window: tuple
    d_posts: N.DeviceArray
    d_xs: N.DeviceArray
    d_ys: N.DeviceArray
    spec: N.Upsample | None

    def launch_posts(self, stream=None):
        """The noise at the window's posts (tg_hip_noise_fill), row by row."""
        src = self.source
        _, _, w, h = self.window
        params = N.KernelParams(w, h, N.NODATA_NONE, 0.0, 0)
        src.dev.call("tg_hip_noise_fill", src.dev.handle, src.plan, src.elevation,
                     self.d_posts.ptr, self.d_xs.ptr, self.d_ys.ptr, C.byref(params), stream)

    def launch_upsample(self, out: N.DeviceArray, stream=None):
        """The grid from the posts (tg_hip_upsample)."""
        src = self.source
        src.dev.call("tg_hip_upsample", src.dev.handle, C.byref(self.spec), self.d_posts.ptr,
                     out.ptr, stream)

    def launch(self, out: N.DeviceArray, stream=None):
        self.launch_posts(stream)
        self.launch_upsample(out, stream)


@dataclass
class RasterPosts:
    """The posts of a north-up raster (`values`, NaN where there is no data) whose north-west
    corner is (origin_x, origin_y) and whose cells are `post` apart, held on the device: a DEM's
    cells as GeoTiffSource serves them. `wrap_columns`: the raster spans every longitude of a body
    and its columns wrap."""
    dev: N.HipDevice
    values: np.ndarray
    origin_x: float
    origin_y: float
    post: float
    method: int = N.TG_RESAMPLE_CUBIC
    wrap_columns: bool = False

    def __post_init__(self):
        v = np.ascontiguousarray(self.values, np.float32)
        if v.ndim != 2 or v.size == 0:
            raise ValueError("raster posts need a non-empty 2-D raster")
        self.values = v
        self.d_posts = N.DeviceArray.from_host(self.dev, v)

    def fill(self, grid: N.Grid, out: N.DeviceArray, spacing: float = 0.0, stream=None):
        spacing = max(spacing, grid.pixel)
        h, w = self.values.shape
        spec = upsample_spec(grid, self.origin_x, self.origin_y, self.post, (0, 0, w, h),
                             self.method, spacing, w if self.wrap_columns else 0)
        self.dev.call("tg_hip_upsample", self.dev.handle, C.byref(spec), self.d_posts.ptr,
                      out.ptr, stream)
        self.dev.synchronize(stream)
        return spec


TURN_DEGREES = 360.0


class GeoTiffPosts:
    """A GeoTiffSource on the device (src/sources/geotiff_source.cpp): one band of a GeoTIFF
    read on the host (tgx_geotiff_read: decoded, scaled, NaN where there is no data), the window
    of it the source's cells need held on the device, and grids filled from it by
    tg_hip_upsample as RasterHeights fills them (src/sources/raster_source.cpp:78-179): NaN
    where a cell's centre lies outside the cells covered (the whole file, or those that meet
    `crop`), a raster in longitude and latitude spanning 360 degrees wrapping in longitude."""

    def __init__(self, dev: N.HipDevice | None, path, band: int = 0,
                 method: int = N.TG_RESAMPLE_CUBIC, crop=None, scale=None, offset=None):
        self.dev = dev
        self.method = method
        self.file = N.GeoTiffFile(path)
        info, file_scale, file_offset = self.file.info(band)
        self.info = info
        t = list(info.transform)  # GDAL order: originX, pixelWidth, rowRot, originY, colRot, ph
        if t[2] != 0.0 or t[4] != 0.0 or not t[1] > 0.0 or not t[5] < 0.0:
            raise ValueError("the file's raster is not north-up")
        self.origin_x, self.post, self.origin_y, self.post_height = t[0], t[1], t[3], -t[5]
        self.scale = file_scale if scale is None else scale
        self.offset = file_offset if offset is None else offset
        self.no_data = info.no_data if info.has_no_data else None
        w, h = info.width, info.height
        self.cells = self.covered_cells(crop)
        # NativeWindow (raster_source.cpp:58-66): the cells and the kernel's reach around them.
        margin = RADIUS[method] + 1
        c0, c1, r0, r1 = (max(0, self.cells[0] - margin), min(w, self.cells[2] + margin),
                          max(0, self.cells[1] - margin), min(h, self.cells[3] + margin))
        # WrapsLongitude (raster_source.cpp:45-50).
        self.wraps = bool(info.geographic) and abs(w * self.post - TURN_DEGREES) <= \
            1e-9 * TURN_DEGREES
        if self.wraps:
            c0, c1 = 0, w
        self.window = (c0, r0, c1 - c0, r1 - r0)
        self.values = self.file.read(band, c0, r0, c1 - c0, r1 - r0, self.scale, self.offset)
        self.d_posts = N.DeviceArray.from_host(dev, self.values) if dev is not None else None

    def covered_cells(self, crop):
        """(col0, row0, col1, row1): the file's cells that meet `crop` (PixelWindowIn of the box's
        corners, cell_loops.cpp:21-28), or all of them."""
        w, h = self.info.width, self.info.height
        if crop is None:
            return (0, 0, w, h)
        x0, y0, x1, y1 = crop
        if not (x0 < x1 and y0 < y1) or not np.all(np.isfinite(crop)):
            raise ValueError("the crop must be finite with min < max")
        ph = -self.post_height
        det = self.post * ph - 0.0 * 0.0
        us = [((x - self.origin_x) * ph - 0.0) / det for x in (x0, x1)]
        vs = [((y - self.origin_y) * self.post - 0.0) / det for y in (y0, y1)]
        clamp = lambda v, n: int(min(max(v, 0.0), float(n)))  # noqa: E731
        cells = (clamp(np.floor(min(us)), w), clamp(np.floor(min(vs)), h),
                 clamp(np.ceil(max(us)), w), clamp(np.ceil(max(vs)), h))
        if cells[0] >= cells[2] or cells[1] >= cells[3]:
            raise ValueError("the crop misses the file")
        return cells

    def spec(self, grid: N.Grid, spacing: float, out_pitch: int = 0) -> N.Upsample:
        return upsample_spec(grid, self.origin_x, self.origin_y, self.post, self.window,
                             self.method, spacing, self.info.width if self.wraps else 0,
                             self.post_height, self.cells, out_pitch)

    def runs(self, grid: N.Grid):
        """The grid's runs of columns a whole number of turns of longitude apart from the
        lattice (RasterHeights::FillGrid, raster_source.cpp:154-179): (first, end, grid of the
        run moved onto the lattice's longitudes)."""
        if not self.wraps:
            return [(0, grid.width, grid)]
        left = min(self.origin_x, self.origin_x + self.info.width * self.post)
        cx = grid.origin_x + (np.arange(grid.width) + 0.5) * grid.pixel + 0.5 * 0.0
        turns = np.floor((cx - left) / TURN_DEGREES)
        if turns[0] == 0.0 and turns[-1] == 0.0:
            return [(0, grid.width, grid)]
        out, begin = [], 0
        while begin < grid.width:
            end = begin + 1
            while end < grid.width and turns[end] == turns[begin]:
                end += 1
            origin = grid.origin_x + begin * grid.pixel - turns[begin] * TURN_DEGREES
            out.append((begin, end, N.Grid(end - begin, grid.height, origin, grid.origin_y,
                                           grid.pixel)))
            begin = end
        return out

    def specs(self, grid: N.Grid, spacing: float = 0.0):
        """(first column, tg_upsample) of each run of the grid's columns: what a fill of `grid` at
        `spacing` (in the lattice's units; 0: the grid's) launches."""
        spacing = max(spacing, grid.pixel)
        return [(begin, self.spec(part, spacing, grid.width))
                for begin, end, part in self.runs(grid)]

    def fill(self, grid: N.Grid, out: N.DeviceArray, spacing: float = 0.0, stream=None):
        """Source::Fill of `grid` (width x height floats into `out`) at `spacing` (0: the
        grid's): NaN where the source has no height."""
        for begin, spec in self.specs(grid, spacing):
            self.dev.call("tg_hip_upsample", self.dev.handle, C.byref(spec), self.d_posts.ptr,
                          out.ptr + 4 * begin, stream)
        self.dev.synchronize(stream)
