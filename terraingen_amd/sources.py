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


def lattice_positions(grid: N.Grid, lattice_x: float, lattice_y: float, post: float):
    """The lattice position (in posts, from post (0, 0)) of each column's and each row's cell
    centres: GeoTransform::WorldToPixel of the lattice less half a post, in the kernel's FP64
    operations."""
    zero = 0.0
    cols = np.arange(grid.width, dtype=np.float64)
    rows = np.arange(grid.height, dtype=np.float64)
    x = grid.origin_x + (cols + 0.5) * grid.pixel + 0.5 * zero
    y = grid.origin_y + 0.5 * zero + (rows + 0.5) * -grid.pixel
    det = post * -post - zero * zero
    u = ((x - lattice_x) * -post - zero) / det - 0.5
    v = ((y - lattice_y) * post - zero) / det - 0.5
    return u, v


def tap_window(grid: N.Grid, lattice_x: float, lattice_y: float, post: float, method: int):
    """The posts a fill of `grid` reads: (col0, row0, width, height) of the lattice, one post to
    spare each side of the kernel's taps."""
    u, v = lattice_positions(grid, lattice_x, lattice_y, post)
    r = RADIUS[method]
    c0 = int(np.floor(u.min())) - r - 1
    c1 = int(np.floor(u.max())) + r + 1
    r0 = int(np.floor(v.min())) - r - 1
    r1 = int(np.floor(v.max())) + r + 1
    return c0, r0, c1 - c0 + 1, r1 - r0 + 1


def at_post_spacing(post: float, spacing: float) -> bool:
    """AtPostSpacing (upsampling.cpp:289-294) of a north-up lattice."""
    return spacing >= abs(post * -post) ** 0.5 * (1.0 - SPACING_TOLERANCE)


def upsample_spec(grid: N.Grid, lattice_x: float, lattice_y: float, post: float, window,
                  method: int, spacing: float, wrap_columns: int = 0) -> N.Upsample:
    c0, r0, w, h = window
    return N.Upsample(lattice_x, lattice_y, post, c0, r0, w, h, 1 if wrap_columns else 0,
                      wrap_columns, grid.origin_x, grid.origin_y, grid.pixel, grid.width,
                      grid.height, method, 1 if at_post_spacing(post, spacing) else 0)


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
        c0, r0, w, h = window
        xs, ys = self.post_coordinates(window)
        out = N.DeviceArray(self.dev, w * h * 4)
        d_xs = N.DeviceArray.from_host(self.dev, xs)
        d_ys = N.DeviceArray.from_host(self.dev, ys)
        params = N.KernelParams(w, h, N.NODATA_NONE, 0.0, 0)
        self.dev.call("tg_hip_noise_fill", self.dev.handle, self.plan, self.elevation, out.ptr,
                      d_xs.ptr, d_ys.ptr, C.byref(params), stream)
        self.dev.synchronize(stream)
        return out

    def fill(self, grid: N.Grid, out: N.DeviceArray, spacing: float = 0.0, stream=None):
        """SyntheticSource::FillSamples of `grid` at `spacing` (0: the grid's, as
        ResolvedSpacing gives it) into `out`; returns the spec and the posts' window."""
        spacing = max(spacing, grid.pixel)
        window = tap_window(grid, self.lattice_x, self.lattice_y, self.post, self.method)
        posts = self.posts(window, stream)
        spec = upsample_spec(grid, self.lattice_x, self.lattice_y, self.post, window,
                             self.method, spacing)
        self.dev.call("tg_hip_upsample", self.dev.handle, C.byref(spec), posts.ptr, out.ptr,
                      stream)
        self.dev.synchronize(stream)
        return spec, posts

    def close(self):
        if getattr(self, "plan", None) is not None:
            self.dev.lib.tg_hip_noise_plan_free(self.plan)
            self.plan = None


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
