# SPDX-License-Identifier: BSD-3-Clause
"""Cases of sources with posts: a SyntheticSource with a pixel size upsampled by each kernel onto
grids finer than, at and coarser than its posts (tests/golden/sources_synthetic.npz holds the
reference's own heights for them, written by tests/golden/make_source_fixtures.py)."""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

from terraingen_amd import native as N
from terraingen_amd import sources as S

GOLDEN = Path(__file__).resolve().parent / "golden"
# Short wavelengths, so neighbouring posts differ by a good share of the amplitude.
FRACTALS = [dict(seed=3, wavelength=9.0, amplitude=4.0, octaves=4, lacunarity=2.0, gain=0.5),
            dict(seed=8, wavelength=40.0, amplitude=10.0, octaves=3, type=1, offset=0.9,
                 blend=0)]
ELEVATION = 2.5

# name: (post, method, width, height, origin_x, origin_y, pixel, spacing)
CASES = {
    "cubic_finer": (2.0, N.TG_RESAMPLE_CUBIC, 97, 65, -13.3, 21.7, 0.5, 0.0),
    "bilinear_finer": (1.5, N.TG_RESAMPLE_BILINEAR, 80, 70, 100.1, -40.2, 0.3, 0.0),
    "nearest_finer": (1.0, N.TG_RESAMPLE_NEAREST, 64, 64, -7.0, -3.0, 0.25, 0.0),
    "lanczos_finer": (4.0, N.TG_RESAMPLE_LANCZOS3, 72, 60, 33.0, 12.0, 1.0, 0.0),
    "bspline_on_lattice": (1.0, N.TG_RESAMPLE_CUBIC_BSPLINE, 50, 40, -10.5, 20.5, 1.0, 0.0),
    "bspline_at_spacing": (1.0, N.TG_RESAMPLE_CUBIC_BSPLINE, 50, 40, -10.2, 20.5, 1.0, 0.0),
    "cubic_coarser": (2.0, N.TG_RESAMPLE_CUBIC, 45, 33, -1.0, 41.0, 3.0, 0.0),
    "cubic_finer_at_post_spacing": (2.0, N.TG_RESAMPLE_CUBIC, 60, 50, 0.3, 0.7, 0.5, 2.0),
}


def case_grid(name: str) -> N.Grid:
    post, method, w, h, ox, oy, pixel, spacing = CASES[name]
    return N.Grid(w, h, ox, oy, pixel)


def oracle_posts(orc, window, post: float) -> np.ndarray:
    """The SyntheticSource's noise at the window's posts on the oracle (bit-exact to the
    reference's FractalStackRows)."""
    c0, r0, w, h = window
    cols = np.arange(c0, c0 + w, dtype=np.float64)
    rows = np.arange(r0, r0 + h, dtype=np.float64)
    xs = (-0.5 * post + (cols + 0.5) * post + 0.0).astype(np.float32)
    ys = (0.5 * post + 0.0 + (rows + 0.5) * -post).astype(np.float32)
    kernel = N.noise_kernel(FRACTALS)
    params = N.KernelParams(w, h, 0, 0.0, 0)
    out = np.full((h, w), ELEVATION, np.float32)
    orc.or_noise_grid(C.byref(kernel), xs.ctypes.data, ys.ctypes.data, w, h, C.byref(params),
                      out.ctypes.data, 1)
    return out


def oracle_heights(orc, name: str):
    """(heights, spec, posts) of a case on the oracle: posts by the oracle's noise, the fill by
    the oracle's upsampling."""
    post, method, w, h, ox, oy, pixel, spacing = CASES[name]
    grid = case_grid(name)
    window = S.tap_window(grid, -0.5 * post, 0.5 * post, post, method)
    posts = oracle_posts(orc, window, post)
    spec = S.upsample_spec(grid, -0.5 * post, 0.5 * post, post, window, method,
                           max(spacing, pixel))
    out = np.zeros((h, w), np.float32)
    orc.or_upsample_grid(C.byref(spec), posts.ctypes.data, out.ctypes.data)
    return out, spec, posts


def reference_heights(ref, name: str) -> np.ndarray:
    post, method, w, h, ox, oy, pixel, spacing = CASES[name]
    kernel = N.noise_kernel(FRACTALS)
    out = np.zeros((h, w), np.float32)
    code = ref.ref_synthetic_grid(C.cast(C.byref(kernel.fractals), C.c_void_p), kernel.count,
                                  ELEVATION, post, method, w, h, ox, oy, pixel, spacing,
                                  out.ctypes.data)
    if code != 0:
        raise RuntimeError(ref.ref_sources_error().decode())
    return out


def golden_heights(name: str) -> np.ndarray:
    return np.load(GOLDEN / "sources_synthetic.npz")[name]


# GeoTiffSource cases: (file of tiff_craft.cases(), method, crop or None, grid (width, height,
# origin_x, origin_y, pixel), spacing). Grids finer than the cells, coarser, cropped, across the
# file's edges, and across the antimeridian of a file spanning every longitude.
GEOTIFF_CASES = {
    "int16_cubic_finer": ("int16_be_strips_deflate_pred2", N.TG_RESAMPLE_CUBIC, None,
                          (90, 70, 495.3, 803.1, 0.75), 0.0),
    "int16_bilinear_cropped": ("int16_be_strips_deflate_pred2", N.TG_RESAMPLE_BILINEAR,
                               (530.0, 760.0, 580.0, 790.0), (80, 60, 520.1, 795.7, 1.1), 0.0),
    "uint8_nearest_coarser": ("uint8_tiles_packbits", N.TG_RESAMPLE_NEAREST, None,
                              (30, 25, 503.0, 797.0, 3.3), 0.0),
    "float64_bspline_point": ("float64_planar_bigtiff_point", N.TG_RESAMPLE_CUBIC_BSPLINE, None,
                              (70, 50, -121.0, 45.3, 0.1), 0.0),
    "float32_lanczos_on_lattice": ("float32_be_pred3_deflate", N.TG_RESAMPLE_LANCZOS3, None,
                                   (40, 30, 510.0, 790.0, 2.0), 0.0),
    "int32_cubic_antimeridian": ("int32_geographic_wrap", N.TG_RESAMPLE_CUBIC, None,
                                 (100, 60, 160.3, 70.2, 0.55), 0.0),
}


# Files in longitude and latitude, evaluated on a sphere of the Moon's radius.
GEOGRAPHIC = {"int32_geographic_wrap": True}
MOON_RADIUS = 1737400.0


def geotiff_path(tmp_dir, file_case: str):
    """The case's file written into `tmp_dir` (the same bytes every time)."""
    import tiff_craft

    path = Path(tmp_dir) / f"{file_case}.tif"
    if not path.exists():
        tiff_craft.write(path, tiff_craft.cases()[file_case])
    return path


def reference_geotiff_heights(ref, name: str, tmp_dir) -> np.ndarray:
    file_case, method, crop, (w, h, ox, oy, pixel), spacing = GEOTIFF_CASES[name]
    path = geotiff_path(tmp_dir, file_case)
    box = np.asarray(crop, np.float64) if crop is not None else None
    out = np.zeros((h, w), np.float32)
    radius = MOON_RADIUS if GEOGRAPHIC.get(file_case) else 0.0
    code = ref.ref_geotiff_grid(str(path).encode(), method,
                                box.ctypes.data if box is not None else None, w, h, ox, oy,
                                pixel, spacing, radius, out.ctypes.data)
    if code != 0:
        raise RuntimeError(ref.ref_sources_error().decode())
    return out


def golden_geotiff_heights(name: str) -> np.ndarray:
    return np.load(GOLDEN / "sources_geotiff.npz")[name]


def marked(heights: np.ndarray, posts) -> np.ndarray:
    """NaN cells as Terrain::Evaluate marks them (MarkMissing, src/sources/source.cpp:10-19): the
    file's no-data value when its band is not scaled, NaN otherwise."""
    if posts.no_data is not None and posts.scale == 1.0 and posts.offset == 0.0:
        return np.where(np.isnan(heights), np.float32(posts.no_data), heights)
    return heights


def geotiff_posts(dev, name: str, tmp_dir):
    from terraingen_amd.sources import GeoTiffPosts

    file_case, method, crop, grid, spacing = GEOTIFF_CASES[name]
    return GeoTiffPosts(dev, geotiff_path(tmp_dir, file_case), method=method, crop=crop)


def oracle_geotiff_heights(orc, name: str, tmp_dir) -> np.ndarray:
    """The case on the CPU: the posts this repository's reader reads, filled by the oracle's
    upsampling run by run."""
    posts = geotiff_posts(None, name, tmp_dir)
    file_case, method, crop, (w, h, ox, oy, pixel), spacing = GEOTIFF_CASES[name]
    out = np.zeros((h, w), np.float32)
    for begin, spec in posts.specs(N.Grid(w, h, ox, oy, pixel), spacing):
        orc.or_upsample_grid(C.byref(spec), posts.values.ctypes.data, out[:, begin:].ctypes.data
                             if begin else out.ctypes.data)
    return marked(out, posts)
