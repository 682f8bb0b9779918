# SPDX-License-Identifier: BSD-3-Clause
"""Sources with posts on the MI355X (tg_hip_upsample): a SyntheticSource with a pixel size filled
on the device equals the reference's own heights bit for bit for the nearest, bilinear, cubic and
B-spline kernels (Lanczos within an ulp, its sine being the device's), at spacings finer than,
at and coarser than the posts; raster posts with holes and wrapped columns equal the oracle's
fill."""
import ctypes as C

import numpy as np
import pytest

from terraingen_amd import native as N
from terraingen_amd import sources as S

from source_cases import CASES, ELEVATION, FRACTALS, case_grid, golden_heights, oracle_posts

pytestmark = pytest.mark.gpu


def ulps(a: np.ndarray, b: np.ndarray) -> int:
    ia = a.view(np.int32).astype(np.int64)
    ib = b.view(np.int32).astype(np.int64)
    return int(np.abs(ia - ib).max())


@pytest.mark.parametrize("name", CASES)
def test_synthetic_source_with_posts_matches_the_reference(hip, name):
    post, method, w, h, ox, oy, pixel, spacing = CASES[name]
    src = S.SyntheticPosts(hip, FRACTALS, ELEVATION, post, method)
    try:
        grid = case_grid(name)
        out = N.DeviceArray(hip, w * h * 4)
        src.fill(grid, out, spacing)
        got = out.download(np.empty((h, w), np.float32))
    finally:
        src.close()
    want = golden_heights(name)
    print(f"[{name}] max |dh| {np.abs(got - want).max():.3g} m, {ulps(got, want)} ulp",
          flush=True)
    if method == N.TG_RESAMPLE_LANCZOS3:
        assert ulps(got, want) <= 1
    else:
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32))


def test_device_posts_are_the_oracle_noise(hip, orc):
    src = S.SyntheticPosts(hip, FRACTALS, ELEVATION, 1.5, N.TG_RESAMPLE_CUBIC)
    try:
        window = (-7, 3, 40, 33)
        got = src.posts(window).download(np.empty((33, 40), np.float32))
    finally:
        src.close()
    assert np.array_equal(got.view(np.uint32), oracle_posts(orc, window, 1.5).view(np.uint32))


@pytest.mark.parametrize("method", [N.TG_RESAMPLE_NEAREST, N.TG_RESAMPLE_BILINEAR,
                                    N.TG_RESAMPLE_CUBIC, N.TG_RESAMPLE_CUBIC_BSPLINE])
@pytest.mark.parametrize("wrap", [False, True])
def test_raster_posts_with_holes_match_the_oracle(hip, orc, method, wrap):
    """A DEM-like raster with no-data holes: NaN exactly where a weighed tap has none, the edges
    repeated beyond the raster, and columns wrapping when the raster spans the lattice."""
    rng = np.random.default_rng(4)
    values = rng.normal(0.0, 50.0, (60, 90)).astype(np.float32)
    values[20:23, 30:34] = np.nan
    values[50, 5] = np.nan
    posts = S.RasterPosts(hip, values, origin_x=-180.0, origin_y=60.0, post=4.0, method=method,
                          wrap_columns=wrap)
    # Cells of 1.3 from west of the raster across its east edge (where wrapping reads its west
    # columns) and past its south edge.
    grid = N.Grid(300, 210, -185.7, 63.1, 1.3)
    out = N.DeviceArray(hip, grid.width * grid.height * 4)
    spec = posts.fill(grid, out)
    got = out.download(np.empty((grid.height, grid.width), np.float32))
    want = np.zeros_like(got)
    orc.or_upsample_grid(C.byref(spec), values.ctypes.data, want.ctypes.data)
    assert np.array_equal(np.isnan(got), np.isnan(want))
    assert np.isnan(got).sum() > 0
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))


from source_cases import GEOTIFF_CASES, geotiff_posts, golden_geotiff_heights, marked  # noqa: E402


@pytest.mark.parametrize("name", GEOTIFF_CASES)
def test_geotiff_source_on_the_device_matches_the_reference(hip, name, tmp_path):
    """A GeoTiffSource's heights on the device (the reader's posts, tg_hip_upsample run by run
    across the antimeridian) against the reference's own (tests/golden/sources_geotiff.npz):
    the same bits and the same cells without data, cropped or not, finer and coarser."""
    posts = geotiff_posts(hip, name, tmp_path)
    _, _, _, (w, h, ox, oy, pixel), spacing = GEOTIFF_CASES[name]
    out = N.DeviceArray(hip, w * h * 4)
    posts.fill(N.Grid(w, h, ox, oy, pixel), out, spacing)
    got = marked(out.download(np.empty((h, w), np.float32)), posts)
    want = golden_geotiff_heights(name)
    assert np.array_equal(np.isnan(got), np.isnan(want))
    ok = ~np.isnan(want)
    if GEOTIFF_CASES[name][1] == N.TG_RESAMPLE_LANCZOS3:
        assert ulps(got[ok], want[ok]) <= 1
    else:
        assert np.array_equal(got[ok].view(np.uint32), want[ok].view(np.uint32))
