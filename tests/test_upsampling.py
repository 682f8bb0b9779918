# SPDX-License-Identifier: BSD-3-Clause
"""The upsampling oracle (oracle/upsample_oracle.cpp) against the reference's own SyntheticSource
heights (tests/golden/sources_synthetic.npz, and oracle/_ref live where it is built), and the
kernels' weights against their closed forms."""
import ctypes as C

import numpy as np
import pytest

from terraingen_amd import native as N

from source_cases import CASES, golden_heights, oracle_heights, reference_heights

KERNELS = [N.TG_RESAMPLE_BILINEAR, N.TG_RESAMPLE_CUBIC, N.TG_RESAMPLE_LANCZOS3,
           N.TG_RESAMPLE_CUBIC_BSPLINE]


@pytest.mark.parametrize("name", CASES)
def test_oracle_reproduces_the_reference_fixture(orc, name):
    got, spec, _ = oracle_heights(orc, name)
    want = golden_heights(name)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
    assert np.unique(got).size > 100  # the posts vary across the grid


@pytest.mark.parametrize("name", ["cubic_finer", "bspline_at_spacing"])
def test_oracle_matches_the_live_reference(orc, ref, name):
    got, _, _ = oracle_heights(orc, name)
    want = reference_heights(ref, name)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))


def weights(orc, method, fraction):
    out = np.zeros(6)
    orc.or_kernel_weights(method, fraction, out.ctypes.data)
    return out


@pytest.mark.parametrize("method", KERNELS)
def test_weights_are_a_partition_of_unity_that_reproduces_lines(orc, method):
    """Every kernel's weights sum to 1 and place a linear ramp of posts exactly (the taps run
    from 1 - radius to radius around the point)."""
    radius = {N.TG_RESAMPLE_BILINEAR: 1, N.TG_RESAMPLE_LANCZOS3: 3}.get(method, 2)
    taps = np.arange(1 - radius, radius + 1)
    for fraction in np.linspace(0.0, 0.999, 37):
        w = weights(orc, method, fraction)[:taps.size]
        assert abs(w.sum() - 1.0) <= 1e-14
        if method != N.TG_RESAMPLE_LANCZOS3:  # the normalized Lanczos only nearly does
            assert abs((w * taps).sum() - fraction) <= 1e-14


def test_weights_closed_forms(orc):
    # Catmull-Rom passes through the posts; the B-spline smooths them (1/6, 2/3, 1/6).
    assert np.array_equal(weights(orc, N.TG_RESAMPLE_CUBIC, 0.0)[:4], [0.0, 1.0, 0.0, 0.0])
    assert np.allclose(weights(orc, N.TG_RESAMPLE_CUBIC_BSPLINE, 0.0)[:4],
                       [1 / 6, 2 / 3, 1 / 6, 0.0], atol=1e-16)
    # Keys' kernel at the midpoint: (-1/16, 9/16, 9/16, -1/16).
    assert np.array_equal(weights(orc, N.TG_RESAMPLE_CUBIC, 0.5)[:4],
                          [-0.0625, 0.5625, 0.5625, -0.0625])
    # Lanczos 3 is 0 at whole offsets but 0.
    assert np.array_equal(weights(orc, N.TG_RESAMPLE_LANCZOS3, 0.0), [0, 0, 1, 0, 0, 0])


def test_nan_posts_spread_to_the_cells_that_weigh_them(orc):
    """A post without data gives NaN to every cell whose taps of non-zero weight reach it, and
    to no other (InterpolateLatticeWeights returns nothing on a missing tap)."""
    posts = np.arange(100, dtype=np.float32).reshape(10, 10)
    posts[5, 5] = np.nan
    spec = N.Upsample(0.0, 10.0, 1.0, 0.0, 0, 0, 10, 10, 0, 0, 0.0, 10.0, 0.25, 40, 40,
                      N.TG_RESAMPLE_BILINEAR, 0)
    out = np.zeros((40, 40), np.float32)
    orc.or_upsample_grid(C.byref(spec), posts.ctypes.data, out.ctypes.data)
    bad = np.argwhere(np.isnan(out))
    # Bilinear: cells whose lattice position lies within one post of (5, 5), open on the far
    # side of a zero weight.
    u = (np.arange(40) + 0.5) * 0.25 - 0.5
    near = np.abs(u - 5.0) < 1.0
    want = np.argwhere(near[:, None] & near[None, :])
    assert np.array_equal(bad, want)


# ---- GeoTIFF DEMs -------------------------------------------------------------------------------

from source_cases import (GEOTIFF_CASES, golden_geotiff_heights,  # noqa: E402
                          oracle_geotiff_heights, reference_geotiff_heights)


def same_heights(a: np.ndarray, b: np.ndarray) -> bool:
    return (np.array_equal(np.isnan(a), np.isnan(b)) and
            np.array_equal(a[~np.isnan(a)].view(np.uint32), b[~np.isnan(b)].view(np.uint32)))


@pytest.mark.parametrize("name", GEOTIFF_CASES)
def test_geotiff_source_on_the_cpu_matches_the_reference(orc, name, tmp_path):
    """This repository's reader and the oracle's upsampling against the reference's
    GeoTiffSource (tests/golden/sources_geotiff.npz): same heights, same cells without data."""
    got = oracle_geotiff_heights(orc, name, tmp_path)
    want = golden_geotiff_heights(name)
    assert same_heights(got, want)
    assert np.isfinite(want).sum() > want.size // 5


def test_geotiff_source_matches_the_live_reference(orc, ref, tmp_path):
    for name in ("int16_bilinear_cropped", "int32_cubic_antimeridian"):
        assert same_heights(oracle_geotiff_heights(orc, name, tmp_path),
                            reference_geotiff_heights(ref, name, tmp_path))


def test_reader_matches_the_reference_reader(ref, tmp_path):
    """Every layout of tests/tiff_craft.py, and this repository's and the reference's own LZW
    GeoTIFFs, read to the reference's ReadGeoTiff samples (no-data samples as NaN)."""
    import tiff_craft

    from source_cases import GOLDEN

    paths = []
    for name, spec in tiff_craft.cases().items():
        path = tmp_path / f"{name}.tif"
        tiff_craft.write(path, spec)
        paths.append((path, spec.values.shape[0]))
    paths.append((GOLDEN / "output_sphere_lzw.tif", 1))
    for path, bands in paths:
        for band in range(bands):
            f = N.GeoTiffFile(path)
            info, _, _ = f.info(band)
            got = f.read(band, 0, 0, info.width, info.height)
            want = np.zeros((info.height, info.width), np.float32)
            transform = np.zeros(6)
            no_data, has = C.c_double(), C.c_int()
            assert ref.ref_geotiff_read(str(path).encode(), band, want.ctypes.data,
                                        transform.ctypes.data, C.byref(no_data),
                                        C.byref(has)) == want.size
            if has.value:
                want[want == np.float32(no_data.value)] = np.nan
            assert same_heights(got, want), (path.name, band)
            assert np.array_equal(np.asarray(info.transform), transform), path.name
            # A window decodes only the chunks it meets, to the same samples.
            win = f.read(band, 3, 5, info.width - 7, info.height - 9)
            assert same_heights(win, got[5:info.height - 4, 3:info.width - 4])
