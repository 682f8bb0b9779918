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
    spec = N.Upsample(0.0, 10.0, 1.0, 0, 0, 10, 10, 0, 0, 0.0, 10.0, 0.25, 40, 40,
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
