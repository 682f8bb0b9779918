# SPDX-License-Identifier: BSD-3-Clause
"""The boulder oracle against the reference: its FP32-stored form reproduces the reference's CPU
heights bit for bit on every fixture, its FP64 form stays within FP32 rounding of them, and the
dome obeys its closed forms."""
import ctypes as C
import math

import numpy as np
import pytest

from terraingen_amd import native as N

from boulder_cases import (NAMES, SPHERE_NAMES, load, load_sphere, oracle_heights,
                          oracle_heights_f64, sphere_oracle_f64)


@pytest.fixture(scope="module", params=NAMES)
def case(request, orc):
    return load(request.param, orc)


def test_oracle_reproduces_the_reference_heights(orc, case):
    got = oracle_heights(orc, case)
    assert np.array_equal(got.view(np.uint32), case.heights.view(np.uint32))
    assert (case.heights > case.base).any()  # the boulders drew something


def test_fp64_oracle_tracks_the_reference(orc, case):
    d = np.abs(oracle_heights_f64(orc, case) - case.heights)
    assert d.max() <= 1e-5


def test_prepared_records_drop_models_and_unresolved_boulders(case):
    b = case.boulders
    spacing = case.spec.pixel
    drawn = (b["model"] == 0) & (b["diameter"] / spacing > 2.0)
    assert case.records.size == int(drawn.sum())
    r = case.records
    assert np.allclose(r["radius"], 0.5 * b["diameter"][drawn])
    assert np.allclose(r["reach"], 0.5 * (b["diameter"][drawn] + spacing))
    # The split centre holds the FP64 centre to the FP32 rounding of the remainder.
    split = r["x"].astype(np.float64) + r["x_low"]
    assert (np.abs(split - b["x"][drawn]) <= 1e-12 * np.maximum(1.0, np.abs(b["x"][drawn]))).all()
    assert case.kernel.width == np.float32(spacing) and case.kernel.count == r.size


def dome_boulder(diameter=2.0, exposed=0.8, facets=()):
    b = np.zeros(1, N.BOULDER_DTYPE)
    b["diameter"], b["exposed"] = diameter, exposed
    for k, f in enumerate(facets):
        b["facets"][0, 3 * k:3 * k + 3] = f
    b["facet_count"] = len(facets)
    return b


def test_dome_closed_forms(orc):
    b = dome_boulder()
    relief = lambda dx, dy, w: orc.or_boulder_relief_f64(b.ctypes.data, dx, dy, w)  # noqa: E731
    # A point sample of the half-ellipsoid: exposed x sqrt(1 - (d / r)^2).
    assert relief(0.0, 0.0, 0.0) == pytest.approx(0.8)
    assert relief(0.6, 0.0, 0.0) == pytest.approx(0.8 * math.sqrt(1 - 0.36))
    assert relief(1.0, 0.0, 0.0) == 0.0
    # The band average of a band narrower than the dome's flat top tends to the point value.
    assert relief(0.3, 0.4, 1e-6) == pytest.approx(relief(0.3, 0.4, 0.0), rel=1e-6)
    # Beyond the radius plus half the band there is nothing.
    assert relief(1.0 + 0.25 + 1e-9, 0.0, 0.5) == 0.0
    # A facet facing east that meets the ground at the centre leaves nothing east of the centre
    # and cuts nothing far west of it at its slope.
    cut = dome_boulder(facets=[(0.0, math.radians(55.0), 0.0)])
    relief_cut = lambda dx, dy: orc.or_boulder_relief_f64(cut.ctypes.data, dx, dy, 0.0)  # noqa
    assert relief_cut(0.2, 0.0) == 0.0
    assert relief_cut(-0.9, 0.0) == pytest.approx(min(0.8 * math.sqrt(1 - 0.81),
                                                      0.9 * math.tan(math.radians(55.0))))


@pytest.fixture(scope="module", params=SPHERE_NAMES)
def sphere_case(request):
    return load_sphere(request.param)


def test_body_oracle_tracks_the_reference_on_a_sphere(orc, sphere_case):
    """The FP64 restatement on a sphere (SphereOffset, the dome at the raster's spacing) against
    the reference's BoulderLayer on Surface::Sphere, to the FP32 rounding of its stored heights."""
    got = sphere_oracle_f64(orc, sphere_case)
    ref = sphere_case.heights
    assert (ref > 0.0).sum() > 50  # the layer drew boulders into the raster
    assert np.abs(got - ref).max() <= 1e-7


def test_sphere_offset_is_the_azimuthal_equidistant_offset(orc):
    """SphereOffset against the great-circle distance (haversine) and the initial bearing
    (atan2(sin dlon cos lat2, cos lat1 sin lat2 - sin lat1 cos lat2 cos dlon)) of textbook
    spherical trigonometry: the offset's length is the arc and its direction the bearing, east
    of north."""
    rng = np.random.default_rng(5)
    radius = 1737400.0
    out = np.zeros(2)
    for _ in range(200):
        lon1, lat1 = rng.uniform(-180, 180), rng.uniform(-89, 89)
        scale = 10.0 ** rng.uniform(-6, 0)
        lon2 = lon1 + rng.uniform(-1, 1) * scale
        lat2 = float(np.clip(lat1 + rng.uniform(-1, 1) * scale, -89.9, 89.9))
        orc.or_sphere_offset(radius, lon1, lat1, lon2, lat2, out.ctypes.data)
        p1, p2, dl = map(math.radians, (lat1, lat2, lon2 - lon1))
        hav = math.sin((p2 - p1) / 2) ** 2 + math.cos(p1) * math.cos(p2) * math.sin(dl / 2) ** 2
        arc = 2 * radius * math.asin(math.sqrt(hav))
        bearing = math.atan2(math.sin(dl) * math.cos(p2),
                             math.cos(p1) * math.sin(p2) - math.sin(p1) * math.cos(p2) * math.cos(dl))
        assert math.hypot(*out) == pytest.approx(arc, rel=1e-9, abs=1e-9)
        if arc > 1e-3:
            assert out[0] == pytest.approx(arc * math.sin(bearing), rel=1e-7, abs=1e-7 * arc)
            assert out[1] == pytest.approx(arc * math.cos(bearing), rel=1e-7, abs=1e-7 * arc)
