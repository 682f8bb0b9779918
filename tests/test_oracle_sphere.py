# SPDX-License-Identifier: BSD-3-Clause
"""The crater oracle on a body pinned: the reference's CPU heights of crater layers on a sphere
(tests/golden/craters_sphere_*.npz) reproduced by the FP32 oracle over the reference's own frames
and split positions; those frames and positions checked against spherical geometry in FP64; and
the FP64 oracle, fed only longitudes and latitudes, within the height budget."""
import ctypes as C

import numpy as np
import pytest

from crater_cases import (SPHERE_NAMES, load_sphere, sphere_frames64, sphere_oracle_heights,
                          sphere_oracle_heights_f64, sphere_position)


@pytest.fixture(scope="module", params=SPHERE_NAMES)
def case(request):
    return load_sphere(request.param)


def test_fixture_shapes_craters(case):
    assert case.records.size > 5
    assert np.count_nonzero(case.heights) > case.heights.size // 4
    assert case.points.shape == (case.side, case.side)


def test_split_points_are_cell_centres_on_the_sphere(case):
    """MakeSplitPoint on a sphere: Surface::Position of the cell centre, split into float pairs."""
    lon, lat = case.cell_lonlat()
    want = sphere_position(lon, lat, case.radius)
    p = case.points
    got = np.stack([p["x"].astype(np.float64) + p["x_low"],
                    p["y"].astype(np.float64) + p["y_low"],
                    p["z"].astype(np.float64) + p["z_low"]], axis=-1)
    assert np.abs(got - want).max() < 1e-6


def test_frames_are_local_east_north(case):
    """MakeCraterFrame: the centre on the sphere and its local east and north unit vectors."""
    f = case.frames
    want = sphere_frames64(case)
    centre = np.stack([f["x"].astype(np.float64) + f["x_low"],
                       f["y"].astype(np.float64) + f["y_low"],
                       f["z"].astype(np.float64) + f["z_low"]], axis=-1)
    assert np.allclose(np.linalg.norm(centre, axis=1), case.radius, rtol=0, atol=1e-6)
    east = np.stack([f["east_x"], f["east_y"], f["east_z"]], axis=-1).astype(np.float64)
    north = np.stack([f["north_x"], f["north_y"], f["north_z"]], axis=-1).astype(np.float64)
    assert np.abs(east - want[:, 3:6]).max() < 1e-7
    assert np.abs(north - want[:, 6:9]).max() < 1e-7
    assert np.all(f["sphere_radius"] == np.float32(case.radius))


def test_fp32_oracle_reproduces_reference_heights(orc, case):
    """The same FP32 operations in the same order as ShapeCratersOnBody (craters.h:1359-1370)."""
    got = sphere_oracle_heights(orc, case)
    d = np.abs(got - case.heights)
    print(f"[sphere {case.name}] FP32 oracle vs reference: max {d.max():.3g} m", flush=True)
    assert d.max() <= 1e-5


def test_fp64_oracle_within_budget(orc, case):
    got = sphere_oracle_heights_f64(orc, case)
    d = np.abs(got - case.heights)
    print(f"[sphere {case.name}] FP64 oracle vs reference FP32: max {d.max():.3g} m, "
          f"rms {np.sqrt((d ** 2).mean()):.3g} m", flush=True)
    assert d.max() <= 1e-3


def test_body_crater_is_the_plane_crater_at_geodesic_polar_offsets(orc, case):
    """A crater on the sphere shapes each cell as the same crater on a plane does at the cell's
    geodesic polar offset: distance the great-circle arc, direction the bearing in
    the crater's tangent plane (the reference's SphereCraterLayer test,
    tests/unit/layers/craters/sphere_test.cpp:126-146, states this through the layer API). The
    FP64 body oracle, one crater at a time, against the FP64 plane oracle."""
    lon, lat = case.cell_lonlat()
    pos = sphere_position(lon, lat, case.radius).reshape(-1, 3)
    frames = sphere_frames64(case)
    k = case.kernel
    across = (case.side + 15) // 16
    tiles = np.asarray([across, 16, 1] + [0, 1] * (across * across), np.uint32)
    rng = np.random.default_rng(3)
    for n in rng.choice(case.records.size, size=min(case.records.size, 3), replace=False):
        c, east, north = frames[n, :3], frames[n, 3:6], frames[n, 6:9]
        # The angle by atan2 of |c x p| and c . p, which keeps its precision at small angles.
        arc = case.radius * np.arctan2(np.linalg.norm(np.cross(c, pos), axis=1), pos @ c)
        d = pos - c
        u, v = d @ east, d @ north
        across_len = np.hypot(u, v)
        dx, dy = arc * u / across_len, arc * v / across_len
        entries = np.asarray([n], np.uint32)
        body = np.zeros(case.side * case.side, np.float64)
        orc.or_shape_grid_body_f64(C.byref(k), case.records.ctypes.data, case.data.ctypes.data,
                                   tiles.ctypes.data, entries.ctypes.data, frames.ctypes.data,
                                   np.ascontiguousarray(pos).ctypes.data, case.side, case.side,
                                   0, body.ctypes.data, 1)
        rec = case.records.ctypes.data + int(n) * 144
        # The cells the crater reaches, and a sample of the rest.
        idx = np.flatnonzero(body != 0.0)
        idx = np.concatenate([idx[:: max(1, idx.size // 400)],
                              rng.choice(body.size, 100, replace=False)])
        plane = np.array([orc.or_apply_crater_f64(C.byref(k), case.data.ctypes.data, rec, 0.0,
                                                  float(dx[i]), float(dy[i])) for i in idx])
        assert np.abs(plane - body[idx]).max() < 1e-5, n
