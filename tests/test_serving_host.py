# SPDX-License-Identifier: BSD-3-Clause
"""The camera request stream (host only): deterministic, Poisson count, quadtree-aligned tiles."""
import numpy as np

from terraingen_amd.serving import CameraStream


def test_stream_is_deterministic_and_sorted():
    a = CameraStream(duration=1.0).requests()
    b = CameraStream(duration=1.0).requests()
    assert np.array_equal(a, b)
    assert np.all(np.diff(a["t"]) >= 0.0) and a["t"][0] >= 0.0 and a["t"][-1] <= 1.0
    # Poisson(2000): within 5 sigma.
    assert abs(a.size - 2000) < 5 * np.sqrt(2000)


def test_tiles_sit_on_their_level_grid():
    s = CameraStream(duration=0.5, seed=9)
    r = s.requests()
    side = s.finest_side * np.exp2(s.levels - r["level"])
    assert np.all((r["level"] >= 0) & (r["level"] <= s.levels))
    assert np.allclose(r["pixel"] * (r["width"] - 1), side)
    assert np.all(np.isin(r["width"], s.edges)) and np.all(r["width"] == r["height"])
    assert np.allclose(np.mod(r["origin_x"], side), 0.0)
    assert np.allclose(np.mod(r["origin_y"], side), 0.0)
    # Finer levels only near the camera: a level-10 tile lies within a few tile sides of the
    # track at its arrival time.
    fine = r["level"] == s.levels
    cam_x = s.path_length * r["t"] / s.duration
    near = np.hypot(r["origin_x"][fine] - cam_x[fine], r["origin_y"][fine])
    assert near.max() < 4 * s.finest_side + np.sqrt(3) * s.altitude
