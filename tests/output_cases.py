# SPDX-License-Identifier: BSD-3-Clause
"""Inputs of the output-file tests and fixtures, seeded so that every run writes the same data."""
import numpy as np

from terraingen_amd import native as N

MOON = 1737400.0


def fixture_raster():
    """A 33 x 65 heightmap on the Moon (longitude-latitude cells of about 3 m), some cells
    no-data, with the sphere's projection keys."""
    rng = np.random.default_rng(5)
    values = (rng.normal(size=(33, 65)) * 40.0).astype(np.float32)
    values[3:6, 10:14] = -32768.0
    step = 3.0 / (MOON * np.pi / 180.0)
    return values, N.north_up_transform(12.25, -3.5, step), -32768.0, N.sphere_geokeys(MOON)


def fixture_craters():
    """Twenty craters with flags and attributes, values spanning many magnitudes."""
    rng = np.random.default_rng(8)
    n = 20
    craters = np.zeros(n, N.CRATER_DTYPE)
    craters["x"] = rng.uniform(-180.0, 180.0, n)
    craters["y"] = rng.uniform(-90.0, 90.0, n) * 10.0 ** rng.integers(-6, 2, n)
    craters["diameter"] = 10.0 ** rng.uniform(-1.0, 4.5, n)
    craters["age_gyr"] = rng.uniform(0.0, 3.9, n)
    craters["age_gyr"][0] = 0.0
    craters["id"] = rng.integers(0, 2 ** 63, n, dtype=np.uint64) * 2 + 1
    manual = (rng.random(n) < 0.3).astype(np.uint8)
    replace = (rng.random(n) < 0.2).astype(np.uint8)
    attributes = np.zeros(n, N.CRATER_ATTRIBUTES_DTYPE)
    attributes["floor_depth"] = craters["diameter"] * rng.uniform(0.05, 0.2, n)
    attributes["rim_height"] = craters["diameter"] * rng.uniform(0.01, 0.05, n)
    attributes["depth_ratio"] = rng.uniform(0.02, 0.2, n)
    attributes["freshness"] = rng.uniform(0.0, 1.0, n)
    attributes["complex_"] = (craters["diameter"] > 1000.0).astype(np.uint8)
    return craters, manual, replace, attributes
