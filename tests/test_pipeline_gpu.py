# SPDX-License-Identifier: BSD-3-Clause
"""A whole terrain on the hip backend against the reference's CPU evaluation, its craters placed,
prepared and binned by this repository's host code: the quick-start terrain (flat source and the
default crater models, with diffused profiles), and synthetic fBm plus power-law craters."""
import numpy as np
import pytest

from terraingen_amd import native as N
from terraingen_amd.pipeline import CraterLayerSpec, PlaneTerrain

from crater_cases import GOLDEN

pytestmark = pytest.mark.gpu

C2_FRACTAL = dict(seed=1, wavelength=512.0, amplitude=20.0, octaves=8, lacunarity=2.0, gain=0.5)


def test_c2_recipe_matches_reference_terrain(hip):
    z = np.load(GOLDEN / "craters_c2_synthetic.npz")
    grid = N.north_up_grid(257, 257, 0.5, (300.0, -150.0))
    terrain = PlaneTerrain(hip, grid, [C2_FRACTAL], 0.0, CraterLayerSpec(seed=1))
    terrain.prepare()
    try:
        assert terrain.crater_list.size == z["x"].size
        assert np.array_equal(terrain.crater_list["id"], z["id"])
        got = terrain.evaluate()
    finally:
        terrain.close()
    d = np.abs(got - z["heights"])
    print(f"[c2_synthetic] GPU terrain vs reference CPU terrain: max {d.max():.3g} m, "
          f"rms {np.sqrt((d ** 2).mean()):.3g} m over {got.size} cells", flush=True)
    # FP32 rounding of the fast outline and of the noise's sum order only (4.5e-6 m measured).
    assert d.max() <= 5e-5


def test_c1_quick_start_matches_reference_terrain(hip):
    """FlatSource 257 x 257 at 1 m, height 0, one CraterLayer with the reference's default models
    and apply options, seed 0: the same crater list, and the reference's CPU heights bit for bit
    (the diffusion tables, states and records are byte-equal to the reference's)."""
    z = np.load(GOLDEN / "craters_c1_default.npz")
    grid = N.north_up_grid(257, 257, 1.0, (0.0, 0.0))
    spec = CraterLayerSpec(model=N.MODELS_DEFAULT, seed=0, min_diameter=2.5, max_diameter=250.0,
                           irregular=1, peaks=1, terraces=1)
    terrain = PlaneTerrain(hip, grid, [], 0.0, spec)
    terrain.prepare()
    try:
        for col, key in [("x", "x"), ("y", "y"), ("diameter", "d"), ("age_gyr", "age"),
                         ("id", "id")]:
            assert np.array_equal(terrain.crater_list[col], z[key]), col
        assert terrain.records.view(np.uint8).tobytes() == z["records"].tobytes()
        got = terrain.evaluate()
    finally:
        terrain.close()
    d = np.abs(got - z["heights"])
    print(f"[c1_default] GPU terrain vs reference CPU terrain: max {d.max():.3g} m over "
          f"{got.size} cells, {terrain.crater_count} craters", flush=True)
    # FP32 rounding of the fast outline only (9.5e-7 m measured).
    assert d.max() <= 1e-5


def test_c1_quick_start_with_the_readme_seed(hip, ref):
    """The quick-start terrain with the seed the reference's README and options default to
    (CraterLayerOptions::seed = 1), against the reference's own CPU pipeline run here (oracle/_ref):
    the same crater list and records byte for byte, the heights within the same bound."""
    import sys

    sys.path.insert(0, str(GOLDEN))
    from make_crater_fixtures import CASES, run_case

    z = run_case("c1_seed1", dict(CASES["c1_default"], seed=1))
    grid = N.north_up_grid(257, 257, 1.0, (0.0, 0.0))
    spec = CraterLayerSpec(model=N.MODELS_DEFAULT, seed=1, min_diameter=2.5, max_diameter=250.0,
                           irregular=1, peaks=1, terraces=1)
    terrain = PlaneTerrain(hip, grid, [], 0.0, spec)
    terrain.prepare()
    try:
        for col, key in [("x", "x"), ("y", "y"), ("diameter", "d"), ("age_gyr", "age"),
                         ("id", "id")]:
            assert np.array_equal(terrain.crater_list[col], z[key]), col
        assert terrain.records.view(np.uint8).tobytes() == z["records"].tobytes()
        got = terrain.evaluate()
    finally:
        terrain.close()
    d = np.abs(got - z["heights"])
    print(f"[c1 seed 1] GPU terrain vs reference CPU terrain: max {d.max():.3g} m over "
          f"{got.size} cells, {terrain.crater_count} craters", flush=True)
    assert terrain.crater_count > 100
    assert d.max() <= 1e-5
