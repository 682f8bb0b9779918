# SPDX-License-Identifier: BSD-3-Clause
"""Writes the boulder fixtures with the reference's own BoulderLayer (oracle/_ref, built from its
sources by oracle/Makefile). Run from the repository root: python tests/golden/make_boulder_fixtures.py

* boulders_<relief case>.npz: on a plane, the boulders the layer draws into a grid (its list, in
  the layer's order: centre, diameter, exposed height, facets, model flag) and the FP32 heights its
  CPU backend evaluates for the source plus the boulder layer;
* boulders_sphere_<case>.npz: the same on a sphere (Surface::Sphere, the Moon's radius) over a
  longitude-latitude raster of zeros;
* boulders_<placement case>.npz: the boulders a power-law boulder layer lists over a region for a
  grid spacing, on a plane and on the Moon (tests/test_boulder_placement.py compares this
  repository's lists with them where oracle/_ref is not built).
"""
import ctypes as C
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

from oracle import (RefBoulderCase, RefBoulderSphereCase, reference,  # noqa: E402
                    reference_boulders)
from terraingen_amd import native as N  # noqa: E402

OUT = ROOT / "tests" / "golden"
C2_FRACTAL = dict(seed=1, wavelength=512.0, amplitude=20.0, octaves=8, lacunarity=2.0, gain=0.5)
MOON = 1737400.0

RELIEF_CASES = {
    # The layer's default sizes (NASA SP-8023 intercrater blocks) on a flat source at 1.5 m.
    "default_flat": dict(width=257, height=257, pixel=0.25, center_x=10.0, center_y=-20.0,
                         elevation=1.5, sizes_kind=0, seed=3),
    # C3's sizes (power law N(>D) = 3.8e-3 D^-3, D in [0.2, 3] m) with ventifacts on half the
    # boulders, over the C2 fractal source, so each cell's base differs.
    "c3_ventifacts": dict(width=193, height=161, pixel=0.2, center_x=-35.0, center_y=12.0,
                          elevation=0.0, source_kind=1, fractal_count=1, sizes_kind=1,
                          sfd_coefficient=3.8e-3, sfd_exponent=3.0, min_d=0.2, max_d=3.0,
                          ventifact_share=0.5, seed=2),
    # C3's sizes at 0.75 m, where the fade from 2 to 4 samples weights most drawn boulders.
    "c3_fade": dict(width=129, height=129, pixel=0.75, center_x=200.0, center_y=300.0,
                    elevation=-4.0, sizes_kind=1, sfd_coefficient=3.8e-3, sfd_exponent=3.0,
                    min_d=0.2, max_d=3.0, seed=5),
}

# C3's sizes on the Moon: cells of 0.2 m (a step of 0.2 m of arc in degrees) at mid latitude, and
# of 0.35 m near the north pole with ventifacts on half of the boulders.
SPHERE_CASES = {
    "midlat": dict(side=160, lon=30.0, lat=-10.0, step=0.2 / (MOON * np.pi / 180.0), seed=2,
                   ventifact_share=0.0),
    "polar_fade": dict(side=192, lon=-120.0, lat=84.0, step=0.35 / (MOON * np.pi / 180.0),
                       seed=9, ventifact_share=0.5),
}

PLACEMENT_CASES = {
    "plane_c3_quarter_m": dict(region=(-30.0, -20.0, 30.0, 25.0), spacing=0.25, seed=2),
    "plane_all_sizes": dict(region=(100.0, 100.0, 118.0, 112.0), spacing=0.0, seed=7),
    "moon_small_box": dict(region=(30.0, -10.0, 30.002, -9.998), spacing=0.5, seed=3,
                           sphere_radius=MOON),
}


def relief_case(lib, spec: dict) -> dict:
    case = RefBoulderCase(threads=8, **spec)
    if spec.get("source_kind") == 1:
        case.fractals[0] = N.make_fractal(**C2_FRACTAL)
    handle = lib.ref_boulders_open(C.byref(case))
    assert handle, lib.ref_boulders_error()
    try:
        n = lib.ref_boulders_list(handle, None)
        assert n >= 0, lib.ref_boulders_error()
        boulders = np.zeros(max(n, 1), N.BOULDER_DTYPE)[:n]
        lib.ref_boulders_list(handle, boulders.ctypes.data)
        heights = np.zeros((case.height, case.width), np.float32)
        assert lib.ref_boulders_evaluate(handle, heights.ctypes.data) == 0, \
            lib.ref_boulders_error()
    finally:
        lib.ref_boulders_close(handle)
    return dict(case=np.frombuffer(bytes(case), np.uint8), boulders=boulders, heights=heights)


def sphere_case(lib, spec: dict) -> dict:
    case = RefBoulderSphereCase(radius=MOON, sfd_coefficient=3.8e-3, sfd_exponent=3.0, min_d=0.2,
                                max_d=3.0, threads=8, **spec)
    heights = np.zeros((case.side, case.side), np.float32)
    n = lib.ref_boulders_sphere(C.byref(case), heights.ctypes.data, None, 0)
    assert n >= 0, lib.ref_boulders_error()
    boulders = np.zeros(max(n, 1), N.BOULDER_DTYPE)
    assert lib.ref_boulders_sphere(C.byref(case), heights.ctypes.data, boulders.ctypes.data,
                                   n) == n, lib.ref_boulders_error()
    return dict(case=np.frombuffer(bytes(case), np.uint8), boulders=boulders[:n],
                heights=heights)


def placement_spec(case: dict):
    return N.BoulderLayerSpec(*case["region"], case.get("sphere_radius", 0.0), case["spacing"],
                              3.8e-3, 3.0, 0.2, 3.0, **N.BOULDER_POSE_DEFAULTS,
                              seed=case["seed"], threads=1)


def main() -> None:
    lib = reference()
    if lib is None:
        raise SystemExit("oracle/_ref is not built (make -C oracle ref)")
    which = set(sys.argv[1:]) or {"relief", "sphere", "placement"}
    if "relief" in which:
        for name, spec in RELIEF_CASES.items():
            data = relief_case(lib, spec)
            np.savez_compressed(OUT / f"boulders_{name}.npz", **data)
            print(name, data["boulders"].size)
    if "sphere" in which:
        for name, spec in SPHERE_CASES.items():
            data = sphere_case(lib, spec)
            np.savez_compressed(OUT / f"boulders_sphere_{name}.npz", **data)
            b = data["boulders"]
            print(f"sphere_{name}: {b.size} boulders (facets "
                  f"{np.bincount(b['facet_count'], minlength=3)}), heights up to "
                  f"{data['heights'].max():.3f} m")
    if "placement" in which:
        for name, case in PLACEMENT_CASES.items():
            boulders = reference_boulders(lib, placement_spec(case))
            np.savez_compressed(OUT / f"boulders_{name}.npz", boulders=boulders)
            print(name, boulders.size)


if __name__ == "__main__":
    main()
