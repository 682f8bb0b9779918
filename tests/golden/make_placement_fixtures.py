# SPDX-License-Identifier: BSD-3-Clause
"""Writes tests/golden/placement_sphere_*.npz from the reference's own crater generation on a
sphere (oracle/_ref, compiled from its sources in this container): GenerateCraters over a
longitude-latitude region with CubeSphereCells, a power-law density and uniform ages. Each
fixture holds the case and the list (x = longitude, y = latitude, diameter, age, ID) in the
reference's order.
Run from the repository root: python tests/golden/make_placement_fixtures.py
"""
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

from oracle import reference, reference_sphere_craters  # noqa: E402

MOON = 1737400.0
BASE = dict(radius=MOON, sfd_coefficient=0.05, sfd_exponent=2.0, min_resolved=0.0, lifetime=1.0,
            threads=1)

CASES = {
    # The dense field's population (slope -2.2, D from 2 m here) over a box on the near side.
    "box_nearside": dict(BASE, min_lon=20.0, min_lat=5.0, max_lon=20.05, max_lat=5.03,
                         sfd_coefficient=0.0156, sfd_exponent=2.2, min_d=2.0, max_d=300.0,
                         margin_radii=3.0, seed=5),
    # A box across the antimeridian.
    "antimeridian": dict(BASE, min_lon=179.95, min_lat=-10.0, max_lon=180.1, max_lat=-9.9,
                         min_d=20.0, max_d=200.0, margin_radii=3.0, seed=7),
    # A cap around the north pole of a 20 km body, whose margins take every longitude.
    "polar_small_body": dict(BASE, radius=20000.0, min_lon=-30.0, min_lat=85.0, max_lon=30.0,
                             max_lat=90.0, min_d=20.0, max_d=2000.0, margin_radii=2.0, seed=3),
    # The whole Moon, 14-20 km craters (cells of 262 and 524 km, many across face edges).
    "global_km": dict(BASE, min_lon=-180.0, min_lat=-90.0, max_lon=180.0, max_lat=90.0,
                      min_d=14000.0, max_d=20000.0, margin_radii=0.0, seed=4),
}


def main() -> None:
    lib = reference()
    if lib is None:
        raise SystemExit("oracle/_ref is not built (make -C oracle ref)")
    for name, spec in CASES.items():
        craters = reference_sphere_craters(lib, **spec)
        out = ROOT / "tests" / "golden" / f"placement_sphere_{name}.npz"
        case = np.array([tuple(spec[k] for k in sorted(spec))],
                        dtype=[(k, np.float64) for k in sorted(spec)])
        np.savez_compressed(out, case=case, x=craters["x"], y=craters["y"],
                            d=craters["diameter"], age=craters["age_gyr"], id=craters["id"])
        print(f"{out.name}: {craters.size} craters")


if __name__ == "__main__":
    main()
