# SPDX-License-Identifier: BSD-3-Clause
"""Writes boulder-list fixtures with the reference's own BoulderLayer (oracle/_ref, built from its
sources by oracle/Makefile): the boulders a power-law boulder layer lists over a region for a grid
spacing, on a plane and on the Moon. tests/test_boulder_placement.py compares this repository's
lists with them where oracle/_ref is not built. Run from the repository root:
python tests/golden/make_boulder_fixtures.py"""
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

from oracle import reference, reference_boulders  # noqa: E402
from terraingen_amd import native as N  # noqa: E402

# C3's boulder sizes: N(>D) = 3.8e-3 D^-3 per m^2, D 0.2-3 m.
CASES = {
    "plane_c3_quarter_m": dict(region=(-30.0, -20.0, 30.0, 25.0), spacing=0.25, seed=2),
    "plane_all_sizes": dict(region=(100.0, 100.0, 118.0, 112.0), spacing=0.0, seed=7),
    "moon_small_box": dict(region=(30.0, -10.0, 30.002, -9.998), spacing=0.5, seed=3,
                           sphere_radius=1737400.0),
}


def spec(case: dict):
    return N.BoulderLayerSpec(*case["region"], case.get("sphere_radius", 0.0), case["spacing"],
                              3.8e-3, 3.0, 0.2, 3.0, **N.BOULDER_POSE_DEFAULTS,
                              seed=case["seed"], threads=1)


def main() -> None:
    lib = reference()
    if lib is None:
        raise SystemExit("oracle/_ref is not built (make -C oracle ref)")
    for name, case in CASES.items():
        boulders = reference_boulders(lib, spec(case))
        np.savez_compressed(ROOT / "tests" / "golden" / f"boulders_{name}.npz", boulders=boulders)
        print(name, boulders.size)


if __name__ == "__main__":
    main()
