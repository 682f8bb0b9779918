# SPDX-License-Identifier: BSD-3-Clause
"""Writes tests/golden/boulders_*.npz from the reference's own CPU pipeline (oracle/_ref, compiled
from its sources in this container): for each case, the boulders BoulderLayer draws into the grid
(its list, in the layer's order: centre, diameter, exposed height, facets, model flag) and the
FP32 heights its CPU backend evaluates for the source plus the boulder layer.
Run from the repository root: python tests/golden/make_boulder_fixtures.py
"""
import ctypes as C
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

from oracle import RefBoulderCase, reference  # noqa: E402
from terraingen_amd import native as N  # noqa: E402

C2_FRACTAL = dict(seed=1, wavelength=512.0, amplitude=20.0, octaves=8, lacunarity=2.0, gain=0.5)

CASES = {
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


def run_case(name: str, spec: dict) -> dict:
    ref = reference()
    assert ref is not None, "oracle/_ref is not built"
    case = RefBoulderCase(threads=8, **spec)
    if spec.get("source_kind") == 1:
        case.fractals[0] = N.make_fractal(**C2_FRACTAL)
    handle = ref.ref_boulders_open(C.byref(case))
    assert handle, ref.ref_boulders_error()
    try:
        n = ref.ref_boulders_list(handle, None)
        assert n >= 0, ref.ref_boulders_error()
        boulders = np.zeros(max(n, 1), N.BOULDER_DTYPE)[:n]
        ref.ref_boulders_list(handle, boulders.ctypes.data)
        heights = np.zeros((case.height, case.width), np.float32)
        assert ref.ref_boulders_evaluate(handle, heights.ctypes.data) == 0, \
            ref.ref_boulders_error()
    finally:
        ref.ref_boulders_close(handle)
    return dict(case=np.frombuffer(bytes(case), np.uint8), boulders=boulders, heights=heights)


def main() -> None:
    out = ROOT / "tests" / "golden"
    for name, spec in CASES.items():
        data = run_case(name, spec)
        np.savez_compressed(out / f"boulders_{name}.npz", **data)
        b = data["boulders"]
        print(f"{name}: {b.size} boulders (facets {np.bincount(b['facet_count'], minlength=3)}), "
              f"heights {data['heights'].min():.3f}..{data['heights'].max():.3f}")


if __name__ == "__main__":
    main()
