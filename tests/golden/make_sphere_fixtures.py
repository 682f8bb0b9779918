# SPDX-License-Identifier: BSD-3-Clause
"""Writes tests/golden/craters_sphere_*.npz from the reference's own CPU pipeline (oracle/_ref,
compiled from its sources in this container): a CraterLayer on Surface::Sphere applied to a
longitude-latitude raster of zeros, as tests/unit/layers/craters/sphere_test.cpp sets one up.
Each fixture holds what the crater kernel reads on a body (the kernel struct, every prepared
crater's record, the data block, each crater's frame, each cell's split position) and the FP32
heights the reference's CPU backend shapes.
Run from the repository root: python tests/golden/make_sphere_fixtures.py
"""
import ctypes as C
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

from oracle import RefSphereCase, reference  # noqa: E402
from terraingen_amd import native as N  # noqa: E402

MOON = 1737400.0

CASES = {
    # sphere_test.cpp's models and sizes (power law 0.05 D^-2, D in [150, 1200] m, FT profile,
    # no degradation) at mid latitude, 0.0008 degrees (~24 m) a cell, with a 2 km crater added.
    "ft_midlat": dict(side=96, lon=30.02, lat=60.03, step=0.0008, radius=MOON, models=0,
                      sfd_coefficient=0.05, sfd_exponent=2.0, min_d=150.0, max_d=1200.0,
                      lifetime=1.0, seed=5, manual=[(30.06, 59.99, 2000.0, 1.0)]),
    # Martin-Parkes-Dunstan with freshness degradation and irregular outlines (half elongated)
    # across the antimeridian, as the reference's antimeridian test.
    "mpd_antimeridian": dict(side=100, lon=179.95, lat=45.05, step=0.001, radius=MOON, models=1,
                             sfd_coefficient=0.05, sfd_exponent=2.0, min_d=150.0, max_d=1200.0,
                             lifetime=1.0, irregular=1, elongated_share=0.5, seed=5,
                             manual=[(179.99, 45.01, 3000.0, 0.5)]),
    # A denser field south of the equator (D from 60 m), so craters overlap and their order
    # matters.
    "ft_dense_south": dict(side=128, lon=-70.04, lat=-19.95, step=0.0006, radius=MOON, models=0,
                           sfd_coefficient=0.05, sfd_exponent=2.0, min_d=60.0, max_d=1200.0,
                           lifetime=1.0, seed=9),
}


def run_case(spec: dict) -> dict:
    ref = reference()
    assert ref is not None, "oracle/_ref is not built"
    spec = dict(spec)
    manual = spec.pop("manual", [])
    case = RefSphereCase(threads=8, **spec)
    case.manual_count = len(manual)
    for k, row in enumerate(manual):
        for j, v in enumerate(row):
            case.manual[k][j] = v
    handle = ref.ref_sphere_open(C.byref(case))
    assert handle, ref.ref_sphere_error()
    try:
        data_count = C.c_long()
        n = ref.ref_sphere_sizes(handle, C.byref(data_count))
        kernel = np.zeros(C.sizeof(N.CraterShape), np.uint8)
        records = np.zeros(max(n, 1), N.RECORD_DTYPE)[:n]
        data = np.zeros(max(data_count.value, 1), np.float32)[:data_count.value]
        frames = np.zeros(max(n, 1), N.FRAME_DTYPE)[:n]
        points = np.zeros((case.side, case.side), N.POINT_DTYPE)
        heights = np.zeros((case.side, case.side), np.float32)
        ref.ref_sphere_copy(handle, kernel.ctypes.data, records.ctypes.data, data.ctypes.data,
                            frames.ctypes.data, points.ctypes.data, heights.ctypes.data)
    finally:
        ref.ref_sphere_close(handle)
    return dict(case=np.frombuffer(bytes(case), np.uint8), kernel=kernel, records=records,
                data=data, frames=frames, points=points, heights=heights)


def main() -> None:
    out = ROOT / "tests" / "golden"
    for name, spec in CASES.items():
        data = run_case(spec)
        np.savez_compressed(out / f"craters_sphere_{name}.npz", **data)
        h = data["heights"]
        print(f"{name}: {data['records'].size} craters, {np.count_nonzero(h)} cells shaped, "
              f"heights {h.min():.3f}..{h.max():.3f}")


if __name__ == "__main__":
    main()
