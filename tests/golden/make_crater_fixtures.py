# SPDX-License-Identifier: BSD-3-Clause
"""Writes tests/golden/craters_*.npz from the reference's own CPU pipeline (oracle/_ref, compiled
from its sources in this container): for each case, the crater list the layer applies (x, y,
diameter, age, id, in generation order), the kernel struct, records and data the reference
prepares for them (PrepareCraters + MakeCraterKernelBatch), and the FP32 heights its CPU backend
evaluates. Run from the repository root: python tests/golden/make_crater_fixtures.py
"""
import ctypes as C
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

from oracle import RefCase, reference  # noqa: E402
from terraingen_amd import native as N  # noqa: E402

CASES = {
    # The quick-start terrain: FlatSource 257 x 257 at 1 m, height 0, the default crater models
    # (Neukum production + Hartmann equilibrium, steady-state ages, diffused Martin-Parkes-Dunstan,
    # matched diffusion), default apply options, seed 0.
    "c1_default": dict(width=257, height=257, pixel=1.0, craters=1, models=0, irregular=1,
                       peaks=1, terraces=1, seed=0),
    # Power-law craters N(>D) = 0.05 D^-2, D in [1, 200] m, uniform ages over 1 Gyr, MPD profile
    # with freshness degradation, irregular outlines, seed 1, at 0.5 m.
    "c2_powerlaw": dict(width=257, height=257, pixel=0.5, craters=1, models=1,
                        sfd_coefficient=0.05, sfd_exponent=2.0, min_d=1.0, max_d=200.0,
                        lifetime=1.0, irregular=1, seed=1),
    # Dense complex craters: slope -2.2, D in [0.5, 300] m, complex from 50 m with peaks and
    # terraces from 150 m (full at 300 m), at 0.25 m, seed 2.
    "c3_complex": dict(width=257, height=257, pixel=0.25, center_x=40.0, center_y=-25.0,
                       craters=1, models=2, sfd_coefficient=0.0156, sfd_exponent=2.2, min_d=0.5,
                       max_d=300.0, lifetime=1.0, transition_diameter=50.0, irregular=1,
                       peaks=1, terraces=1, terrace_onset=150.0, terrace_full=300.0, seed=2),
    # Every optional kernel feature: rim smoothing, footprint and rim noise, wall streaks and
    # ejecta texture, on Fassett-Thomson craters.
    "features_ft": dict(width=161, height=161, pixel=0.5, craters=1, models=3,
                        sfd_coefficient=0.05, sfd_exponent=2.0, min_d=2.0, max_d=60.0,
                        lifetime=0.2, irregular=1, seed=7, rim_sigma=0.08,
                        noise_amplitude=0.01, noise_rim_amplitude=0.005, wall_streaks=0.004,
                        ejecta=0.3),
    # The C2 recipe at 257 x 257: synthetic source of 8-octave Perlin fBm (512 m, lacunarity 2,
    # gain 0.5, 20 m, seed 1) plus power-law craters, MPD with freshness degradation, seed 1.
    "c2_synthetic": dict(width=257, height=257, pixel=0.5, center_x=300.0, center_y=-150.0,
                         source_kind=1, fractal_count=1, craters=1, models=1,
                         sfd_coefficient=0.05, sfd_exponent=2.0, min_d=1.0, max_d=200.0,
                         lifetime=1.0, irregular=1, seed=1),
    # Wall terraces on simple craters: a dense power-law field (N(>D) = 0.5 D^-2, D in [4, 50] m)
    # terraced from 12 m (full at 40 m), at 0.25 m, with the Martin-Parkes-Dunstan profile and
    # freshness degradation, then with the Fassett-Thomson profile.
    "terraces_mpd": dict(width=257, height=257, pixel=0.25, center_x=-60.0, center_y=35.0,
                         craters=1, models=1, sfd_coefficient=0.5, sfd_exponent=2.0, min_d=4.0,
                         max_d=50.0, lifetime=1.0, irregular=1, terraces=1, terrace_onset=12.0,
                         terrace_full=40.0, seed=11),
    "terraces_ft": dict(width=257, height=257, pixel=0.25, center_x=75.0, center_y=10.0,
                        craters=1, models=3, sfd_coefficient=0.5, sfd_exponent=2.0, min_d=4.0,
                        max_d=50.0, lifetime=0.5, irregular=1, terraces=1, terrace_onset=12.0,
                        terrace_full=40.0, seed=12),
}

C2_FRACTAL = dict(seed=1, wavelength=512.0, amplitude=20.0, octaves=8, lacunarity=2.0, gain=0.5)


def run_case(name: str, spec: dict) -> dict:
    ref = reference()
    assert ref is not None, "oracle/_ref is not built"
    case = RefCase(threads=8, **spec)
    if spec.get("source_kind") == 1:
        case.fractals[0] = N.make_fractal(**C2_FRACTAL)
    h = ref.ref_case_open(C.byref(case))
    if not h:
        raise RuntimeError(ref.ref_last_error().decode())
    try:
        n = ref.ref_case_craters(h, None, None, None, None, None)
        cols = [np.zeros(n) for _ in range(4)] + [np.zeros(n, np.uint64)]
        ref.ref_case_craters(h, *[a.ctypes.data for a in cols])
        kernel = N.CraterShape()
        data_count = C.c_long()
        nrec = ref.ref_case_batch(h, C.byref(kernel), C.byref(data_count))
        assert nrec >= 0, ref.ref_last_error()
        records = np.zeros(nrec, N.RECORD_DTYPE)
        data = np.zeros(data_count.value, np.float32)
        ref.ref_case_batch_copy(h, records.ctypes.data, data.ctypes.data)
        heights = np.zeros((case.height, case.width), np.float32)
        assert ref.ref_case_evaluate(h, heights.ctypes.data) == 0, ref.ref_last_error()
    finally:
        ref.ref_case_close(h)
    spec_arr = np.frombuffer(bytes(case), np.uint8)
    return dict(x=cols[0], y=cols[1], d=cols[2], age=cols[3], id=cols[4],
                kernel=np.frombuffer(bytes(kernel), np.uint8), records=records.view(np.uint8),
                data=data, heights=heights, case=spec_arr)


def main() -> None:
    out = Path(__file__).resolve().parent
    # Names on the command line write only those fixtures.
    wanted = sys.argv[1:] or list(CASES)
    for name, spec in CASES.items():
        if name not in wanted:
            continue
        arrays = run_case(name, spec)
        np.savez_compressed(out / f"craters_{name}.npz", **arrays)
        print(f"{name}: {arrays['x'].size} craters, {arrays['records'].size // 144} records, "
              f"{arrays['data'].size} data floats, heights "
              f"[{arrays['heights'].min():.3f}, {arrays['heights'].max():.3f}]")


if __name__ == "__main__":
    main()
