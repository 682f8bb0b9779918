# SPDX-License-Identifier: BSD-3-Clause
"""The crater oracle pinned: bit for bit against the reference's own ApplyCraterKernel
(oracle/_ref) on the reference's prepared records, end to end against the reference's CPU heights
(tests/golden/craters_*.npz), and against closed forms of the profiles."""
import ctypes as C

import numpy as np
import pytest

from terraingen_amd import native as N

from crater_cases import NAMES, load, oracle_heights, oracle_heights_f64


@pytest.fixture(scope="module", params=NAMES)
def case(request):
    return load(request.param)


def test_fixture_layouts(case):
    assert case.records.itemsize == 144
    assert C.sizeof(case.kernel) == 188
    assert case.tiles[0] == (case.grid.width + 15) // 16 and case.tiles[1] == 16


def test_oracle_pairs_match_reference_bitwise(orc, ref, case):
    """Every record, offsets across its reach: identical bits (same FP32 operation order, same
    libm, contraction off)."""
    rng = np.random.default_rng(11)
    k = case.kernel
    data = case.data.ctypes.data
    base = case.records.ctypes.data
    n = case.records.size
    picks = rng.choice(n, size=min(n, 150), replace=False)
    for i in picks:
        rec = base + int(i) * 144
        r = 0.5 * float(case.records["diameter"][i]) * float(k.reach_radii) * 1.3
        for dx, dy in rng.uniform(-r, r, (12, 2)):
            h = float(np.float32(rng.normal()))
            a = orc.or_apply_crater(C.byref(k), data, rec, h, float(np.float32(dx)),
                                    float(np.float32(dy)))
            b = ref.ref_apply_crater(C.byref(k), data, rec, h, float(np.float32(dx)),
                                     float(np.float32(dy)))
            assert np.float32(a).tobytes() == np.float32(b).tobytes(), (i, dx, dy, a, b)


def test_oracle_grid_matches_reference_heights(orc, case):
    """This repository's binning plus the FP32 oracle reproduce the reference's CPU heights."""
    got = oracle_heights(orc, case)
    d = np.abs(got - case.heights)
    print(f"[{case.name}] oracle FP32 vs reference: max {d.max():.3g} m", flush=True)
    assert d.max() <= 1e-5


def test_fp64_oracle_within_budget(orc, case):
    got = oracle_heights_f64(orc, case)
    d = np.abs(got - case.heights)
    print(f"[{case.name}] FP64 oracle vs reference FP32: max {d.max():.3g} m, "
          f"rms {np.sqrt((d ** 2).mean()):.3g} m", flush=True)
    assert d.max() <= 1e-3


def test_far_cells_unchanged(orc, case):
    """Beyond its reach a crater leaves the height as it was (craters.h:1191-1193)."""
    k = case.kernel
    for i in range(min(50, case.records.size)):
        rec = case.records.ctypes.data + i * 144
        far = 0.5 * float(case.records["diameter"][i]) * float(k.reach_radii) * 4.0
        assert orc.or_apply_crater(C.byref(k), case.data.ctypes.data, rec, 1.25, far, 0.0) == 1.25


def test_mpd_center_is_rim_minus_depth(orc):
    """A round Martin-Parkes-Dunstan crater at its center: resolution x (rim - depth) x D
    (the interior law at x = -1, craters.h:398-404)."""
    k = N.CraterShape()
    k.profile.simple = 1
    k.reach_radii = 2.0
    k.profile_reach_radii = 2.0
    rec = N.CraterRecord()
    rec.diameter, rec.resolution = 40.0, 0.75
    rec.fresh_depth, rec.fresh_rim = 0.2, 0.04
    rec.current_depth, rec.current_rim = 0.12, 0.025
    rec.ellipticity = 1.0
    data = np.zeros(1, np.float32)
    h = orc.or_apply_crater(C.byref(k), data.ctypes.data, C.byref(rec), 0.0, 0.0, 0.0)
    assert h == pytest.approx(0.75 * (0.025 - 0.12) * 40.0, rel=1e-6)
    # At the rim crest (x = 0) the height is the current rim.
    h = orc.or_apply_crater(C.byref(k), data.ctypes.data, C.byref(rec), 0.0, 20.0, 0.0)
    assert h == pytest.approx(0.75 * 0.025 * 40.0, rel=1e-6)
