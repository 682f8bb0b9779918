# SPDX-License-Identifier: BSD-3-Clause
"""Host placement and preparation against the reference: the same crater lists in the same order
(x, y, diameter, age, ID bit for bit), and the same kernel struct, records and data bytes."""
import ctypes as C
import math

import numpy as np
import pytest

from terraingen_amd import native as N

from crater_cases import GOLDEN

# (fixture, models, apply, layer seed, spacing, sfd (c, e), diameters, lifetime, grid)
CASES = {
    "c2_powerlaw": (N.CraterModels(1, 0, 15000.0, 1.0), N.Apply(1, 0, 0), 1, 0.5, (0.05, 2.0),
                    (1.0, 200.0), 1.0),
    "features_ft": (N.CraterModels(0, 0, 15000.0, 0.2), N.Apply(1, 0, 0, 0.08, 0.01, 0.005,
                                                                0.004, 0.3), 7, 0.5,
                    (0.05, 2.0), (2.0, 60.0), 0.2),
    "c3_complex": (N.CraterModels(1, 1, 50.0, 1.0), N.Apply(1, 1, 0), 2, 0.25, (0.0156, 2.2),
                   (0.5, 300.0), 1.0),
}


def fixture(name):
    return np.load(GOLDEN / f"craters_{name}.npz")


def grid_bounds(z):
    spec = z["case"].tobytes()
    w, h = np.frombuffer(spec[:8], np.uint32)
    pixel, cx, cy = np.frombuffer(spec[8:32], np.float64)
    return cx - 0.5 * w * pixel, cy - 0.5 * h * pixel, cx + 0.5 * w * pixel, cy + 0.5 * h * pixel


@pytest.mark.parametrize("name", list(CASES))
def test_generated_list_is_the_references(name):
    models, apply, seed, spacing, (c, e), (dmin, dmax), life = CASES[name]
    z = fixture(name)
    x0, y0, x1, y1 = grid_bounds(z)
    g = N.Generation(min_x=x0, min_y=y0, max_x=x1, max_y=y1, sfd_coefficient=c, sfd_exponent=e,
                     min_diameter=dmin, max_diameter=dmax,
                     margin_radii=N.margin_radii(models, apply), cell_size_factor=16.0,
                     min_resolved_diameter=2.0 * spacing, lifetime_gyr=life, seed=seed)
    got = N.generate_craters(g)
    assert got.size == z["x"].size
    for col, key in [("x", "x"), ("y", "y"), ("diameter", "d"), ("age_gyr", "age"), ("id", "id")]:
        assert np.array_equal(got[col], z[key]), col


@pytest.mark.parametrize("name", ["c2_powerlaw", "features_ft"])
def test_prepared_batch_is_the_references(name):
    models, apply, seed, spacing, *_ = CASES[name]
    z = fixture(name)
    cr = np.zeros(z["x"].size, N.CRATER_DTYPE)
    cr["x"], cr["y"], cr["diameter"], cr["age_gyr"], cr["id"] = z["x"], z["y"], z["d"], z["age"], z["id"]
    kernel, records, data, foot = N.prepare_craters(cr, models, apply, seed, spacing)
    assert bytes(kernel) == z["kernel"].tobytes()
    assert records.view(np.uint8).tobytes() == z["records"].tobytes()
    assert data.view(np.uint32).tobytes() == z["data"].view(np.uint32).tobytes()
    assert foot.shape == (records.size, 4) and np.all(foot[:, 2] > 0)
    assert np.array_equal(foot[:, 3].astype(np.float32), records["diameter"])


def test_prepared_peaks_are_the_references():
    """Complex craters' central peak massifs (records and data up to the first terraced crater,
    terraces are not restated here)."""
    models, apply, seed, spacing, *_ = CASES["c3_complex"]
    z = fixture("c3_complex")
    cr = np.zeros(z["x"].size, N.CRATER_DTYPE)
    cr["x"], cr["y"], cr["diameter"], cr["age_gyr"], cr["id"] = z["x"], z["y"], z["d"], z["age"], z["id"]
    kernel, records, data, _ = N.prepare_craters(cr, models, apply, seed, spacing)
    ref = z["records"].view(N.RECORD_DTYPE)
    first_terraced = int(np.argmax(ref["terrace_count"] > 0))
    assert ref["terrace_count"][first_terraced] > 0
    assert (ref["peak_count"][:first_terraced] > 0).any() or (records["peak_count"] > 0).any()
    n = first_terraced
    assert records[:n].tobytes() == ref[:n].tobytes()
    end = int(ref["harmonic_offset"][n])
    assert np.array_equal(data[:end].view(np.uint32), z["data"][:end].view(np.uint32))


@pytest.mark.parametrize("counter,key,want", [
    ((0, 0, 0, 0), (0, 0), (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)),
    ((0xFFFFFFFF,) * 4, (0xFFFFFFFF,) * 2, (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)),
    ((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0),
     (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1)),
])
def test_philox_known_answers(counter, key, want):
    """Random123's Philox4x32-10 known answers (tests/unit/random/philox_test.cpp:14-23)."""
    lib = N.host_library()
    c = (C.c_uint32 * 4)(*counter)
    k = (C.c_uint32 * 2)(*key)
    out = (C.c_uint32 * 4)()
    lib.tgx_philox4x32(c, k, out)
    assert tuple(out) == want
