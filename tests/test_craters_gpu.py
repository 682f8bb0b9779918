# SPDX-License-Identifier: BSD-3-Clause
"""The crater kernel on the MI355X against the reference's CPU heights and the oracles, on the
reference's own prepared records (tests/golden/craters_*.npz)."""
import numpy as np
import pytest

from terraingen_amd import native as N

from crater_cases import NAMES, binding, load, oracle_heights, oracle_heights_f64

pytestmark = pytest.mark.gpu

# The north-star bound: max |dh| <= 1e-3 m against the FP64 oracle. The reference itself allows
# 2e-3 m between its GPU and CPU crater heights (crater_kernel_test.cpp:212-272).
TOL_FP64 = 1e-3


@pytest.fixture(scope="module", params=NAMES)
def case(request):
    return load(request.param)


def run(hip, case, heights=None, rows=None, first_row=0):
    h, w = case.shape
    out = np.zeros((h, w), np.float32) if heights is None else heights
    params = N.KernelParams(w, out.shape[0], N.NODATA_NONE, 0.0, first_row)
    ys = case.ys if rows is None else np.ascontiguousarray(case.ys[rows])
    hip.run(N.TG_KERNEL_CRATERS, case.kernel, out, case.xs, ys, params, binding(case))
    return out


def test_crater_kernel_matches_reference_and_oracles(hip, orc, case):
    got = run(hip, case)
    d_ref = np.abs(got - case.heights)
    f64 = oracle_heights_f64(orc, case)
    d64 = np.abs(got - f64)
    d32 = np.abs(got - oracle_heights(orc, case))
    print(f"[{case.name}] GPU vs reference CPU: max {d_ref.max():.3g} m; vs FP32 oracle max "
          f"{d32.max():.3g}; vs FP64 oracle max {d64.max():.3g} rms "
          f"{np.sqrt((d64 ** 2).mean()):.3g} m", flush=True)
    assert d64.max() <= TOL_FP64
    # The reference's FP32 arithmetic, apart from FP32 rounding of the fast outline (1.6e-6 m
    # measured): far inside the bound.
    assert d_ref.max() <= 2e-5


def test_strips_join(hip, case):
    whole = run(hip, case)
    h, w = case.shape
    joined = np.zeros((h, w), np.float32)
    for r0 in range(0, h, 45):
        r1 = min(h, r0 + 45)
        strip = np.zeros((r1 - r0, w), np.float32)
        run(hip, case, strip, slice(r0, r1), r0)
        joined[r0:r1] = strip
    assert np.array_equal(whole.view(np.uint32), joined.view(np.uint32))


def test_no_data_cells_untouched(hip, case):
    h, w = case.shape
    base = np.zeros((h, w), np.float32)
    base[::7, ::5] = -32768.0
    params = N.KernelParams(w, h, N.NODATA_VALUE, -32768.0, 0)
    hip.run(N.TG_KERNEL_CRATERS, case.kernel, base, case.xs, case.ys, params, binding(case))
    assert np.all(base[::7, ::5] == -32768.0)
