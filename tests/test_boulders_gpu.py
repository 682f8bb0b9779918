# SPDX-License-Identifier: BSD-3-Clause
"""The boulder relief kernel on the MI355X against the reference's CPU heights and the FP64
oracle, on the reference's own boulder lists (tests/golden/boulders_*.npz)."""
import numpy as np
import pytest

from terraingen_amd import native as N

from boulder_cases import NAMES, binding, load, oracle_heights_f64

pytestmark = pytest.mark.gpu

TOL_FP64 = 1e-3  # the north-star bound against the FP64 oracle


@pytest.fixture(scope="module", params=NAMES)
def case(request, orc):
    return load(request.param, orc)


def run(hip, case, heights=None, rows=None, first_row=0):
    out = case.base.copy() if heights is None else heights
    params = N.KernelParams(case.grid.width, out.shape[0], N.NODATA_NONE, 0.0, first_row)
    ys = case.ys if rows is None else np.ascontiguousarray(case.ys[rows])
    hip.run(N.TG_KERNEL_BOULDERS, case.kernel, out, case.xs, ys, params, binding(case))
    return out


def test_boulder_kernel_matches_reference_and_oracle(hip, orc, case):
    got = run(hip, case)
    d_ref = np.abs(got - case.heights)
    d64 = np.abs(got.astype(np.float64) - oracle_heights_f64(orc, case))
    print(f"[{case.name}] {case.records.size} boulders: GPU vs reference CPU max "
          f"{d_ref.max():.3g} m; vs FP64 oracle max {d64.max():.3g} rms "
          f"{np.sqrt((d64 ** 2).mean()):.3g} m", flush=True)
    assert d64.max() <= TOL_FP64 and d_ref.max() <= TOL_FP64
    assert d_ref.max() <= 1e-4  # FP32 rounding of the dome, far inside the bound


def test_strips_join(hip, case):
    whole = run(hip, case)
    h = case.grid.height
    joined = np.empty_like(whole)
    for r0 in range(0, h, 37):
        r1 = min(h, r0 + 37)
        strip = case.base[r0:r1].copy()  # a view would let the run write into the fixture
        run(hip, case, strip, slice(r0, r1), r0)
        joined[r0:r1] = strip
    assert np.array_equal(whole.view(np.uint32), joined.view(np.uint32))


def test_no_data_cells_untouched(hip, case):
    base = case.base.copy()
    base[::5, ::3] = -32768.0
    out = base.copy()
    params = N.KernelParams(case.grid.width, case.grid.height, N.NODATA_VALUE, -32768.0, 0)
    hip.run(N.TG_KERNEL_BOULDERS, case.kernel, out, case.xs, case.ys, params, binding(case))
    assert (out[::5, ::3] == -32768.0).all()


def test_run_kernel_runs_the_boulder_program(case):
    values = case.base.copy()
    assert N.run_kernel(6, N.TG_KERNEL_BOULDERS, case.kernel, case.grid, values, (0.0, 0.0),
                        binding(case))
    assert np.abs(values - case.heights).max() <= 1e-4
