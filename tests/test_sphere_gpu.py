# SPDX-License-Identifier: BSD-3-Clause
"""The crater kernel's body branch (ShapeCratersOnBody, craters.h:1359-1370: each cell's split
position, each crater's frame) on the MI355X against the reference's CPU heights of crater layers
on a sphere (tests/golden/craters_sphere_*.npz) and the FP64 body oracle."""
import numpy as np
import pytest

from terraingen_amd import native as N

from crater_cases import (SPHERE_NAMES, load_sphere, sphere_binding, sphere_oracle_heights,
                          sphere_oracle_heights_f64)

pytestmark = pytest.mark.gpu

TOL_FP64 = 1e-3  # max |dh| against the FP64 oracle, meters


@pytest.fixture(scope="module", params=SPHERE_NAMES)
def case(request):
    return load_sphere(request.param)


def run(hip, case, rows=None, first_row=0, heights=None):
    r = slice(None) if rows is None else rows
    out = np.zeros(case.shape, np.float32)[r].copy() if heights is None else heights
    params = N.KernelParams(case.side, out.shape[0], N.NODATA_NONE if heights is None else
                            N.NODATA_VALUE, -32768.0, first_row)
    # Columns and rows are unused on a body; RunKernel binds them all the same.
    xs = np.zeros(case.side, np.float32)
    ys = np.zeros(out.shape[0], np.float32)
    hip.run(N.TG_KERNEL_CRATERS, case.kernel, out, xs, ys, params, sphere_binding(case, rows))
    return out


def test_body_kernel_matches_reference_and_oracles(hip, orc, case):
    print(f"[sphere {case.name}] {case.records.size} craters, {case.side}^2 cells", flush=True)
    got = run(hip, case)
    d_ref = np.abs(got - case.heights)
    d32 = np.abs(got - sphere_oracle_heights(orc, case))
    d64 = np.abs(got - sphere_oracle_heights_f64(orc, case))
    print(f"[sphere {case.name}] GPU vs reference CPU: max {d_ref.max():.3g} m; vs FP32 oracle "
          f"max {d32.max():.3g}; vs FP64 oracle max {d64.max():.3g} rms "
          f"{np.sqrt((d64 ** 2).mean()):.3g} m", flush=True)
    assert d64.max() <= TOL_FP64
    # FP32 rounding of the fast outline and an ulp of the cell positions (3.1e-5 m measured).
    assert d_ref.max() <= 3e-4


def test_body_strips_join(hip, case):
    whole = run(hip, case)
    joined = np.zeros_like(whole)
    for r0 in range(0, case.side, 37):
        r1 = min(case.side, r0 + 37)
        joined[r0:r1] = run(hip, case, slice(r0, r1), r0)
    assert np.array_equal(whole.view(np.uint32), joined.view(np.uint32))


def test_body_no_data_cells_untouched(hip, case):
    base = np.zeros(case.shape, np.float32)
    base[::7, ::5] = -32768.0
    got = run(hip, case, heights=base)
    assert np.all(got[::7, ::5] == -32768.0)
    ref = run(hip, case)
    keep = np.ones(case.shape, bool)
    keep[::7, ::5] = False
    assert np.array_equal(got[keep].view(np.uint32), ref[keep].view(np.uint32))


def test_body_through_run_kernel(hip, case):
    """RunKernel's mirror slices the per-cell split positions with each strip."""
    grid = N.Grid(case.side, case.side, case.lon, case.lat, case.step)
    whole = run(hip, case)
    strips = np.zeros(case.shape, np.float32)
    assert N.run_kernel(N.backend_named("hip"), N.TG_KERNEL_CRATERS, case.kernel, grid, strips,
                        (0.0, 0.0), sphere_binding(case), max_strip_bytes=case.side * 4 * 29)
    assert np.array_equal(whole.view(np.uint32), strips.view(np.uint32))
