# SPDX-License-Identifier: BSD-3-Clause
"""The datum relief kernels (TG_KERNEL_CRATER_DATUMS, TG_KERNEL_CRATER_BODY_DATUMS): the oracle's
restatement of KernelDatumRelief and KernelBodyDatumRelief (craters.h:1225-1326) gives the
reference's own values bit for bit (oracle/_ref) on a prepared batch with complex craters and
central peaks, on a plane and on the Moon; the GPU kernels, run through tg_hip_run as RunKernel
binds them, give the oracle's values within TOL (the device's powf, atan2f and cosf differ from the
host's libm by an ulp or two, which the profiles, outlines and peak massifs pass on), also over
strips; malformed indices are refused on the host."""
import ctypes as C

import numpy as np
import pytest

from terraingen_amd import native as N

from datum_cases import body_case, params, plane_case, ptrs

# m: max |dh| of the GPU kernels against the oracle's FP32 restatement; achieved 2.9e-6 on the
# plane case and 4.8e-6 on the Moon.
TOL = 1e-5


def oracle_plane(orc, kd, extras, heights, first_row=0):
    out = heights.copy()
    orc.or_datum_grid(C.byref(kd), *ptrs(extras), C.byref(params(out, first_row)),
                      out.ctypes.data, 4)
    return out


def oracle_body(orc, kd, extras, heights, first_row=0, no_data="nan"):
    out = heights.copy()
    orc.or_body_datum_grid(C.byref(kd), *ptrs(extras), C.byref(params(out, first_row, no_data)),
                           out.ctypes.data, 4)
    return out


def test_plane_oracle_equals_the_reference(orc, ref):
    kd, extras, heights = plane_case()
    want = heights.copy()
    ref.ref_datum_grid(C.byref(kd), *ptrs(extras), C.byref(params(want)), want.ctypes.data)
    got = oracle_plane(orc, kd, extras, heights)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
    moved = np.abs(got - heights)
    assert (moved > 1e-3).mean() > 0.3, "the case must have relief at most lattice points"
    # Points outside the circular footprint add nothing.
    side = 7
    c = (2.0 * np.arange(side) + 1.0) / side - 1.0
    outside = (c[None, :] ** 2 + c[:, None] ** 2).ravel() >= 1.0
    assert (moved[:, outside] == 0.0).all()


def test_body_oracle_equals_the_reference(orc, ref):
    kd, extras, heights = body_case()
    want = heights.copy()
    ref.ref_body_datum_grid(C.byref(kd), *ptrs(extras), C.byref(params(want, 0, "nan")),
                            want.ctypes.data)
    got = oracle_body(orc, kd, extras, heights)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
    assert np.isnan(got[0, 5])
    assert (np.abs(np.nan_to_num(got - heights)) > 1e-3).mean() > 0.3


@pytest.mark.gpu
def test_plane_kernel_equals_the_oracle(hip, orc):
    kd, extras, heights = plane_case()
    want = oracle_plane(orc, kd, extras, heights)
    got = heights.copy()
    xs, ys = np.zeros(got.shape[1], np.float32), np.zeros(got.shape[0], np.float32)
    hip.run(N.TG_KERNEL_CRATER_DATUMS, kd, got, xs, ys, params(got),
            [(e, 0) for e in extras])
    d = np.abs(got - want)
    print(f"[datums] plane: {got.shape[0]} craters x {got.shape[1]} points, max |dh| "
          f"{d.max():.3g} m, {np.mean(got.view(np.uint32) == want.view(np.uint32)):.4f} bit-equal",
          flush=True)
    assert d.max() <= TOL


@pytest.mark.gpu
def test_plane_kernel_strips_join(hip, orc):
    """A strip from row 5 reads targets[5 + r], as RunKernel's strips do."""
    kd, extras, heights = plane_case()
    want = oracle_plane(orc, kd, extras, heights)
    part = heights[5:17].copy()
    hip.run(N.TG_KERNEL_CRATER_DATUMS, kd, part, np.zeros(49, np.float32),
            np.zeros(12, np.float32), params(part, 5), [(e, 0) for e in extras])
    whole = heights.copy()
    hip.run(N.TG_KERNEL_CRATER_DATUMS, kd, whole, np.zeros(49, np.float32),
            np.zeros(whole.shape[0], np.float32), params(whole), [(e, 0) for e in extras])
    assert np.array_equal(part.view(np.uint32), whole[5:17].view(np.uint32))
    assert np.abs(part - want[5:17]).max() <= TOL


@pytest.mark.gpu
def test_body_kernel_equals_the_oracle(hip, orc):
    kd, extras, heights = body_case()
    want = oracle_body(orc, kd, extras, heights)
    got = heights.copy()
    hip.run(N.TG_KERNEL_CRATER_BODY_DATUMS, kd, got, np.zeros(got.shape[1], np.float32),
            np.zeros(got.shape[0], np.float32), params(got, 0, "nan"),
            [(e, 1 if i == 6 else 0) for i, e in enumerate(extras)])
    assert np.isnan(got[0, 5]) and np.isnan(want[0, 5])
    d = np.abs(np.nan_to_num(got - want))
    print(f"[datums] body: {got.shape[0]} craters, max |dh| {d.max():.3g} m, "
          f"{np.mean(got.view(np.uint32) == want.view(np.uint32)):.4f} bit-equal", flush=True)
    assert d.max() <= TOL


@pytest.mark.gpu
def test_malformed_indices_are_refused(hip):
    kd, extras, heights = plane_case(rows=4)
    xs, ys = np.zeros(49, np.float32), np.zeros(4, np.float32)
    bad = list(extras)
    bad[6] = np.array([10 ** 6] * 4, np.uint32)  # a target past the records
    with pytest.raises(N.HipError, match="target"):
        hip.run(N.TG_KERNEL_CRATER_DATUMS, kd, heights.copy(), xs, ys, params(heights),
                [(e, 0) for e in bad])
    bad = list(extras)
    cells = extras[4].copy()
    cells[1] = 10 ** 7  # a cell listing entries past the end
    bad[4] = cells
    with pytest.raises(N.HipError, match="cell"):
        hip.run(N.TG_KERNEL_CRATER_DATUMS, kd, heights.copy(), xs, ys, params(heights),
                [(e, 0) for e in bad])
