# SPDX-License-Identifier: BSD-3-Clause
"""Inputs of the datum relief kernels (crater_datums.slang, crater_body_datums.slang): a prepared
crater batch with complex craters and central peaks, the size-band index of cells the plane kernel
walks (laid out as the reference's DatumKernelIndex: per band its cell size and first cell
indices, its cells across and down and first cell, per cell its first entry and count, entries
ascending), the fitted craters' lattice rows, and on a sphere each row's listed craters, bounds,
frames and lattice positions. Seeded: every run builds the same inputs."""
import ctypes as C

import numpy as np

from terraingen_amd import native as N
from terraingen_amd.pipeline import CraterLayerSpec

SIDE = 7  # lattice points per side (kDatumLatticeSide)
MOON = 1737400.0


def spec():
    return CraterLayerSpec(sfd_coefficient=0.05, sfd_exponent=2.0, min_diameter=2.0,
                           max_diameter=120.0, complex_=1, transition_diameter=40.0, peaks=1,
                           seed=7)


def plane_batch(region=400.0, seed=3):
    s = spec()
    craters = N.generate_craters(N.Generation(
        min_x=0.0, min_y=0.0, max_x=region, max_y=region, sfd_coefficient=s.sfd_coefficient,
        sfd_exponent=s.sfd_exponent, min_diameter=s.min_diameter, max_diameter=s.max_diameter,
        lifetime_gyr=1.0, seed=seed))
    kernel, records, data, foot = N.prepare_craters(craters, s.models(), s.apply(), s.seed, 0.5)
    return kernel, records, data, foot


def band_index(kernel, records, foot):
    """Bands of square cells by diameter octave, each as wide as its craters' profile reach, the
    craters listed in the cells holding their centres, ascending."""
    d = foot[:, 3]
    band = np.floor(np.log2(d)).astype(np.int64)
    floats, cells_meta, cells, entries = [], [], [], []
    for b in np.unique(band):
        members = np.nonzero(band == b)[0]
        size = np.float32(1.0001 * kernel.profile_reach_radii * 2.0 ** b)
        x = records["x"][members].astype(np.float64) + records["x_low"][members]
        y = records["y"][members].astype(np.float64) + records["y_low"][members]
        ci = np.floor(x / size).astype(np.int64)
        cj = np.floor(y / size).astype(np.int64)
        i0, j0 = ci.min(), cj.min()
        across, down = int(ci.max() - i0 + 1), int(cj.max() - j0 + 1)
        first_cell = len(cells) // 2
        floats += [size, np.float32(i0), np.float32(j0)]
        cells_meta += [across, down, first_cell]
        slot = (cj - j0) * across + (ci - i0)
        for c in range(across * down):
            listed = np.sort(members[slot == c])
            cells += [len(entries), listed.size]
            entries += listed.tolist()
    return (np.array(floats, np.float32), np.array(cells_meta, np.uint32),
            np.array(cells, np.uint32), np.array(entries if entries else [0], np.uint32),
            len(np.unique(band)))


def plane_case(rows=None):
    """(kernel struct, extras, heights, params) of the plane kernel over every third crater."""
    kernel, records, data, foot = plane_batch()
    floats, meta, cells, entries, bands = band_index(kernel, records, foot)
    targets = np.arange(2, records.size, 3, dtype=np.uint32)
    kd = N.CraterDatumKernel(shape=kernel, fit_radii=1.5, lattice_side=SIDE, band_count=bands)
    n = targets.size if rows is None else rows
    heights = np.random.default_rng(1).normal(size=(n, SIDE * SIDE)).astype(np.float32)
    extras = [records, data, floats, meta, cells, entries, targets]
    return kd, extras, heights


def body_case():
    """(kernel struct, extras, heights) of the body kernel: craters on the Moon near (30, -10),
    each row's listed craters those before it whose reach meets its footprint."""
    s = spec()
    craters = N.generate_sphere_craters(MOON, (30.0, -10.0, 30.01, -9.99), s.sfd_coefficient,
                                        s.sfd_exponent, s.min_diameter, s.max_diameter,
                                        lifetime_gyr=1.0, seed=5)
    kernel, records, data, foot = N.prepare_craters(craters, s.models(), s.apply(), s.seed, 0.5)
    frames = N.sphere_frames(foot[:, 0], foot[:, 1], MOON)
    bounds = (1.0001 * kernel.profile_reach_radii * 0.5 * foot[:, 3] * 1.3).astype(np.float32)
    unit = MOON * np.pi / 180.0
    targets = np.arange(3, records.size, 4)
    lists, entries, positions = [], [], []
    for k in targets:
        fit = 1.5 * 0.5 * foot[k, 3]
        step = 2.0 * fit / SIDE / unit
        g = N.Grid(SIDE, SIDE, foot[k, 0] - fit / unit, foot[k, 1] + fit / unit, step)
        positions.append(N.sphere_split_points(g, MOON))
        near = np.hypot((foot[:k, 0] - foot[k, 0]) * np.cos(np.radians(foot[k, 1])),
                        foot[:k, 1] - foot[k, 1]) * unit < fit * 1.5 + bounds[:k]
        listed = np.nonzero(near)[0]
        lists += [len(entries), listed.size]
        entries += listed.tolist()
    heights = np.random.default_rng(2).normal(size=(targets.size, SIDE * SIDE)).astype(
        np.float32)
    heights[0, 5] = np.nan  # a point with no data stays as it is
    kd = N.CraterBodyDatumKernel(shape=kernel)
    extras = [records, data, frames, bounds, np.array(lists, np.uint32),
              np.array(entries if entries else [0], np.uint32), np.concatenate(positions)]
    return kd, extras, heights


def params(heights, first_row=0, no_data=None):
    p = N.KernelParams(heights.shape[1], heights.shape[0], 0, 0.0, first_row)
    if no_data == "nan":
        p.no_data = 1  # TG_NODATA_NAN
    return p


def ptrs(extras):
    return [np.ascontiguousarray(e).ctypes.data for e in extras]
