# SPDX-License-Identifier: BSD-3-Clause
"""Throughput of the hip backend's noise + crater shaping hot path on one MI355X.

Workload: the synthetic source 4097 x 4097 at 0.5 m/px (north-up, centred on 0, source origin
{0, 0}), 8-octave Perlin fBm (base wavelength 512 m, lacunarity 2, gain 0.5, amplitude 20 m,
seed 1), plus craters from the power-law cumulative SFD N(>D) = 0.05 D^-2 per m^2, D in [1, 200] m,
uniform centres, ages uniform over 1 Gyr, seed 1 (Martin-Parkes-Dunstan profile, freshness
degradation, irregular outlines), FP32 heights. Host placement, preparation and binning run once
(reported as host_ms); one step is the device evaluation of the whole grid with every input
resident in HBM: the source fill with noise, then the crater kernel. Prints one JSON line.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import platform
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

# Algorithmic work, in lane operations (each + - x /, compare/select, integer op, conversion,
# floor/sqrt/transcendental counts 1), from the reference's source:
#   a Perlin fBm octave per cell: rotation 6, scale and offset 4, lattice split 6, 4 corner hashes
#   x 18, 4 gradients x 21, offsets 2, fades 14, lerps 9, sqrt(2) 1, weighted sum 2
#   (include/terraingen/noise/fractal.h:228-232, perlin.h:17-33,72-90, hash.h:40-57);
#   a (cell, crater) pair inside the reach: 110 (split offset, distance, reach test, profile,
#   blend; craters.h:1186-1211) + 12 per outline harmonic (craters.h:787-803);
#   a pair outside the reach: the same up to the reach test, 40 + 12 per harmonic.
# Frozen; never tuned to a measurement.
OPS_PER_PERLIN_OCTAVE = 200
OPS_PAIR_IN = 110
OPS_PAIR_OUT = 40
OPS_PER_HARMONIC = 12
VALU_PEAK = 256 * 64 * 2.4e9  # lane-ops/s: 256 CUs x 64 lanes x 2.4 GHz
HBM_PEAK_GBS = 8000.0

C2_FRACTAL = dict(seed=1, wavelength=512.0, amplitude=20.0, octaves=8, lacunarity=2.0, gain=0.5)


def pair_ops(terrain) -> tuple[float, int, int]:
    """Algorithmic lane ops of the crater kernel over the grid: every (cell, listed crater) pair,
    split by whether the cell lies within the crater's round reach (counted on the host from the
    tile lists and the records, not by the kernel)."""
    g = terrain.grid
    side = int(terrain.tiles[1])
    tiles_x = int(terrain.tiles[0])
    tiles_y = (g.height + side - 1) // side
    rec = terrain.records
    reach_r = float(terrain.kernel.reach_radii) * 0.5 * rec["diameter"].astype(np.float64)
    cx = rec["x"].astype(np.float64) + rec["x_low"]
    cy = rec["y"].astype(np.float64) + rec["y_low"]
    harm = rec["harmonic_count"].astype(np.float64)
    xs, ys = N_grid_centers(g)
    total_in = total_out = 0
    ops = 0.0
    for ty in range(tiles_y):
        r0, r1 = ty * side, min(g.height, (ty + 1) * side)
        for tx in range(tiles_x):
            t = ty * tiles_x + tx
            first, count = int(terrain.tiles[3 + 2 * t]), int(terrain.tiles[4 + 2 * t])
            if count == 0:
                continue
            c0, c1 = tx * side, min(g.width, (tx + 1) * side)
            idx = terrain.entries[first:first + count]
            dx = xs[c0:c1][None, None, :] - cx[idx][:, None, None]
            dy = ys[r0:r1][None, :, None] - cy[idx][:, None, None]
            inside = (dx * dx + dy * dy) < (reach_r[idx] ** 2)[:, None, None]
            n_in = inside.sum(axis=(1, 2))
            cells = (r1 - r0) * (c1 - c0)
            n_out = cells - n_in
            ops += float((n_in * (OPS_PAIR_IN + OPS_PER_HARMONIC * harm[idx])).sum())
            ops += float((n_out * (OPS_PAIR_OUT + OPS_PER_HARMONIC * harm[idx])).sum())
            total_in += int(n_in.sum())
            total_out += int(n_out.sum())
    return ops, total_in, total_out


def N_grid_centers(g):
    from terraingen_amd import native as N

    return N.grid_centers(g)


def cpu_baseline(terrain, rows: int) -> dict:
    """The reference's CPU path (FP32 kernel math over all host threads: noise rows, then crater
    tiles), timed by the oracle restatement on a bounded sample of the grid's rows; also the
    hip result's error on those rows against the FP64 oracle."""
    from oracle import oracle
    from terraingen_amd import native as N

    orc = oracle()
    g = terrain.grid
    cores = os.cpu_count() or 1
    r0 = (g.height - rows) // 2
    sub_ys = np.ascontiguousarray(terrain.ys[r0:r0 + rows])
    kernel = N.noise_kernel([C2_FRACTAL])
    params = N.KernelParams(g.width, rows, 0, 0.0, r0)
    out = np.zeros((rows, g.width), np.float32)

    def run_once():
        out[:] = 0.0
        orc.or_noise_grid(C.byref(kernel), terrain.xs.ctypes.data, sub_ys.ctypes.data, g.width,
                          rows, C.byref(params), out.ctypes.data, cores)
        orc.or_shape_grid(C.byref(terrain.kernel), terrain.records.ctypes.data,
                          terrain.data.ctypes.data, terrain.tiles.ctypes.data,
                          terrain.entries.ctypes.data, terrain.xs.ctypes.data,
                          sub_ys.ctypes.data, terrain.lows.ctypes.data, g.width, rows,
                          C.byref(params), out.ctypes.data, cores)

    run_once()
    best = float("inf")
    for _ in range(3):
        t0 = time.perf_counter()
        run_once()
        best = min(best, time.perf_counter() - t0)
    # FP64 oracle on the same rows.
    cx64, cy64 = N.grid_centers(g)
    xs64 = np.ascontiguousarray(cx64)
    ys64 = np.ascontiguousarray(cy64[r0:r0 + rows])
    ref64 = np.zeros((rows, g.width), np.float64)
    orc.or_noise_grid_f64(C.byref(kernel), xs64.ctypes.data, ys64.ctypes.data, g.width, rows,
                          ref64.ctypes.data, cores)
    orc.or_shape_grid_f64(C.byref(terrain.kernel), terrain.records.ctypes.data,
                          terrain.data.ctypes.data, terrain.tiles.ctypes.data,
                          terrain.entries.ctypes.data, xs64.ctypes.data, ys64.ctypes.data,
                          g.width, rows, r0, ref64.ctypes.data, cores)
    cells = rows * g.width
    return {"value": round(cells / best / 1e9, 5), "unit": "Gcells/s", "cores": cores,
            "kind": "port",
            "sample": f"rows {r0}..{r0 + rows - 1} (x {g.width} cells) of the 4097^2 grid: noise "
                      f"then craters, best of 3, {platform.processor() or platform.machine()}",
            "_ref64": ref64, "_r0": r0, "_fp32": out.copy()}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=30)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--size", type=int, default=4097)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU; tiles shard across GPUs without "
                                   "exchange, but multi-GPU runs are not part of it"}))
        return 2

    from terraingen_amd import native as N
    from terraingen_amd.pipeline import CraterLayerSpec, PlaneTerrain

    dev = N.HipDevice(0)
    n = args.size
    grid = N.north_up_grid(n, n, 0.5)
    terrain = PlaneTerrain(dev, grid, [C2_FRACTAL], 0.0, CraterLayerSpec(seed=1))
    t0 = time.perf_counter()
    terrain.prepare()
    host_ms = (time.perf_counter() - t0) * 1e3
    heights = N.DeviceArray(dev, n * n * 4)
    sv = C.c_void_p()
    dev.call("tg_hip_stream_create", dev.handle, C.byref(sv))
    stream = sv.value

    for _ in range(args.warmup):
        terrain.launch(heights, stream)
    dev.synchronize(stream)
    start, stop = N.Event(dev), N.Event(dev)
    t0 = time.perf_counter()
    start.record(stream)
    for _ in range(args.steps):
        terrain.launch(heights, stream)
    stop.record(stream)
    dev.synchronize(stream)
    wall = time.perf_counter() - t0
    step_ms = start.elapsed_ms(stop) / args.steps

    # Each kernel alone, on the same stream, for the roofline of the dominant one.
    def time_kernel(fn):
        a, b = N.Event(dev), N.Event(dev)
        a.record(stream)
        for _ in range(args.steps):
            fn()
        b.record(stream)
        dev.synchronize(stream)
        return a.elapsed_ms(b) / args.steps

    noise_ms = time_kernel(lambda: terrain.launch_noise(heights, stream))
    crater_ms = time_kernel(lambda: terrain.launch_craters(heights, stream))
    terrain.launch(heights, stream)
    dev.synchronize(stream)
    got = heights.download(np.empty((n, n), np.float32))

    cells = n * n
    gcells = cells / (step_ms * 1e-3) / 1e9
    noise_ops = cells * C2_FRACTAL["octaves"] * OPS_PER_PERLIN_OCTAVE
    crater_ops, pairs_in, pairs_out = pair_ops(terrain)
    dominant = "craters" if crater_ms >= noise_ms else "noise"
    dom_ms = max(crater_ms, noise_ms)
    dom_ops = crater_ops if dominant == "craters" else noise_ops
    achieved = dom_ops / (dom_ms * 1e-3)
    result = {
        "metric": "Gcells/s batched vs plain HIP at <=1e-3 m err; p50/p99 tile latency",
        "value": round(gcells, 3),
        "unit": "Gcells/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(step_ms, 4),
        "wall_ms_per_step": round(wall * 1e3 / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+u32",
        "data": "synthetic",
        "config": {"workload": f"synthetic source {n}x{n} @0.5 m: 8-octave Perlin fBm (512 m, "
                               "lac 2, gain 0.5, 20 m, seed 1) + power-law craters "
                               "N(>D)=0.05 D^-2, D 1-200 m, seed 1",
                   "cells": cells, "craters": int(terrain.crater_count),
                   "pairs_in_reach": pairs_in, "pairs_listed": pairs_in + pairs_out},
        "kernels_ms": {"noise_fill": round(noise_ms, 4), "craters": round(crater_ms, 4)},
        "host_ms": {k: round(v, 1) for k, v in terrain.timings.items()} | {
            "total_prepare": round(host_ms, 1)},
        "roofline": {"bound": "valu", "kernel": dominant, "achieved": round(achieved / 1e12, 3),
                     "peak": round(VALU_PEAK / 1e12, 3), "unit": "Tops/s",
                     "frac": round(achieved / VALU_PEAK, 4), "traffic": None,
                     "hbm_write_GBs": round(cells * 4 / (step_ms * 1e-3) / 1e9, 1)},
    }
    if not args.no_cpu_baseline:
        base = cpu_baseline(terrain, 16)
        r0 = base.pop("_r0")
        ref64 = base.pop("_ref64")
        fp32 = base.pop("_fp32")
        d = np.abs(got[r0:r0 + ref64.shape[0]] - ref64)
        result["cpu_baseline"] = base
        result["error_vs_fp64_oracle_m"] = {"max": float(f"{d.max():.3g}"),
                                            "rms": float(f"{np.sqrt((d ** 2).mean()):.3g}"),
                                            "rows": int(ref64.shape[0])}
        result["max_diff_vs_fp32_oracle_m"] = float(
            f"{np.abs(got[r0:r0 + fp32.shape[0]] - fp32).max():.3g}")
    terrain.close()
    print(json.dumps(result), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
