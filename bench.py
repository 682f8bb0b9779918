# SPDX-License-Identifier: BSD-3-Clause
"""Throughput of the hip backend's noise + crater shaping hot path on one MI355X.

Workload (one step): the synthetic source 4097 x 4097 at 0.5 m/px, 8-octave Perlin fBm (base
wavelength 512 m, lacunarity 2, gain 0.5, amplitude 20 m, seed 1), evaluated into FP32 heights
already resident in HBM. Prints one JSON line.
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

# Algorithmic work of one Perlin fBm octave per cell, in lane operations: rotation 6, scale and
# offset 4, lattice split 6, 4 corner hashes x 18, 4 gradients x 21, offsets 2, fades 14, lerps 9,
# sqrt(2) 1, weighted sum 2 (include/terraingen/noise/fractal.h:228-232, perlin.h:17-33,72-90,
# hash.h:40-57). Frozen; never tuned to a measurement.
OPS_PER_PERLIN_OCTAVE = 200
VALU_PEAK = 256 * 64 * 2.4e9  # lane-ops/s: 256 CUs x 64 lanes x 2.4 GHz
HBM_PEAK_GBS = 8000.0


def cpu_baseline(kernel, xs, ys, rows_sample):
    """The reference's CPU noise path (FP32 kernel math, rows over all host threads), timed by
    the oracle restatement on a bounded sample of the same grid's rows."""
    from oracle import oracle
    from terraingen_amd import native as N

    orc = oracle()
    cores = os.cpu_count() or 1
    sub = np.ascontiguousarray(ys[:rows_sample])
    out = np.zeros((rows_sample, xs.size), np.float32)
    params = N.KernelParams(xs.size, rows_sample, 0, 0.0, 0)
    orc.or_noise_grid(C.byref(kernel), xs.ctypes.data, sub.ctypes.data, xs.size, rows_sample,
                      C.byref(params), out.ctypes.data, cores)  # warm-up
    best = float("inf")
    for _ in range(3):
        out[:] = 0
        t0 = time.perf_counter()
        orc.or_noise_grid(C.byref(kernel), xs.ctypes.data, sub.ctypes.data, xs.size,
                          rows_sample, C.byref(params), out.ctypes.data, cores)
        best = min(best, time.perf_counter() - t0)
    cells = rows_sample * xs.size
    return {"value": cells / best / 1e9, "unit": "Gcells/s", "cores": cores, "kind": "port",
            "sample": f"{rows_sample} rows x {xs.size} cells of the 4097^2 grid, best of 3, "
                      f"{platform.processor() or platform.machine()}"}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU; tiles shard without exchange "
                                   "across GPUs but multi-GPU runs are not part of it"}))
        return 2

    from terraingen_amd import native as N

    dev = N.HipDevice(0)
    n = 4097
    pixel = 0.5
    fractal = dict(seed=1, wavelength=512.0, amplitude=20.0, octaves=8, lacunarity=2.0, gain=0.5)
    kernel = N.noise_kernel([fractal])
    half = 0.5 * n * pixel
    xs = (-half + (np.arange(n) + 0.5) * pixel).astype(np.float32)
    ys = (half - (np.arange(n) + 0.5) * pixel).astype(np.float32)

    err = C.create_string_buffer(512)
    plan = C.c_void_p()
    if dev.lib.tg_hip_noise_plan_create(dev.handle, C.byref(kernel), C.byref(plan), err, 512):
        raise RuntimeError(err.value.decode())
    heights = N.DeviceArray(dev, n * n * 4)
    heights.zero()
    dxs = N.DeviceArray.from_host(dev, xs)
    dys = N.DeviceArray.from_host(dev, ys)
    params = N.KernelParams(n, n, 0, 0.0, 0)
    sv = C.c_void_p()
    dev.call("tg_hip_stream_create", dev.handle, C.byref(sv))
    stream = sv.value

    def step():
        code = dev.lib.tg_hip_noise_launch(dev.handle, plan, heights.ptr, dxs.ptr, dys.ptr,
                                           C.byref(params), stream, err, 512)
        if code:
            raise RuntimeError(err.value.decode())

    for _ in range(args.warmup):
        step()
    dev.synchronize(stream)
    start, stop = N.Event(dev), N.Event(dev)
    t0 = time.perf_counter()
    start.record(stream)
    for _ in range(args.steps):
        step()
    stop.record(stream)
    dev.synchronize(stream)
    wall = time.perf_counter() - t0
    kernel_ms = start.elapsed_ms(stop) / args.steps
    cells = n * n
    gcells = cells / (kernel_ms * 1e-3) / 1e9
    ops = cells * fractal["octaves"] * OPS_PER_PERLIN_OCTAVE
    achieved = ops / (kernel_ms * 1e-3)
    bytes_moved = cells * 4 * 2  # read-modify-write of the heights
    result = {
        "metric": "Gcells/s batched vs plain HIP at <=1e-3 m err; p50/p99 tile latency",
        "value": round(gcells, 3),
        "unit": "Gcells/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(kernel_ms, 4),
        "wall_ms_per_step": round(wall * 1e3 / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+u32",
        "data": "synthetic",
        "config": {"workload": "synthetic source 4097x4097 @0.5 m, 8-octave Perlin fBm "
                               "(512 m, lac 2, gain 0.5, 20 m, seed 1): noise stage",
                   "cells": cells},
        "roofline": {"bound": "valu", "achieved": round(achieved / 1e12, 3),
                     "peak": round(VALU_PEAK / 1e12, 3), "unit": "Tops/s",
                     "frac": round(achieved / VALU_PEAK, 4), "traffic": None,
                     "hbm_GBs": round(bytes_moved / (kernel_ms * 1e-3) / 1e9, 1)},
    }
    if not args.no_cpu_baseline:
        result["cpu_baseline"] = cpu_baseline(kernel, xs, ys, 64)
    dev.lib.tg_hip_noise_plan_free(plan)
    print(json.dumps(result), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
