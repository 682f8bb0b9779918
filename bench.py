# SPDX-License-Identifier: BSD-3-Clause
"""Throughput of the hip backend's noise + crater shaping hot path on one MI355X.

Default workload (``--workload sphere``): 1024 tiles of mixed size and cell size on the Moon,
evaluated in one persistent launch. Tiles are longitude-latitude rasters on a 1737.4 km sphere:
edges 2^k + 1 with k uniform in 6..11, cells log-uniform in 0.25-8 m (in degrees over the
sphere's unit length), north-west corners uniform on the sphere. Each cell's body-fixed position
is computed in FP64 on the device. The noise is an 8-octave 3D Perlin fBm at that position
(EvaluateFractals3: base wavelength 512 m, lacunarity 2, gain 0.5, amplitude 20 m, seed 1). A
global catalogue of about 1e6 craters follows the power-law cumulative SFD N(>D) = 0.05 D^-2 per
m^2 over the sphere (D from 1374 m to 20 km), ages uniform over 1 Gyr, seed 4, placed as the
reference places them (GenerateCraters in cube-sphere cells over the whole Moon, bit for bit);
each crater is placed by its FP64 frame and offsets are taken along the sphere. The
profile is Martin-Parkes-Dunstan with freshness degradation and irregular outlines. Heights are
FP32.

``--workload tiles`` is the same batch on a plane: west/north edges uniform over a 4472 m square,
the 2D noise with origin {0, 0}, and a catalogue of about 1e6 craters with D in [1, 200] m over
the square.

* ``value``: Gcells/s of the batched path. One step is one graph launch of the persistent kernel
  over every tile, with the catalogue, bins and outputs resident in HBM.
* ``plain``: Gcells/s of the plain per-tile path over the same tiles (every 4th tile on the
  sphere). Each tile goes through ``tg_hip_run``, once for the noise layer and once for the
  craters, synchronous with host memory in and out, as RunKernel drives a device. Host preparation
  of its inputs is excluded.
* ``serving``: the served stream, run through the tile server (continuous batching, several
  batches in flight, a 64 GB HBM admission pool). Tile requests arrive as a Poisson process at
  2000/s for 10 s from a camera flying 50 km at 200 m altitude. Each request is the quadtree
  tile, at levels 0-10 by slant range, that holds a ground point near the camera, with edges of
  129, 257, 513 or 1025 cells. With the default workload the track is a great circle on the
  Moon and the tiles are longitude-latitude ones (3D body noise, frame-placed craters); with
  ``tiles`` it is a straight track on a plane. The crater catalogue (C3's population: N(>D) =
  0.0156 D^-2.2 per m^2, D 0.5-300 m, complex from 50 m with central peaks and wall terraces from
  150 m, seed 5) covers a 2 km swath along the track, placed on the sphere as the reference places
  it (one generation over the union of the track's boxes), with boulders placed as the
  reference's BoulderLayer places them and drawn after the craters. It reports the
  p50/p99 request-to-host latency (submission to heights in the caller's memory) for 513 x 513
  tiles and for all tiles, and the served Gcells/s.

``--workload grid`` is one 4097 x 4097 grid at 0.5 m, with the same noise and craters (seed 1):
``value`` is the grid as one tile of the batched kernel, and the strip kernels' passes (the plain
path) are timed beside it and checked bit for bit against it. ``--workload c3`` is one 8193 x 8193
grid at 0.25 m with complex craters and boulders, measured the same way. ``--workload sources`` is the 4097 x 4097 grid at 0.5 m filled by a
SyntheticSource whose noise lies on posts 2 m apart, upsampled by the kernel upsamplers.

Prints one JSON line.
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
#   a (cell, crater) pair whose cell lies inside the crater's footprint (reach radii x outline
#   bound x radius): 110 (split offset, distance, reach test, profile, blend;
#   craters.h:1186-1211) + 12 per outline harmonic (craters.h:787-803) + 12 for an elongated
#   outline. Pairs outside the footprint count nothing.
# Frozen; never tuned to a measurement.
OPS_PER_PERLIN_OCTAVE = 200
OPS_PAIR = 110
OPS_PER_HARMONIC = 12
OPS_ELLIPSE = 12
#   a (cell, boulder) pair inside the dome's reach: 40 (offset, hypot, two clamps, two
#   sqrt + asin band areas, the max composite; boulder_relief.cpp:44-57,244-263).
OPS_BOULDER_PAIR = 40
#   on a sphere: a Perlin3 octave per cell: rotation 15 and scale 6 in FP64, lattice split 12,
#   8 corners x (key 4, hash 20, offsets 3, gradient 18), fades 21, lerps 21, clamp 2, weighted
#   sum 2 (fractal.h:284-317, perlin.h:38-63,113-128, hash.h:40-57, lattice.h:59-69); the cell's
#   body-fixed position 15 (two sines, two cosines, five products; surface.cpp:45-50); a
#   (cell, crater) pair adds to the plane's the frame offset, 35 (three split differences, two
#   dot products, the chord, its arcsine ratio and the rescale; craters.h:168-209).
OPS_PER_PERLIN3_OCTAVE = 440
OPS_BODY_POSITION = 15
OPS_BODY_PAIR_EXTRA = 35
# MI355X's published FP32 vector peak. It counts a packed FMA as 4 flops per lane (2 lanes x 2
# flops), so it bounds any count of one lane-op per operation from above: the kernels are built
# without packed FP32, and their issue ceiling is a quarter of it, 39.3e12 lane-instructions/s
# (256 CUs x 64 lanes x 2.4 GHz).
VALU_PEAK_TFLOPS = 157.3
ISSUE_CEILING = 256 * 64 * 2.4e9
HBM_PEAK_GBS = 8000.0
# MI355X's published FP64 vector peak (an FMA as 2 flops a lane).
FP64_PEAK_TFLOPS = 78.6
#   kernel upsampling per cell, all in FP64 (src/raster/kernels.h:31-78, resample.cpp:19-135,
#   geo_transform.h:63-79): the cell centre and its lattice position 26 (two PixelToWorld rows,
#   two differences, the 2x2 inverse, the half-post shift); floor and fraction 4; per tap and
#   axis the kernel's weight (cubic and B-spline 9: offset, fabs, two compares, five-term
#   polynomial; Lanczos 3 15: offset, fabs, three compares, floor, two sinc of mul, sin, div, the
#   third, the product, the normalising division; bilinear 1 an axis); per tap of the square 12
#   (the weight product and its zero test, two clamped post indices, the read's NaN test, the
#   widening, the weighted product and the sum); the narrowing 1. Nearest: position 26, two
#   rounds, one tap 12.
OPS_UPSAMPLE_POSITION = 27
OPS_UPSAMPLE_TAP = 12
OPS_UPSAMPLE_WEIGHT = {"cubic": 9, "cubic_bspline": 9, "lanczos3": 15}
METRIC = "Gcells/s batched vs plain HIP at <=1e-3 m err; p50/p99 tile latency"

C2_FRACTAL = dict(seed=1, wavelength=512.0, amplitude=20.0, octaves=8, lacunarity=2.0, gain=0.5)
TILES_SIDE = 4472.0  # m: 0.05 x 4472^2 = 1.0e6 craters of D >= 1 m
MOON = 1737400.0  # m
GLOBAL_CRATERS = 1_000_000
GLOBAL_MAX_D = 20000.0  # m


def roofline(ops: float, kernel_ms: float, bytes_written: float) -> dict:
    """The VALU roofline of one launch: algorithmic lane operations per second against the
    published FP32 vector peak, and the compulsory height writes against HBM bandwidth."""
    achieved = ops / (kernel_ms * 1e-3)
    return {"bound": "valu", "achieved": round(achieved / 1e12, 3),
            "peak": VALU_PEAK_TFLOPS, "unit": "TFLOP/s",
            "frac": round(achieved / (VALU_PEAK_TFLOPS * 1e12), 4),
            "peak_basis": "MI355X published FP32 vector peak (157.3 TFLOP/s, packed FMA as 4 "
                          "flops a lane); achieved counts one per algorithmic lane operation",
            "traffic": None,
            "hbm_write_GBs": round(bytes_written / (kernel_ms * 1e-3) / 1e9, 1),
            "hbm_write_frac": round(bytes_written / (kernel_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 5)}


def counted_traffic(kernel_ms: float, name: str) -> dict | None:
    """HBM bytes per launch of a 1024-tile workload, from its committed counter runs
    (profiles/`name`: FETCH_SIZE and WRITE_SIZE, each scaled by a 1 GiB copy's reading), with the
    VALU instructions those runs counted, scaled to this run's launch time."""
    path = ROOT / "profiles" / name
    if not path.exists():
        return None
    for line in path.read_text().splitlines():
        if line.strip().startswith("traffic: "):
            meta = json.loads(line.strip()[len("traffic: "):])
            # Counts of another build of the kernel would not describe this one: they are used
            # only when the counted launch took the time this one takes.
            if abs(meta["kernel_ns"] * 1e-6 / kernel_ms - 1.0) > 0.05:
                return {"traffic": None, "counters_stale": str(path.relative_to(ROOT))}
            lane = meta["valu_lane_instructions"]
            out = {"traffic": round(meta["read_bytes"] + meta["write_bytes"]),
                   "traffic_read": round(meta["read_bytes"]),
                   "traffic_write": round(meta["write_bytes"]),
                   "valu_lane_instructions": lane,
                   "counters": str(path.relative_to(ROOT))}
            if "probe_valu_rate" in meta:
                # Executed work against the issue rate the same counter reads for the probe's
                # independent chains, in the same run.
                rate = lane / (meta["kernel_ns"] * 1e-9)
                out["executed_frac"] = round(rate / meta["probe_valu_rate"], 4)
                out["executed_basis"] = ("SQ_INSTS_VALU x 64 per second of the kernel over the "
                                         "same for the VALU probe's chains, one counter run")
            return out
    return None


def valu_ceiling(N, dev) -> dict:
    """The measured vector-ALU issue ceiling (tg_hip_valu_probe): lane-ops/s of FP32 FMA chains
    and of 32-bit integer xor/add chains at full occupancy."""
    out = {}
    for kind, name in ((0, "fp32_fma"), (1, "int32_xor_add")):
        ms, ops = C.c_float(), C.c_double()
        dev.call("tg_hip_valu_probe", dev.handle, kind, 20000, 8192, C.byref(ms), C.byref(ops))
        out[name] = round(ops.value / (ms.value * 1e-3) / 1e12, 2)
    out["unit"] = "T lane-ops/s"
    return out


def percentile(values, q) -> float:
    return float(np.percentile(np.asarray(values), q))


# ---- the CPU baseline leg and the FP64 error (oracle/: test infrastructure) ---------------------

def host_cpu() -> tuple[str, int]:
    """The host CPU's model name and the threads this process may run on."""
    model = platform.processor() or platform.machine()
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count() or 1
    return model, cores


def cpu_baseline(batch, sample: list[int], sphere: bool, repeats: int = 5) -> dict:
    """The reference's CPU path over sampled tiles, on this host's cores: oracle/_ref, the
    reference's own code built from its sources, runs its noise rows (FractalStackRows on
    ParallelFor, noise_layer.cpp:28-43) then its crater tiles (ShapeCraters / ShapeCratersOnBody
    on ForEachTile, apply.cpp:199-222). The oracle's restatement runs the same tiles beside it.
    Inputs (crater batches, bins, positions) are prepared before the timing; one warm-up pass,
    then the best of `repeats` passes of each. Falls back to the restatement alone (kind "port")
    where oracle/_ref was not built."""
    from oracle import oracle, reference
    from terraingen_amd import native as N

    orc, ref = oracle(), reference()
    model, cores = host_cpu()
    inputs = [(i, batch.plain_inputs(i)) for i in sample]
    cells = sum(inp["grid"].width * inp["grid"].height for _, inp in inputs)
    k = batch.noise_kernel
    fractals = C.cast(C.byref(k.fractals), C.c_void_p)

    def run_reference():
        outs = {}
        for i, inp in inputs:
            g = inp["grid"]
            out = np.empty((g.height, g.width), np.float32)
            cr = inp["craters"]
            if sphere:
                args = ((C.byref(cr[0]), cr[1].ctypes.data, cr[2].ctypes.data, cr[5].ctypes.data,
                         cr[3].ctypes.data, cr[4].ctypes.data) if cr is not None
                        else (None, None, None, None, None, None))
                code = ref.ref_body_tile(fractals, k.count, g.origin_x, g.origin_y, g.pixel,
                                         g.width, g.height, batch.radius, *args, cores,
                                         out.ctypes.data)
            else:
                args = ((C.byref(cr[0]), cr[1].ctypes.data, cr[2].ctypes.data, cr[3].ctypes.data,
                         cr[4].ctypes.data) if cr is not None else (None,) * 5)
                code = ref.ref_plane_tile(fractals, k.count, g.origin_x, g.origin_y, g.pixel,
                                          g.width, g.height, *args, cores, out.ctypes.data)
            if code != 0:
                raise RuntimeError(ref.ref_bench_error().decode())
            outs[i] = out
        return outs

    def run_port():
        outs = {}
        for i, inp in inputs:
            g = inp["grid"]
            params = N.KernelParams(g.width, g.height, 0, 0.0, 0)
            out = np.zeros((g.height, g.width), np.float32)
            cr = inp["craters"]
            if sphere:
                pos = inp["positions"]
                orc.or_body_noise_grid(C.byref(k), pos.ctypes.data, pos.shape[0],
                                       C.byref(params), out.ctypes.data, cores)
                if cr is not None:
                    orc.or_shape_grid_body(C.byref(cr[0]), cr[1].ctypes.data, cr[2].ctypes.data,
                                           cr[3].ctypes.data, cr[4].ctypes.data,
                                           cr[5].ctypes.data, cr[6].ctypes.data, g.width,
                                           g.height, C.byref(params), out.ctypes.data, cores)
            else:
                orc.or_noise_grid(C.byref(k), inp["xs"].ctypes.data, inp["ys"].ctypes.data,
                                  g.width, g.height, C.byref(params), out.ctypes.data, cores)
                if cr is not None:
                    orc.or_shape_grid(C.byref(cr[0]), cr[1].ctypes.data, cr[2].ctypes.data,
                                      cr[3].ctypes.data, cr[4].ctypes.data, inp["xs"].ctypes.data,
                                      inp["ys"].ctypes.data, inp["lows"].ctypes.data, g.width,
                                      g.height, C.byref(params), out.ctypes.data, cores)
            outs[i] = out
        return outs

    def best(fn):
        fn()
        times = []
        for _ in range(repeats):
            t0 = time.perf_counter()
            outs = fn()
            times.append(time.perf_counter() - t0)
        return min(times), outs

    port_s, port_out = best(run_port)
    result = {"unit": "Gcells/s", "cores": cores, "cpu_model": model,
              "sample": f"{len(sample)} tiles ({cells} cells): noise then craters, inputs "
                        f"prepared beforehand; best of {repeats} after 1 warm-up",
              "port_value": round(cells / port_s / 1e9, 6)}
    if ref is not None:
        ref_s, ref_out = best(run_reference)
        diff = max(float(np.abs(ref_out[i] - port_out[i]).max()) for i in ref_out)
        result |= {"value": round(cells / ref_s / 1e9, 6), "kind": "reference",
                   "calibration_ratio": round(port_s / ref_s, 3),
                   "calibration_basis": "restatement time / reference time, same tiles, "
                                        "same cores",
                   "port_vs_reference_max_m": diff}
    else:
        result |= {"value": result["port_value"], "kind": "port"}
    result["_fp32"] = port_out
    return result


def bands_of(height: int, rows: int = 8) -> list[tuple[int, int]]:
    """Three bands of `rows` rows (top, middle, bottom) of a tile, or all of it when small."""
    if height <= 3 * rows:
        return [(0, height)]
    mid = height // 2 - rows // 2
    return [(0, rows), (mid, rows), (height - rows, rows)]


DIAMETER_BANDS = [("none in reach", 0.0, 1.0), ("< 2 km", 1.0, 2000.0),
                  ("2-5 km", 2000.0, 5000.0), ("5-10 km", 5000.0, 10000.0),
                  (">= 10 km", 10000.0, np.inf)]


def fp64_error(batch, got: np.ndarray, regular: list[int], sphere: bool, cores: int) -> dict:
    """Max/RMS |hip - FP64 oracle| over row bands (bands_of) of the `regular` tiles and of every
    tile a crater of 5 km or more reaches, reported per band of the largest crater in reach."""
    from oracle import oracle
    from terraingen_amd import native as N

    orc = oracle()
    largest = batch.largest_in_reach()
    big = [i for i in range(batch.tiles.size) if largest[i] >= 5000.0]
    chosen = sorted(set(regular) | set(big))
    k = batch.noise_kernel
    stats = {name: [0.0, 0.0, 0, 0] for name, _, _ in DIAMETER_BANDS}  # max, sum sq, cells, tiles
    for i in chosen:
        t = batch.tiles[i]
        w, h = int(t["width"]), int(t["height"])
        prepared = batch.tile_craters(i)
        mine = batch.tile_heights(got, i)
        if prepared is not None:
            kernel, records, data, foot = prepared
            if sphere:
                lists, entries = batch.band_lists(foot, float(t["origin_x"]),
                                                  float(t["origin_y"]), float(t["pixel"]), w, h)
                frames = frames64(foot[:, 0], foot[:, 1], batch.radius)
            else:
                grid = N.Grid(w, h, float(t["origin_x"]), float(t["origin_y"]),
                              float(t["pixel"]))
                lists, entries = N.bin_plane(grid, foot[:, 0], foot[:, 1], foot[:, 2], 16)
        worst, sq, n = 0.0, 0.0, 0
        for r0, rows in bands_of(h):
            band = N.Grid(w, rows, float(t["origin_x"]),
                          float(t["origin_y"]) - r0 * float(t["pixel"]), float(t["pixel"]))
            f64 = np.zeros((rows, w), np.float64)
            if sphere:
                pos = N.sphere_positions(band, batch.radius)
                orc.or_body_noise_grid_f64(C.byref(k), pos.ctypes.data, pos.shape[0],
                                           f64.ctypes.data, cores)
                if prepared is not None:
                    orc.or_shape_grid_body_f64(C.byref(kernel), records.ctypes.data,
                                               data.ctypes.data, lists.ctypes.data,
                                               entries.ctypes.data, frames.ctypes.data,
                                               pos.ctypes.data, w, rows, r0, f64.ctypes.data,
                                               cores)
            else:
                cx64, cy64 = (np.ascontiguousarray(a) for a in N.grid_centers(band))
                orc.or_noise_grid_f64(C.byref(k), cx64.ctypes.data, cy64.ctypes.data, w, rows,
                                      f64.ctypes.data, cores)
                if prepared is not None:
                    orc.or_shape_grid_f64(C.byref(kernel), records.ctypes.data,
                                          data.ctypes.data, lists.ctypes.data,
                                          entries.ctypes.data, cx64.ctypes.data, cy64.ctypes.data,
                                          w, rows, r0, f64.ctypes.data, cores)
            d = np.abs(mine[r0:r0 + rows].astype(np.float64) - f64)
            worst = max(worst, float(d.max()))
            sq += float((d ** 2).sum())
            n += d.size
        for name, lo, hi in DIAMETER_BANDS:
            if lo <= largest[i] < hi:
                st = stats[name]
                st[0] = max(st[0], worst)
                st[1] += sq
                st[2] += n
                st[3] += 1
    total = [max(v[0] for v in stats.values()), sum(v[1] for v in stats.values()),
             sum(v[2] for v in stats.values()), sum(v[3] for v in stats.values())]
    return {"max": float(f"{total[0]:.3g}"),
            "rms": float(f"{np.sqrt(total[1] / max(total[2], 1)):.3g}"),
            "cells": total[2], "tiles": total[3],
            "tiles_with_craters_of_5km_or_more": len(big),
            "sampled": "3 bands of 8 rows per tile (all rows of small tiles): every tile a "
                       "crater of 5 km or more reaches, and every 7th tile within the CPU budget",
            "by_largest_crater_in_reach": {
                name: {"max": float(f"{v[0]:.3g}"), "tiles": v[3], "cells": v[2]}
                for name, v in stats.items() if v[3] > 0}}


def frames64(lon, lat, radius):
    """MakeCraterFrame in FP64 (centre, east, north, radius; 10 values per crater)."""
    lam, phi = np.radians(lon), np.radians(lat)
    centre = radius * np.stack([np.cos(phi) * np.cos(lam), np.cos(phi) * np.sin(lam),
                                np.sin(phi)], axis=1)
    east = np.stack([-np.sin(lam), np.cos(lam), np.zeros_like(lam)], axis=1)
    north = np.stack([-np.sin(phi) * np.cos(lam), -np.sin(phi) * np.sin(lam), np.cos(phi)],
                     axis=1)
    return np.ascontiguousarray(np.concatenate([centre, east, north,
                                                np.full((lam.size, 1), radius)], axis=1))


def cpu_grid(terrain, rows: int, cores: int) -> dict:
    """As cpu_tiles for the single grid, on a band of its rows: the source, the craters and,
    when the terrain has them, the boulders."""
    from oracle import oracle
    from terraingen_amd import native as N

    orc = oracle()
    g = terrain.grid
    r0 = (g.height - rows) // 2
    sub_ys = np.ascontiguousarray(terrain.ys[r0:r0 + rows])
    kernel = N.noise_kernel(terrain.source_fractals) if terrain.source_fractals else None
    params = N.KernelParams(g.width, rows, 0, 0.0, r0)
    out = np.full((rows, g.width), terrain.elevation, np.float32)
    band = N.Grid(g.width, rows, g.origin_x, g.origin_y - r0 * g.pixel, g.pixel)
    t0 = time.perf_counter()
    if kernel is not None:
        orc.or_noise_grid(C.byref(kernel), terrain.xs.ctypes.data, sub_ys.ctypes.data, g.width,
                          rows, C.byref(params), out.ctypes.data, cores)
    orc.or_shape_grid(C.byref(terrain.kernel), terrain.records.ctypes.data,
                      terrain.data.ctypes.data, terrain.tiles.ctypes.data,
                      terrain.entries.ctypes.data, terrain.xs.ctypes.data, sub_ys.ctypes.data,
                      terrain.lows.ctypes.data, g.width, rows, C.byref(params), out.ctypes.data,
                      cores)
    if terrain.boulders is not None:
        orc.or_boulder_grid(terrain.boulders.ctypes.data, terrain.boulders.size, C.byref(band),
                            1, out.ctypes.data)
    elapsed = time.perf_counter() - t0
    cx64, cy64 = N.grid_centers(g)
    xs64 = np.ascontiguousarray(cx64)
    ys64 = np.ascontiguousarray(cy64[r0:r0 + rows])
    ref = np.full((rows, g.width), float(terrain.elevation), np.float64)
    if kernel is not None:
        orc.or_noise_grid_f64(C.byref(kernel), xs64.ctypes.data, ys64.ctypes.data, g.width,
                              rows, ref.ctypes.data, cores)
    orc.or_shape_grid_f64(C.byref(terrain.kernel), terrain.records.ctypes.data,
                          terrain.data.ctypes.data, terrain.tiles.ctypes.data,
                          terrain.entries.ctypes.data, xs64.ctypes.data, ys64.ctypes.data,
                          g.width, rows, r0, ref.ctypes.data, cores)
    if terrain.boulders is not None:
        orc.or_boulder_grid_f64(terrain.boulders.ctypes.data, terrain.boulders.size,
                                C.byref(band), 1, ref.ctypes.data)
    cells = rows * g.width
    return {"value": round(cells / elapsed / 1e9, 6), "unit": "Gcells/s", "cores": cores,
            "kind": "port",
            "cpu_model": host_cpu()[0],
            "sample": f"rows {r0}..{r0 + rows - 1} (x {g.width} cells) of the grid: source, "
                      f"craters{', boulders' if terrain.boulders is not None else ''}, once",
            "_r0": r0, "_fp32": out, "_fp64": ref}


def error_summary(pairs) -> tuple[dict, float]:
    """Max/RMS |hip - FP64 oracle| and max |hip - FP32 oracle| over (hip, fp32, fp64) triples."""
    worst = worst32 = 0.0
    sq = 0.0
    n = 0
    for got, f32, f64 in pairs:
        d = np.abs(got.astype(np.float64) - f64)
        worst = max(worst, float(d.max()))
        sq += float((d ** 2).sum())
        n += d.size
        worst32 = max(worst32, float(np.abs(got - f32).max()))
    return ({"max": float(f"{worst:.3g}"), "rms": float(f"{np.sqrt(sq / max(n, 1)):.3g}"),
             "cells": n}, float(f"{worst32:.3g}"))


# ---- workloads --------------------------------------------------------------------------------

def time_launches(args, N, dev, batch, out) -> tuple[float, float, float]:
    """W untimed graph launches, then K timed ones on the batch's stream with an event pair around
    each: (ms per step between the first and last event, wall ms per step, mean launch ms)."""
    stream = batch.stream
    for _ in range(args.warmup):
        batch.run(out, use_graph=True)
    batch.wait()
    events = [(N.Event(dev), N.Event(dev)) for _ in range(args.steps)]
    start, stop = N.Event(dev), N.Event(dev)
    t0 = time.perf_counter()
    start.record(stream)
    for a, b in events:
        a.record(stream)
        batch.run(out, use_graph=True)
        b.record(stream)
    stop.record(stream)
    batch.wait()
    wall_ms = (time.perf_counter() - t0) * 1e3 / args.steps
    step_ms = start.elapsed_ms(stop) / args.steps
    kernel_ms = float(np.mean([a.elapsed_ms(b) for a, b in events]))
    print(f"[bench] batched {step_ms:.3f} ms/step", file=sys.stderr, flush=True)
    return step_ms, wall_ms, kernel_ms


def cpu_and_error(args, batch, got: np.ndarray, tiles: np.ndarray, sphere: bool) -> dict:
    """The CPU baseline over every 7th tile within the CPU budget, the FP32 restatement's
    difference there, and the banded FP64 error (fp64_error)."""
    if args.no_cpu_baseline:
        return {}
    budget = args.cpu_cells
    sample, total = [], 0
    for i in range(0, tiles.size, 7):
        n = int(tiles["width"][i]) * int(tiles["height"][i])
        if total + n > budget:
            continue
        sample.append(i)
        total += n
    base = cpu_baseline(batch, sample, sphere)
    fp32 = base.pop("_fp32")
    diff32 = max(float(np.abs(batch.tile_heights(got, i) - fp32[i]).max()) for i in sample)
    print(f"[bench] cpu baseline {base}", file=sys.stderr, flush=True)
    err = fp64_error(batch, got, sample, sphere, base["cores"])
    print(f"[bench] fp64 error {err}", file=sys.stderr, flush=True)
    return {"cpu_baseline": base, "error_vs_fp64_oracle_m": err,
            "max_diff_vs_fp32_oracle_m": float(f"{diff32:.3g}"),
            "fp32_oracle_note": "the FP32 restatement of the reference's arithmetic; craters of "
                                "1 km and more are evaluated here in FP64, so the difference "
                                "there is the FP32 arithmetic's own rounding"}


def stage_and_tune(args, N, dev, batch, tiles, walks) -> tuple:
    """Stages `tiles` with host binning (16-cell bins, kept for the pair counts), then with device
    binning; autotunes the launch choices (crater walk, workgroups per CU, bins per queue take,
    culled-walk depth, bin side) unless told not to; restages at the chosen bin side and returns
    (output array, tuning, host bins, host entries, binning timings)."""
    batch.stage(tiles)
    host = {"host_bin_ms": round(batch.timings["bin_ms"], 1)}
    descs16, bins16, entries16 = batch.descs, batch.bins, batch.entries
    batch.stage(tiles, device_bins=True, bin_side=16)
    out = N.DeviceArray(dev, batch.cells * 4)
    tuning = None
    names = ["together", "per_lane", "culled"]
    if not args.no_autotune:
        timings = batch.autotune(out, walks=walks, depths=(32, 64, 128), bin_sides=(8, 16, 32))
        side, walk, g, chunk, depth = batch.tuned
        tuning = {"chosen": {"bin_side": side, "walk": names[walk], "groups_per_cu": g,
                             "chunk": chunk, "cull_depth": depth},
                  "ms": {f"{k[0]}/{names[k[1]]}/{k[2]}/{k[3]}/{k[4]}": round(ms, 3)
                         for k, ms in timings.items()},
                  "key": "bin side/walk/groups per CU/chunk/cull depth"}
        print(f"[bench] autotune {tuning}", file=sys.stderr, flush=True)
    # The chosen bin side's binning, timed: the host's windows, then the device's binning.
    batch.stage(tiles, device_bins=True, bin_side=getattr(batch, "tuned", (16,))[0])
    host |= {"windows_ms": round(batch.timings["bin_ms"], 1),
             "device_bin_ms": round(batch.timings["stage_ms"], 1)}
    return out, tuning, descs16, bins16, entries16, host


def bench_tiles(args, N, dev) -> dict:
    from terraingen_amd.pipeline import CraterLayerSpec, TileBatch, c4_tiles, crater_pair_cost

    t0 = time.perf_counter()
    batch = TileBatch(dev, [C2_FRACTAL], CraterLayerSpec(seed=4),
                      region=(0.0, 0.0, TILES_SIDE, TILES_SIDE))
    batch.build()
    tiles = c4_tiles(args.tiles, TILES_SIDE, seed=4)
    host_ms = (time.perf_counter() - t0) * 1e3
    out, tuning, descs16, bins16, entries16, binning = stage_and_tune(
        args, N, dev, batch, tiles,
        walks=(N.TG_WALK_TOGETHER, N.TG_WALK_PER_LANE, N.TG_WALK_CULLED))
    print(f"[bench] staged {tiles.size} tiles, {batch.cells} cells, {batch.records.size} craters,"
          f" {batch.bin_count} bins of {batch.bin_side} cells", file=sys.stderr, flush=True)
    stream = batch.stream
    step_ms, wall_ms, kernel_ms = time_launches(args, N, dev, batch, out)
    got = out.download(np.empty(batch.cells, np.float32))

    if args.copy_calibration_mib > 0:
        # A copy of known size on the same stream, for scaling the HBM byte counters.
        nbytes = args.copy_calibration_mib << 20
        src, dst = N.DeviceArray(dev, nbytes), N.DeviceArray(dev, nbytes)
        dev.call("tg_hip_memcpy_d2d_async", dev.handle, dst.ptr, src.ptr, nbytes, stream)
        batch.wait()
        del src, dst

    # The plain path over every tile: tg_hip_run per layer with host buffers.
    plain_s = 0.0
    plain_cells = 0
    worst_plain = 0.0
    for i in range(tiles.size if not args.no_plain else 0):
        inp = batch.plain_inputs(i)
        t1 = time.perf_counter()
        h = batch.plain_run(inp)
        plain_s += time.perf_counter() - t1
        plain_cells += h.size
        worst_plain = max(worst_plain, float(np.abs(h - batch.tile_heights(got, i)).max()))
    plain = plain_cells / plain_s / 1e9 if plain_cells else None
    if plain is not None:
        print(f"[bench] plain {plain:.3f} Gcells/s", file=sys.stderr, flush=True)

    # Algorithmic work of one launch, counted on the host from the bins and footprints.
    cells = batch.cells
    ops_noise = cells * C2_FRACTAL["octaves"] * OPS_PER_PERLIN_OCTAVE
    cost = crater_pair_cost(batch.records, OPS_PAIR, OPS_PER_HARMONIC, OPS_ELLIPSE)
    listed, inside, ops_craters = N.batch_pair_ops(descs16, bins16, entries16, batch.foot, cost)
    value = cells / (step_ms * 1e-3) / 1e9

    extra = cpu_and_error(args, batch, got, tiles, sphere=False)

    result = {
        "metric": METRIC, "value": round(value, 3), "unit": "Gcells/s",
        "ms_per_step": round(step_ms, 4), "wall_ms_per_step": round(wall_ms, 4),
        "end_to_end_gcells_per_s": round(
            cells / ((step_ms + binning["windows_ms"] + binning["device_bin_ms"]) * 1e-3) / 1e9,
            3),
        "end_to_end_note": "one launch plus its binning: the host's crater windows and the "
                           "device's binning of them, both timed at the chosen bin side",
        "dtype": "f32+u32",
        "config": {"workload": f"{tiles.size} mixed tiles, edges 2^k+1 (k 6-11), pixel "
                               f"0.25-8 m, over a {TILES_SIDE:.0f} m square; 8-octave Perlin fBm "
                               "(512 m, lac 2, gain 0.5, 20 m, seed 1) + shared catalogue of "
                               "power-law craters N(>D)=0.05 D^-2, D 1-200 m, seed 4",
                   "tiles": int(tiles.size), "cells": cells,
                   "craters": int(batch.records.size), "bins": int(batch.bin_count),
                   "bin_side": int(batch.bin_side), "entries_16_cell_bins": int(entries16.size),
                   "pairs_listed": listed, "pairs_in_reach": inside},
        "plain": {"value": round(plain, 4) if plain else None, "unit": "Gcells/s",
                  "path": "tg_hip_run per tile and layer, host buffers, synchronous",
                  "max_diff_vs_batched_m": worst_plain if plain else None},
        "batched_over_plain": round(value / plain, 2) if plain else None,
        "kernels_ms": {"batch_tiles": round(kernel_ms, 4)},
        "autotune": tuning,
        "host_ms": {k: round(v, 1) for k, v in batch.timings.items()
                    if k not in ("bin_ms", "stage_ms")} | binning | {
            "total_setup": round(host_ms, 1)},
        "roofline": roofline(ops_noise + ops_craters, kernel_ms, cells * 4.0) | {
            "kernel": "BatchTiles", "ops_noise": ops_noise, "ops_craters": ops_craters},
    }
    if tiles.size == 1024 and (counted := counted_traffic(
            kernel_ms, "batch_tiles_c4_1024_mixed_counters.md")) is not None:
        result["roofline"] |= counted
    result["roofline"]["valu_ceiling_measured"] = valu_ceiling(N, dev)
    result |= extra
    batch.close()
    return result


def bench_sphere(args, N, dev) -> dict:
    """The batched tiles on the Moon: tile origins uniform on a 1737.4 km sphere in FP64, each
    cell's noise at its body-fixed position, a global catalogue of 1e6 craters."""
    from terraingen_amd.pipeline import (CraterLayerSpec, SphereTileBatch, crater_pair_cost,
                                         placed_global_craters, sphere_min_diameter,
                                         sphere_tiles)

    t0 = time.perf_counter()
    craters = placed_global_craters(GLOBAL_CRATERS, MOON, max_diameter=GLOBAL_MAX_D, seed=4)
    place_ms = (time.perf_counter() - t0) * 1e3
    batch = SphereTileBatch(dev, [C2_FRACTAL], CraterLayerSpec(seed=4), region=None, radius=MOON,
                            craters_given=craters)
    batch.build()
    tiles = sphere_tiles(args.tiles, MOON, seed=4)
    host_ms = (time.perf_counter() - t0) * 1e3
    out, tuning, descs16, bins16, entries16, binning = stage_and_tune(
        args, N, dev, batch, tiles, walks=(N.TG_WALK_CULLED,))
    print(f"[bench] staged {tiles.size} sphere tiles, {batch.cells} cells, "
          f"{batch.records.size} craters, {batch.bin_count} bins of {batch.bin_side} cells",
          file=sys.stderr, flush=True)
    step_ms, wall_ms, kernel_ms = time_launches(args, N, dev, batch, out)
    got = out.download(np.empty(batch.cells, np.float32))
    if args.copy_calibration_mib > 0:
        # A copy of known size on the same stream, for scaling the HBM byte counters.
        nbytes = args.copy_calibration_mib << 20
        src, dst = N.DeviceArray(dev, nbytes), N.DeviceArray(dev, nbytes)
        dev.call("tg_hip_memcpy_d2d_async", dev.handle, dst.ptr, src.ptr, nbytes, batch.stream)
        batch.wait()
        del src, dst

    # The plain path over a sample of the tiles (every 4th): tg_hip_run per layer, host buffers.
    plain_s = 0.0
    plain_cells = 0
    worst_plain = 0.0
    for i in range(0, tiles.size if not args.no_plain else 0, 4):
        inp = batch.plain_inputs(i)
        t1 = time.perf_counter()
        h = batch.plain_run(inp)
        plain_s += time.perf_counter() - t1
        plain_cells += h.size
        worst_plain = max(worst_plain, float(np.abs(h - batch.tile_heights(got, i)).max()))
        del inp
    plain = plain_cells / plain_s / 1e9 if plain_cells else None
    if plain is not None:
        print(f"[bench] plain {plain:.3f} Gcells/s", file=sys.stderr, flush=True)

    cells = batch.cells
    ops_noise = cells * (C2_FRACTAL["octaves"] * OPS_PER_PERLIN3_OCTAVE + OPS_BODY_POSITION)
    cost = crater_pair_cost(batch.records, OPS_PAIR + OPS_BODY_PAIR_EXTRA, OPS_PER_HARMONIC,
                            OPS_ELLIPSE)
    listed, inside, ops_craters = N.sphere_pair_ops(descs16, bins16, entries16, batch.foot,
                                                    cost, MOON)
    value = cells / (step_ms * 1e-3) / 1e9

    extra = cpu_and_error(args, batch, got, tiles, sphere=True)

    d_min = sphere_min_diameter(GLOBAL_CRATERS, MOON, 0.05, 2.0, GLOBAL_MAX_D)
    result = {
        "metric": METRIC, "value": round(value, 3), "unit": "Gcells/s",
        "ms_per_step": round(step_ms, 4), "wall_ms_per_step": round(wall_ms, 4),
        "end_to_end_gcells_per_s": round(
            cells / ((step_ms + binning["windows_ms"] + binning["device_bin_ms"]) * 1e-3) / 1e9,
            3),
        "end_to_end_note": "one launch plus its binning: the host's crater windows and the "
                           "device's binning of them, both timed at the chosen bin side",
        "dtype": "f32+f64+u32",
        "config": {"workload": f"{tiles.size} mixed tiles on a 1737.4 km sphere, edges 2^k+1 "
                               "(k 6-11), pixel 0.25-8 m, north-west corners uniform on the "
                               "sphere, FP64 body-fixed cell positions; 8-octave 3D Perlin fBm "
                               "(512 m, lac 2, gain 0.5, 20 m, seed 1) + a global catalogue of "
                               f"{craters.size} power-law craters N(>D)=0.05 D^-2, D "
                               f"{d_min:.0f}-{GLOBAL_MAX_D:.0f} m, seed 4, placed as the "
                               "reference places them (cube-sphere cells over the whole Moon)",
                   "tiles": int(tiles.size), "cells": cells,
                   "craters": int(batch.records.size), "bins": int(batch.bin_count),
                   "bin_side": int(batch.bin_side), "entries_16_cell_bins": int(entries16.size),
                   "pairs_listed": listed, "pairs_in_reach": inside},
        "plain": {"value": round(plain, 4) if plain else None, "unit": "Gcells/s",
                  "path": "tg_hip_run per tile and layer (body noise, crater body branch), host "
                          "buffers, synchronous, over every 4th tile",
                  "max_diff_vs_batched_m": worst_plain if plain else None},
        "batched_over_plain": round(value / plain, 2) if plain else None,
        "kernels_ms": {"batch_body_tiles": round(kernel_ms, 4)},
        "autotune": tuning,
        "host_ms": {k: round(v, 1) for k, v in batch.timings.items()
                    if k not in ("bin_ms", "stage_ms")} | binning | {
            "total_setup": round(host_ms, 1)},
        "roofline": roofline(ops_noise + ops_craters, kernel_ms, cells * 4.0) | {
            "kernel": "BatchBodyTiles", "ops_noise": ops_noise, "ops_craters": ops_craters,
            "note": "achieved counts the reference's algorithm (eight corner hashes per 3D "
                    "octave); the kernel hashes each lattice point once per wave where a "
                    "square's corners fit a 4x4x4 block, so it issues fewer instructions than "
                    "that count. executed_frac is what the hardware issued against the issue "
                    "rate the same counter reads for the VALU probe"},
    }
    if tiles.size == 1024 and (counted := counted_traffic(
            kernel_ms, "batch_body_tiles_c4_sphere_1024_counters.md")) is not None:
        result["roofline"] |= counted
    result["roofline"]["valu_ceiling_measured"] = valu_ceiling(N, dev)
    result |= extra
    batch.close()
    return result


def bench_serving(args, N, dev) -> dict:
    from terraingen_amd.pipeline import CraterLayerSpec, SphereTileBatch, TileBatch, TileServer
    from terraingen_amd.serving import CameraStream, GreatCircleStream, serve, served_boulders

    sphere = args.workload == "sphere"
    stream = (GreatCircleStream(duration=args.serve_seconds, radius=MOON) if sphere
              else CameraStream(duration=args.serve_seconds))
    requests = stream.requests()
    t0 = time.perf_counter()
    # The crater population of the dense C3 field: slope -2.2, D 0.5-300 m, complex from 50 m
    # with central peaks, wall terraces from 150 m (full at 300 m). Prepared at the finest cell
    # served (a level-10 tile of 1025 cells: 0.125 m), so no crater is dropped.
    spec = CraterLayerSpec(sfd_coefficient=0.0156, sfd_exponent=2.2, min_diameter=0.5,
                           max_diameter=300.0, complex_=1, transition_diameter=50.0, peaks=1,
                           terraces=1, terrace_onset=150.0, terrace_full=300.0, seed=5)
    finest = stream.finest_side / (max(stream.edges) - 1)
    # The boulder field of C3 (N(>D) = 3.8e-3 D^-3 per m^2, D 0.2-3 m, seed 5) as the
    # reference's BoulderLayer lists it for each served tile.
    t_b = time.perf_counter()
    boulders, boulder_stats = served_boulders(stream, requests, 3.8e-3, 3.0, 0.2, 3.0, seed=5)
    boulder_place_ms = (time.perf_counter() - t_b) * 1e3
    if sphere:
        craters = stream.placed_swath(1000.0, 0.0156, 2.2, 0.5, 300.0,
                                      margin_radii=N.margin_radii(spec.models(), spec.apply()),
                                      seed=5)
        place_ms = (time.perf_counter() - t0) * 1e3 - boulder_place_ms
        swath = SphereTileBatch(dev, [C2_FRACTAL], spec, region=None, finest_pixel=finest,
                                radius=MOON, craters_given=craters, boulders=boulders)
    else:
        swath = TileBatch(dev, [C2_FRACTAL], spec,
                          region=(-2000.0, -1000.0, stream.path_length + 2000.0, 1000.0),
                          finest_pixel=finest, boulders=boulders)
    swath.build()
    server = TileServer(swath, slots=4, max_batch_tiles=32, max_batch_cells=1 << 24,
                        pool_bytes=64 << 30, bin_threads=8)
    setup_ms = (time.perf_counter() - t0) * 1e3
    # Warm the server's slots before the timed stream.
    for i in range(8):
        r = requests[i]
        server.wait(server.submit(float(r["origin_x"]), float(r["origin_y"]),
                                  float(r["pixel"]), int(r["width"]), int(r["height"])))
    print(f"[bench] serving {requests.size} requests over {stream.duration} s, "
          f"{swath.records.size} craters, {swath.boulder_records.size} boulders",
          file=sys.stderr, flush=True)
    # Every 97th served tile is kept and checked against the same tiles in one batched launch.
    kept = {}

    def keep(i, heights):
        if i % 97 == 0:
            kept[i] = heights.copy()

    rec = serve(server, requests, keep=keep)
    stats = server.stats()
    server.close()
    sample = np.array(sorted(kept))
    check = np.zeros(sample.size, N.TILE_DTYPE)
    for name in ("origin_x", "origin_y", "pixel", "width", "height"):
        check[name] = requests[name][sample]
    swath.stage(check)
    out = N.DeviceArray(dev, swath.cells * 4)
    swath.run(out, use_graph=False)
    swath.wait()
    batched = out.download(np.empty(swath.cells, np.float32))
    served_diff = max((float(np.abs(kept[i] - swath.tile_heights(batched, k)).max())
                       for k, i in enumerate(sample)), default=0.0)
    del out
    swath.close()
    lat_ms = rec["latency_s"] * 1e3
    is513 = requests["width"] == 513
    cells = float((requests["width"].astype(np.float64) * requests["height"]).sum())
    return {
        "stream": ("a 50 km great-circle track on the 1737.4 km sphere at 200 m, lon/lat "
                   "quadtree tiles, 3D body noise" if sphere
                   else "a 50 km straight track on a plane at 200 m") +
                  "; a 2 km swath of C3's craters (N(>D)=0.0156 D^-2.2, D 0.5-300 m, complex "
                  "from 50 m with peaks, terraces from 150 m" +
                  (", placed as the reference places them on the sphere)" if sphere else ")") +
                  "; C3's boulders (N(>D)=3.8e-3 D^-3, D 0.2-3 m) as the reference's "
                  "BoulderLayer lists them for each tile at its spacing, drawn after the craters",
        "requests": int(requests.size), "rate_per_s": stream.rate,
        "duration_s": stream.duration,
        "latency_513_ms": {"p50": round(percentile(lat_ms[is513], 50), 3),
                           "p99": round(percentile(lat_ms[is513], 99), 3),
                           "count": int(is513.sum())},
        "latency_all_ms": {"p50": round(percentile(lat_ms, 50), 3),
                           "p99": round(percentile(lat_ms, 99), 3),
                           "max": round(float(lat_ms.max()), 3)},
        "served_gcells_per_s": round(cells / rec["wall_s"] / 1e9, 3),
        "batches": int(stats["batches"]), "max_in_flight": int(stats["max_in_flight"]),
        "mean_batch_tiles": round(float(rec["batch_tiles"].mean()), 2),
        "pool_peak_bytes": int(stats["pool_peak_bytes"]),
        "levels": sorted(int(v) for v in np.unique(requests["level"])),
        "craters": int(swath.records.size), "setup_ms": round(setup_ms, 1),
        "placement_ms": round(place_ms, 1) if sphere else None,
        "boulders": int(swath.boulder_records.size), "boulder_placement_ms":
            round(boulder_place_ms, 1), "boulders_by_spacing": boulder_stats,
        "served_vs_batched": {"tiles": int(sample.size), "max_diff_m": served_diff,
                              "note": "every 97th served tile against the same tiles in one "
                                      "batched launch (which the tests hold bit-equal to the "
                                      "plain per-tile kernels)"},
    }


def grid_as_one_tile(args, N, dev, terrain, spec, strip_heights: np.ndarray) -> dict:
    """The grid through the batched kernel as one tile: the same source noise, crater list
    (generated over the grid's box at its spacing) and boulder list as the strip passes, staged as
    one tile of 16 x 16 bins and run as a captured graph on the batch's stream; checked bit for bit
    against the strip passes' heights."""
    from terraingen_amd.pipeline import TileBatch

    g = terrain.grid
    region = (g.origin_x, g.origin_y - g.height * g.pixel, g.origin_x + g.width * g.pixel,
              g.origin_y)
    t0 = time.perf_counter()
    batch = TileBatch(dev, list(terrain.source_fractals or []), spec, region=region,
                      finest_pixel=g.pixel, elevation=terrain.elevation,
                      boulders=terrain.boulders)
    batch.build()
    one = np.zeros(1, N.TILE_DTYPE)
    one[0] = (g.origin_x, g.origin_y, g.pixel, g.width, g.height)
    batch.stage(one)
    setup_ms = (time.perf_counter() - t0) * 1e3
    out = N.DeviceArray(dev, batch.cells * 4)
    tuned = None
    if not args.no_autotune:
        # Boulders are binned into 16-cell bins, so only a grid without them tries other sides.
        batch.autotune(out, depths=(32, 64, 128),
                       bin_sides=None if terrain.boulders is not None else (8, 16, 32))
        tuned = dict(zip(("bin_side", "walk", "groups_per_cu", "chunk", "cull_depth"),
                         batch.tuned))
    for _ in range(max(args.warmup, 1)):
        batch.run(out)
    batch.wait()
    start, stop = N.Event(dev), N.Event(dev)
    t1 = time.perf_counter()
    start.record(batch.stream)
    for _ in range(args.steps):
        batch.run(out)
    stop.record(batch.stream)
    batch.wait()
    wall = time.perf_counter() - t1
    ms = start.elapsed_ms(stop) / args.steps
    got = out.download(np.empty((g.height, g.width), np.float32))
    same = bool(np.array_equal(got.view(np.uint32), strip_heights.view(np.uint32)))
    diff = float(np.abs(got - strip_heights).max())
    strip_ms = terrain.timings.get("_strip_ms")
    boulders = terrain.boulders is not None
    # The full-evaluation variant for complex craters, peaks or terraces; else the all-fast one.
    general = bool(spec.complex_ or spec.peaks or spec.terraces or spec.profile != 1)
    kind = "culled, " + ("general" if general else "all-fast") + (
        ", boulders" if boulders else "")
    result = {"kernel": f"BatchTiles<{kind}>", "ms": round(ms, 4),
              "bins": int(batch.bin_count), "craters": int(batch.records.size),
              "boulders": int(batch.boulder_records.size) if boulders else 0,
              "setup_ms": round(setup_ms, 1), "autotune": tuned,
              "bit_equal_to_strip_passes": same, "max_diff_vs_strip_m": diff,
              "strip": {"path": "noise fill, then CraterStrip" + (
                            " with the boulders in the same pass (tg_hip_crater_boulder_launch)"
                            if boulders else "") + ", one thread per cell walking its 16 x 16 "
                        "tile's lists",
                        "ms": round(strip_ms, 4) if strip_ms else None,
                        "value": round(g.width * g.height / (strip_ms * 1e-3) / 1e9, 3)
                        if strip_ms else None, "unit": "Gcells/s"},
              "_ms": ms, "_wall_s": wall}
    del out
    batch.close()
    return result


def upsample_ops(method: str) -> int:
    """Algorithmic FP64 operations of one upsampled cell (the counts above)."""
    if method == "nearest":
        return OPS_UPSAMPLE_POSITION + 2 + OPS_UPSAMPLE_TAP
    if method == "bilinear":
        return OPS_UPSAMPLE_POSITION + 4 + 2 + 4 * OPS_UPSAMPLE_TAP
    taps = 6 if method == "lanczos3" else 4
    return (OPS_UPSAMPLE_POSITION + 4 + 2 * taps * OPS_UPSAMPLE_WEIGHT[method]
            + taps * taps * OPS_UPSAMPLE_TAP)


def bench_sources(args, N, dev) -> dict:
    """A SyntheticSource with a pixel size: C2's fractal on posts 2 m apart, upsampled onto the
    4097 x 4097 grid at 0.5 m by the reference's default upsampler (cubic). One step is the noise
    at the posts the grid needs, then UpsampleGrid over the grid, inputs and outputs in HBM. The
    other kernels are timed on the same grid; the CPU baseline is the reference's own
    SyntheticSource::Fill on a band of rows."""
    from terraingen_amd.sources import METHODS, SyntheticPosts

    n = args.size if args.size else 4097
    pixel, post = 0.5, 2.0
    grid = N.north_up_grid(n, n, pixel)
    fractal = [C2_FRACTAL]
    sv = C.c_void_p()
    dev.call("tg_hip_stream_create", dev.handle, C.byref(sv))
    stream = sv.value
    out = N.DeviceArray(dev, n * n * 4)
    names = ["cubic", "nearest", "bilinear", "cubic_bspline", "lanczos3"]
    sources = {m: SyntheticPosts(dev, fractal, 0.0, post, METHODS[m]) for m in names}
    fills = {m: sources[m].prepare(grid) for m in names}

    def timed(fn, steps):
        for _ in range(args.warmup):
            fn()
        dev.synchronize(stream)
        a, b = N.Event(dev), N.Event(dev)
        t0 = time.perf_counter()
        a.record(stream)
        for _ in range(steps):
            fn()
        b.record(stream)
        dev.synchronize(stream)
        return a.elapsed_ms(b) / steps, (time.perf_counter() - t0) / steps

    main = fills["cubic"]
    step_ms, wall_s = timed(lambda: main.launch(out, stream), args.steps)
    posts_ms, _ = timed(lambda: main.launch_posts(stream), args.steps)
    per_method = {}
    for m in names:
        fills[m].launch_posts(stream)
        ms, _ = timed(lambda f=fills[m]: f.launch_upsample(out, stream), args.steps)
        per_method[m] = ms
    main.launch(out, stream)
    dev.synchronize(stream)
    got = out.download(np.empty((n, n), np.float32))
    cells = n * n
    _, _, pw, ph = main.window
    ups_ms = per_method["cubic"]
    ops = cells * upsample_ops("cubic")
    achieved = ops / (ups_ms * 1e-3)
    bytes_moved = cells * 4.0 + pw * ph * 4.0
    result = {
        "metric": METRIC, "value": round(cells / (step_ms * 1e-3) / 1e9, 3), "unit": "Gcells/s",
        "ms_per_step": round(step_ms, 4), "wall_ms_per_step": round(wall_s * 1e3, 4),
        "dtype": "f64 (f32 heights)",
        "config": {"workload": f"synthetic source with 2 m posts on a {n}x{n} grid @0.5 m: "
                               "8-octave Perlin fBm (512 m, lac 2, gain 0.5, 20 m, seed 1) at "
                               "the posts, cubic upsampling (the reference's DefaultUpsampler)",
                   "cells": cells, "posts": pw * ph, "post_spacing_m": post},
        "kernels_ms": {"NoiseFillStrip (posts)": round(posts_ms, 4),
                       "UpsampleGrid (cubic)": round(ups_ms, 4)},
        "upsample_ms": {m: round(v, 4) for m, v in per_method.items()},
        "upsample_gcells_per_s": {m: round(cells / (v * 1e-3) / 1e9, 2)
                                  for m, v in per_method.items()},
        "roofline": {"bound": "valu", "achieved": round(achieved / 1e12, 3),
                     "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": round(achieved / (FP64_PEAK_TFLOPS * 1e12), 4),
                     "peak_basis": "MI355X published FP64 vector peak (78.6 TFLOP/s, FMA as 2 "
                                   "flops a lane); achieved counts one per algorithmic FP64 "
                                   "operation",
                     "traffic": None, "kernel": "UpsampleGrid",
                     "ops_per_cell": upsample_ops("cubic"),
                     "hbm_algorithmic_GBs": round(bytes_moved / (ups_ms * 1e-3) / 1e9, 1),
                     "hbm_frac": round(bytes_moved / (ups_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)},
    }
    if not args.no_cpu_baseline:
        result |= cpu_sources(grid, fractal, post, got, args.cpu_rows)
    for src in sources.values():
        src.close()
    return result


def cpu_sources(grid, fractal, post: float, got: np.ndarray, rows: int) -> dict:
    """The reference's SyntheticSource::Fill with its KernelUpsampler(cubic) (oracle/_ref) on the
    grid or a band of its rows, on this host's cores: one warm-up, best of 3. Its heights are
    compared with the device's bit for bit. Reports no baseline where oracle/_ref was not built."""
    from oracle import reference
    from terraingen_amd import native as N

    model, cores = host_cpu()
    rows = rows if 0 < rows < grid.height else grid.height
    r0 = (grid.height - rows) // 2
    band = N.Grid(grid.width, rows, grid.origin_x, grid.origin_y - r0 * grid.pixel, grid.pixel)
    want = got[r0:r0 + rows]
    ref = reference()
    if ref is None:
        return {"cpu_baseline": {"value": None, "note": "oracle/_ref not built"}}
    out = np.empty((rows, grid.width), np.float32)
    kernel = N.noise_kernel(fractal)
    ref.ref_sources_threads(cores)

    def run():
        code = ref.ref_synthetic_grid(C.cast(C.byref(kernel.fractals), C.c_void_p), kernel.count,
                                      0.0, post, N.TG_RESAMPLE_CUBIC, band.width, band.height,
                                      band.origin_x, band.origin_y, band.pixel, 0.0,
                                      out.ctypes.data)
        if code != 0:
            raise RuntimeError(ref.ref_sources_error().decode())

    run()
    times = []
    for _ in range(3):
        t0 = time.perf_counter()
        run()
        times.append(time.perf_counter() - t0)
    cells = rows * grid.width
    same = bool(np.array_equal(out.view(np.uint32), want.view(np.uint32)))
    return {"cpu_baseline": {"value": round(cells / min(times) / 1e9, 6), "unit": "Gcells/s",
                             "cores": cores, "kind": "reference", "cpu_model": model,
                             "sample": f"rows {r0}..{r0 + rows - 1} (x {grid.width} cells) of "
                                       "the grid: the reference's SyntheticSource with 2 m "
                                       "posts, KernelUpsampler(cubic); best of 3 after 1 "
                                       "warm-up"},
            "bit_equal_to_reference_band": same,
            "max_diff_vs_reference_m": float(np.abs(out.astype(np.float64) - want).max())}


def bench_grid(args, N, dev) -> dict:
    from terraingen_amd.pipeline import (CraterLayerSpec, PlaneTerrain, crater_pair_cost,
                                         plane_batch_view)

    c3 = args.workload == "c3"
    n = args.size if args.size else (8193 if c3 else 4097)
    pixel = 0.25 if c3 else 0.5
    grid = N.north_up_grid(n, n, pixel)
    if c3:
        spec = CraterLayerSpec(sfd_coefficient=0.0156, sfd_exponent=2.2, min_diameter=0.5,
                               max_diameter=300.0, complex_=1, transition_diameter=50.0,
                               peaks=1, terraces=1, terrace_onset=150.0, terrace_full=300.0,
                               seed=2)
        terrain = PlaneTerrain(dev, grid, [], 0.0, spec)
    else:
        terrain = PlaneTerrain(dev, grid, [C2_FRACTAL], 0.0, CraterLayerSpec(seed=1))
    t0 = time.perf_counter()
    terrain.prepare()
    if c3:
        # The boulder layer as the reference's BoulderLayer lists it for the grid (its box grown
        # by half a sample, at the grid's spacing): C3's sizes, uniform background abundance,
        # drawn as dome relief at the grid spacing.
        half = 0.5 * n * pixel + 0.5 * pixel
        t_b = time.perf_counter()
        placed = N.place_boulders((-half, -half, half, half), pixel, 3.8e-3, 3.0, 0.2, 3.0, seed=2)
        terrain.timings["boulder_placement_ms"] = (time.perf_counter() - t_b) * 1e3
        terrain.boulders = N.relief_boulders(placed)
        terrain.all_boulders = terrain.boulders.size
        terrain._prepare_boulders()
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
    terrain.timings["_strip_ms"] = step_ms

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
    boulder_ms = time_kernel(lambda: terrain.launch_boulders(heights, stream)) if c3 else 0.0
    fused_ms = (time_kernel(lambda: terrain.launch_craters_and_boulders(heights, stream))
                if c3 else 0.0)
    terrain.launch(heights, stream)
    dev.synchronize(stream)
    got = heights.download(np.empty((n, n), np.float32))
    cells = n * n
    noise_ops = 0 if c3 else cells * C2_FRACTAL["octaves"] * OPS_PER_PERLIN_OCTAVE
    descs, bins, entries = plane_batch_view(grid, terrain.tiles, terrain.entries)
    cost = crater_pair_cost(terrain.records, OPS_PAIR, OPS_PER_HARMONIC, OPS_ELLIPSE)
    listed, inside, crater_ops = N.batch_pair_ops(descs, bins, entries, terrain.foot, cost)
    kernels = {"NoiseFillStrip" if not c3 else "FillValue": (noise_ms, noise_ops),
               "CraterStrip": (crater_ms, crater_ops)}
    config = {"cells": cells, "craters": int(terrain.crater_count), "pairs_listed": listed,
              "pairs_in_reach": inside}
    if c3:
        bdescs, bbins, bentries = plane_batch_view(grid, terrain.boulder_tiles,
                                                   terrain.boulder_entries)
        bcost = np.full(terrain.boulder_count, float(OPS_BOULDER_PAIR))
        _, b_inside, boulder_ops = N.batch_pair_ops(bdescs, bbins, bentries, terrain.boulder_foot,
                                                    bcost)
        # The step runs the craters and boulders in one pass over the grid; the separate
        # passes are timed for comparison.
        kernels = {"FillValue": (noise_ms, 0),
                   "CraterStrip<boulders>": (fused_ms, crater_ops + boulder_ops)}
        config |= {"boulders": int(terrain.all_boulders),
                   "boulders_drawn": int(terrain.boulder_count),
                   "boulder_pairs_in_reach": b_inside,
                   "separate_passes_ms": {"CraterStrip": round(crater_ms, 4),
                                          "BoulderStrip": round(boulder_ms, 4)}}
        config["workload"] = (f"flat source {n}x{n} @0.25 m + power-law craters N(>D)=0.0156 "
                              "D^-2.2, D 0.5-300 m, complex from 50 m with central peaks, "
                              "terraces from 150 m (full at 300 m), seed 2 + boulders "
                              "N(>D)=3.8e-3 D^-3, D 0.2-3 m, seed 2, placed as the reference's "
                              "BoulderLayer places them (uniform background, no clusters), "
                              "dome relief, drawn in the crater pass")
    else:
        config["workload"] = (f"synthetic source {n}x{n} @0.5 m: 8-octave Perlin fBm (512 m, "
                              "lac 2, gain 0.5, 20 m, seed 1) + power-law craters "
                              "N(>D)=0.05 D^-2, D 1-200 m, seed 1")
    # The value: the grid as one tile of the batched kernel, noise, craters and boulders in one
    # launch, bit-equal to the strip passes.
    batched = grid_as_one_tile(args, N, dev, terrain, spec if c3 else CraterLayerSpec(seed=1),
                               got)
    step_ms, wall = batched.pop("_ms"), batched.pop("_wall_s")
    kernels = {batched["kernel"]: (step_ms, noise_ops + crater_ops +
                                   (boulder_ops if c3 else 0))}
    dominant = max(kernels, key=lambda k: kernels[k][0])
    dom_ms, dom_ops = kernels[dominant]
    result = {
        "metric": METRIC, "value": round(cells / (step_ms * 1e-3) / 1e9, 3), "unit": "Gcells/s",
        "ms_per_step": round(step_ms, 4), "wall_ms_per_step": round(wall * 1e3 / args.steps, 4),
        "dtype": "f32+u32",
        "config": {"workload": config.pop("workload")} | config,
        "kernels_ms": {k: round(v[0], 4) for k, v in kernels.items()},
        "host_ms": {k: round(v, 1) for k, v in terrain.timings.items()
                    if not k.startswith("_")} | {"total_prepare": round(host_ms, 1)},
        "roofline": roofline(dom_ops, dom_ms, cells * 4.0) | {"kernel": dominant},
    }
    if batched is not None:
        result["plain"] = batched.pop("strip")
        result["batched"] = batched
    if not args.no_cpu_baseline:
        base = cpu_grid(terrain, 16, os.cpu_count() or 1)
        r0, fp32, fp64 = base.pop("_r0"), base.pop("_fp32"), base.pop("_fp64")
        err, diff32 = error_summary([(got[r0:r0 + fp32.shape[0]], fp32, fp64)])
        result["cpu_baseline"] = base
        result["error_vs_fp64_oracle_m"] = err
        result["max_diff_vs_fp32_oracle_m"] = diff32
    terrain.close()
    return result


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", choices=("sphere", "tiles", "grid", "c3", "sources"),
                    default="sphere")
    ap.add_argument("--tiles", type=int, default=1024)
    ap.add_argument("--size", type=int, default=0,
                    help="grid workload edge (0: 4097 for grid, 8193 for c3)")
    ap.add_argument("--serve-seconds", type=float, default=10.0,
                    help="length of the served stream (0: skip it)")
    ap.add_argument("--cpu-cells", type=int, default=6_000_000,
                    help="cells of the tiles workload timed on the host (CPU baseline)")
    ap.add_argument("--cpu-rows", type=int, default=0,
                    help="rows of the sources workload timed on the host (0: the whole grid)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-plain", action="store_true", help="skip the plain per-tile path")
    ap.add_argument("--no-autotune", action="store_true",
                    help="launch with the default crater walk and workgroups per CU")
    ap.add_argument("--copy-calibration-mib", type=int, default=0,
                    help="also run one device copy of this size (for counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU: tiles would shard across GPUs "
                                   "with no exchange between them, but runs on more than one "
                                   "GPU are not part of it"}), flush=True)
        return 2

    from terraingen_amd import native as N

    dev = N.HipDevice(0)
    if args.workload == "tiles":
        body = bench_tiles(args, N, dev)
    elif args.workload == "sphere":
        body = bench_sphere(args, N, dev)
    elif args.workload == "sources":
        body = bench_sources(args, N, dev)
    else:
        body = bench_grid(args, N, dev)
    if args.workload in ("tiles", "sphere") and args.serve_seconds > 0:
        body["serving"] = bench_serving(args, N, dev)
    result = {"metric": body.pop("metric"), "value": body.pop("value"), "unit": body.pop("unit"),
              "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
              "ms_per_step": body.pop("ms_per_step"), "higher_is_better": True,
              "scaling": "weak", "vs_baseline": None, "dtype": body.pop("dtype"),
              "data": "synthetic", "config": body.pop("config")} | body
    print(json.dumps(result), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
