# SPDX-License-Identifier: BSD-3-Clause
"""Measurement aids (no counterpart in the reference).

``python -m terraingen_amd.probes stores`` writes 64 tiles of 2049 x 2049 floats (rows off the 64 B
sectors, as the batched kernels' tiles of 2^k + 1 cells) and of 2048 x 2048 (rows on them) in the
three store shapes of ``tg_hip_store_probe``, then copies 1 GiB device to device, printing one JSON
line of launch times. Run under ``rocprofv3 --pmc WRITE_SIZE`` it gives the HBM bytes each shape
writes, the copy's reading calibrating the counter.
"""
from __future__ import annotations

import ctypes as C
import json
import sys

from . import native as N

SHAPES = {0: "8x8 squares", 1: "16x4 strips", 2: "64-float rows"}


def stores(tiles: int = 64) -> dict:
    dev = N.HipDevice(0)
    cap = tiles * 2049 * 2049
    out = N.DeviceArray(dev, cap * 4)
    runs = []
    for width in (2049, 2048):
        for shape in SHAPES:
            ms = C.c_float()
            dev.call("tg_hip_store_probe", dev.handle, shape, width, tiles, out.ptr, cap,
                     C.byref(ms))
            runs.append({"width": width, "shape": SHAPES[shape], "ms": round(ms.value, 4),
                         "bytes": tiles * width * width * 4})
    nbytes = 1 << 30
    src, dst = N.DeviceArray(dev, nbytes), N.DeviceArray(dev, nbytes)
    src.zero()
    dev.call("tg_hip_memcpy_d2d_async", dev.handle, dst.ptr, src.ptr, nbytes, None)
    dev.synchronize(None)
    return {"probe": "stores", "tiles": tiles, "runs": runs, "calibration_copy_bytes": nbytes}


def main(argv: list[str]) -> int:
    if argv[:1] != ["stores"]:
        print("usage: python -m terraingen_amd.probes stores", file=sys.stderr)
        return 2
    print(json.dumps(stores()), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main(sys.argv[1:]))
