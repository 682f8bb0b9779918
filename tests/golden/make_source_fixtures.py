# SPDX-License-Identifier: BSD-3-Clause
"""Writes tests/golden/sources_synthetic.npz and sources_geotiff.npz: the reference's own heights
(oracle/_ref, built from its sources by oracle/Makefile) for each case of tests/source_cases.py,
a SyntheticSource with a pixel size or a GeoTiffSource of a file tests/tiff_craft.py writes,
upsampled by a KernelUpsampler, evaluated on the CPU. Run from the repository root:
python tests/golden/make_source_fixtures.py"""
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "tests"))

from oracle import reference  # noqa: E402
from source_cases import (CASES, GEOTIFF_CASES, reference_geotiff_heights,  # noqa: E402
                          reference_heights)


def main() -> None:
    import tempfile

    lib = reference()
    if lib is None:
        raise SystemExit("oracle/_ref is not built (make -C oracle ref)")
    out = {name: reference_heights(lib, name) for name in CASES}
    np.savez_compressed(ROOT / "tests" / "golden" / "sources_synthetic.npz", **out)
    for name, h in out.items():
        print(name, h.shape, float(h.min()), float(h.max()))
    # GeoTiffSource over the files tests/tiff_craft.py writes (the same bytes every time).
    with tempfile.TemporaryDirectory() as tmp:
        out = {name: reference_geotiff_heights(lib, name, tmp) for name in GEOTIFF_CASES}
    np.savez_compressed(ROOT / "tests" / "golden" / "sources_geotiff.npz", **out)
    for name, h in out.items():
        print(name, h.shape, int(np.isnan(h).sum()), float(np.nanmin(h)), float(np.nanmax(h)))


if __name__ == "__main__":
    main()
