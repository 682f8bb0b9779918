# SPDX-License-Identifier: BSD-3-Clause
"""Writes the output-file fixtures with the reference's own writers (oracle/_ref, built from its
sources by oracle/Makefile): a small LZW GeoTIFF on a sphere with a no-data value, and a crater
list CSV with attributes. tests/test_outputs.py compares this repository's files with them where
oracle/_ref is not built. Run from the repository root: python tests/golden/make_output_fixtures.py
"""
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "tests"))

from oracle import ref_write_crater_csv, ref_write_geotiff, reference  # noqa: E402
from output_cases import fixture_craters, fixture_raster  # noqa: E402


def main() -> None:
    lib = reference()
    if lib is None:
        raise SystemExit("oracle/_ref is not built (make -C oracle ref)")
    out = ROOT / "tests" / "golden"
    values, transform, nodata, keys = fixture_raster()
    ref_write_geotiff(lib, out / "output_sphere_lzw.tif", values, transform, nodata=nodata,
                      geokeys=keys, compression=1, tile_size=16)
    craters, manual, replace, attributes = fixture_craters()
    ref_write_crater_csv(lib, out / "output_craters.csv", craters, manual, replace, attributes)


if __name__ == "__main__":
    main()
