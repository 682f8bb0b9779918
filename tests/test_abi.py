# SPDX-License-Identifier: BSD-3-Clause
"""The C ABI library loads and exports every entry point include/*.h declares (no GPU calls)."""
import ctypes as C
import re
from pathlib import Path

from terraingen_amd import native as N

ROOT = Path(__file__).resolve().parent.parent


def declared(header: str, prefix: str) -> set[str]:
    text = (ROOT / "include" / header).read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return set(re.findall(rf"\b({prefix}\w+)\s*\(", text))


def test_hip_library_exports_every_declared_symbol():
    lib = C.CDLL(str(N.HIP_SO))
    names = declared("tg_hip.h", "tg_hip_")
    assert names, "no entry points parsed"
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, missing


def test_python_signatures_cover_the_header():
    assert declared("tg_hip.h", "tg_hip_") == set(N.HIP_ABI)


def test_layout_sizes():
    for name, size in N.SIZES.items():
        assert C.sizeof(getattr(N, name)) == size
