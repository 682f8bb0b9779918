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


def test_host_library_exports_every_declared_symbol():
    lib = N.host_library()
    names = declared("terraingen_amd.h", "tgx_")
    assert names, "no entry points parsed"
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, missing
    assert declared("terraingen_amd.h", "tgx_") == set(N.HOST_ABI)


def test_host_library_opens_the_kernel_library_lazily():
    """The kernel library opens on first use; here it opens (no device call is made)."""
    assert N.host_library().tgx_hip_unavailable_reason() == b""


def test_server_rejects_bad_arguments():
    cfg = N.ServerConfig(0, 8, 1 << 20, 1 << 24, 1, 0)
    out = C.c_void_p()
    err = C.create_string_buffer(256)
    code = N.host_library().tgx_server_create(None, None, 0.0, None, None, C.byref(cfg),
                                              C.byref(out), err, 256)
    assert code == 1 and b"null" in err.value and not out.value
