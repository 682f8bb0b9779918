# SPDX-License-Identifier: BSD-3-Clause
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the hip backend's kernels)")


@pytest.fixture(scope="session")
def hip():
    """The hip backend's device 0; a GPU test fails (never falls back) when it cannot open."""
    from terraingen_amd.native import HipDevice

    dev = HipDevice(0)
    print(f"[hip] device: {dev.name}", flush=True)
    return dev


@pytest.fixture(scope="session")
def orc():
    from oracle import oracle

    return oracle()


@pytest.fixture(scope="session")
def ref():
    from oracle import reference

    lib = reference()
    if lib is None:
        pytest.skip("oracle/_ref (the reference compiled from its sources) is not built here")
    return lib
