# SPDX-License-Identifier: BSD-3-Clause
"""Boulder placement (tgx_place_boulders): the boulders a BoulderLayer of power-law sizes lists
over a region for a grid spacing are the reference's own, every field bit for bit and in order
(BoulderLayer::Boulders, src/layers/boulders/boulder_layer.cpp:263-420: candidates drawn as
craters are with the layer's candidate sizes, each thinned and posed from its own Philox streams),
against committed fixtures the reference made and, where oracle/_ref is built, live over planes,
spheres, spacings, seeds and pose ranges; on any number of threads. The sizes follow the power law
(a distribution test pins the candidate sampler's inversion on its own)."""
from pathlib import Path

import numpy as np
import pytest

from terraingen_amd import native as N

GOLDEN = Path(__file__).resolve().parent / "golden"
FIELDS = ["x", "y", "diameter", "height", "burial", "yaw", "lean", "wear", "id", "model"]

import sys  # noqa: E402

sys.path.insert(0, str(GOLDEN))
from make_boulder_fixtures import PLACEMENT_CASES as CASES  # noqa: E402
from make_boulder_fixtures import placement_spec as spec  # noqa: E402


def mine(case: dict, threads: int = 0, **pose) -> np.ndarray:
    return N.place_boulders(case["region"], case["spacing"], 3.8e-3, 3.0, 0.2, 3.0,
                            seed=case["seed"], sphere_radius=case.get("sphere_radius", 0.0),
                            threads=threads, **pose)


def same(got, want):
    assert got.size == want.size
    for f in FIELDS:
        assert np.array_equal(got[f], want[f]), f


@pytest.mark.parametrize("name", CASES)
def test_boulders_equal_the_reference_fixtures(name):
    want = np.load(GOLDEN / f"boulders_{name}.npz")["boulders"]
    same(mine(CASES[name]), want)


@pytest.mark.parametrize("case", [
    dict(region=(0.0, 0.0, 200.0, 150.0), spacing=0.25, seed=1),
    dict(region=(-5000.0, 2000.0, -4880.0, 2090.0), spacing=0.1, seed=9),
    dict(region=(10.0, 10.0, 40.0, 30.0), spacing=0.0, seed=4),
    dict(region=(-120.01, 60.0, -119.995, 60.01), spacing=0.3, seed=5, sphere_radius=1737400.0),
    dict(region=(179.998, -1.0, 180.0, -0.998), spacing=0.4, seed=6, sphere_radius=1737400.0),
])
def test_boulders_equal_the_reference_live(ref, case):
    from oracle import reference_boulders

    want = reference_boulders(ref, spec(case))
    assert want.size > 5
    same(mine(case), want)


def test_pose_ranges_follow_the_layer(ref):
    from oracle import reference_boulders

    case = dict(region=(0.0, 0.0, 50.0, 50.0), spacing=0.2, seed=12)
    pose = dict(min_height_ratio=0.2, max_height_ratio=0.9, min_burial=0.0, max_burial=0.7,
                max_lean_degrees=40.0)
    s = spec(case)
    for k, v in pose.items():
        setattr(s, k, v)
    same(mine(case, **pose), reference_boulders(ref, s))


def test_thread_count_does_not_change_the_list():
    case = dict(region=(0.0, 0.0, 300.0, 300.0), spacing=0.25, seed=3)
    one, many = mine(case, threads=1), mine(case, threads=8)
    assert one.size > 1000
    same(one, many)


def test_sizes_and_poses_follow_their_laws():
    """Diameters follow N(>=D) ~ D^-3 between the resolved size and 3 m; heights 0.4-0.6 D;
    burial 0.1-0.4; yaw in [0, 2 pi); lean under 0.26 rad; IDs unique."""
    got = mine(dict(region=(0.0, 0.0, 400.0, 400.0), spacing=0.25, seed=8))
    d = got["diameter"]
    assert d.min() >= 0.5 and d.max() <= 3.0
    # The fraction above 1 m among those above 0.5 m: (1^-3 - 3^-3) / (0.5^-3 - 3^-3).
    expect = (1.0 - 3.0 ** -3) / (0.5 ** -3 - 3.0 ** -3)
    assert abs(np.mean(d >= 1.0) - expect) < 4.0 * np.sqrt(expect / d.size)
    ratio = got["height"] / d
    assert ratio.min() >= 0.4 and ratio.max() <= 0.6
    assert got["burial"].min() >= 0.1 and got["burial"].max() <= 0.4
    assert got["yaw"].min() >= 0.0 and got["yaw"].max() < 2.0 * np.pi
    assert got["lean"].max() < 0.26
    assert np.unique(got["id"]).size == got.size


@pytest.mark.parametrize("radius, boxes", [
    (1737400.0, [(20.0, 5.0, 20.004, 5.004), (20.003, 5.003, 20.008, 5.006),
                 (-60.0, 40.0, -59.997, 40.002)]),
    (0.0, [(0.0, 0.0, 40.0, 30.0), (35.0, 20.0, 70.0, 60.0), (-100.0, 5.0, -80.0, 9.0)]),
])
def test_boxes_list_the_union_of_their_regions(radius, boxes):
    """tgx_place_boulders_boxes lists each boulder of every region once, with the draws it has
    in that region's own list (a served batch's catalogue is the union of its tiles' lists)."""
    args = (0.25, 3.8e-3, 3.0, 0.2, 3.0)
    got = N.place_boulders_boxes(np.array(boxes), *args, seed=5, sphere_radius=radius)
    assert np.unique(got["id"]).size == got.size
    by_id = {int(b["id"]): b for b in got}
    union = set()
    for box in boxes:
        alone = N.place_boulders(box, *args, seed=5, sphere_radius=radius)
        assert alone.size > 0
        for b in alone:
            mine = by_id[int(b["id"])]
            assert mine.tobytes() == b.tobytes()
            union.add(int(b["id"]))
    assert union == set(by_id)


def test_coarser_spacing_lists_the_resolved_subset():
    """A boulder a coarser spacing resolves (D above twice it) is listed at every finer spacing
    with the same draws, so one catalogue at the finest spacing serves tiles of every spacing."""
    box, radius = (20.0, 5.0, 20.01, 5.01), 1737400.0
    fine = N.place_boulders(box, 0.125, 3.8e-3, 3.0, 0.2, 3.0, seed=5, sphere_radius=radius)
    for spacing in (0.25, 0.5, 1.0):
        coarse = N.place_boulders(box, spacing, 3.8e-3, 3.0, 0.2, 3.0, seed=5,
                                  sphere_radius=radius)
        a = fine[fine["diameter"] > 2.0 * spacing]
        b = coarse[coarse["diameter"] > 2.0 * spacing]
        assert a.size > 0 and a.tobytes() == b.tobytes()
