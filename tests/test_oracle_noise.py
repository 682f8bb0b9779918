# SPDX-License-Identifier: BSD-3-Clause
"""The noise oracle pinned before it is trusted: the reference's golden values
(tests/golden/noise/kernels_test.cpp:21-103 of the reference), the reference's own headers
compiled beside it (oracle/_ref), and invariants the mathematics fixes."""
import ctypes as C

import numpy as np
import pytest

from terraingen_amd import native as N

hx = float.fromhex


def frac(**kw):
    return N.make_fractal(**kw)


def test_pcg_hash_golden(orc):
    # kernels_test.cpp:21-25
    assert orc.or_pcg_hash(0) == 129708002
    assert orc.or_pcg_hash(1) == 2831084092
    assert orc.or_pcg_hash(0xDEADBEEF) == 1730779506


def test_pcg3d_golden(orc):
    # kernels_test.cpp:27-29
    out = (C.c_uint32 * 3)()
    orc.or_pcg3d((C.c_uint32 * 3)(1, 2, 3), out)
    assert list(out) == [4204755366, 1223881804, 1500469937]


def test_hash_to_unit_float(orc):
    # tests/unit/noise/hash_test.cpp:61-66
    assert orc.or_hash_to_unit_float(0) == 0.0
    assert orc.or_hash_to_unit_float(0xFFFFFFFF) == 1.0 - 2.0 ** -24
    assert orc.or_hash_to_unit_float(0x80000000) == 0.5


@pytest.mark.parametrize("x,y,want", [
    ("0x1p-2", "0x1.8p-1", "0x1.93a3e6p-4"),
    ("-0x1.9p+3", "0x1.9p+1", "-0x1.459948p-1"),
    ("0x1.f42666p+9", "-0x1.f56666p+7", "0x1.0de226p-1"),
])
def test_perlin2_golden(orc, x, y, want):
    # kernels_test.cpp:31-35
    assert orc.or_perlin2(hx(x), hx(y), 1234) == hx(want)


@pytest.mark.parametrize("x,y,want", [
    ("0x0p+0", "0x0p+0", "0x1.822c5cp+3"),
    ("0x1.004p+9", "-0x1.001p+10", "-0x1.a18ed6p+0"),
    ("-0x1.77p+11", "0x1.000cp+12", "0x1.972ec2p+3"),
])
def test_fbm_golden(orc, x, y, want):
    # kernels_test.cpp:37-43: FractalParams defaults with seed 42
    f = frac(seed=42)
    assert orc.or_fractal2(hx(x), hx(y), C.byref(f)) == hx(want)


def test_simplex_value_worley_golden(orc):
    # kernels_test.cpp:45-54
    assert orc.or_simplex2(0.25, 0.75, 1234) == hx("0x1.bac986p-4")
    assert orc.or_simplex2(hx("-0x1.9p+3"), hx("0x1.9p+1"), 1234) == hx("0x1.88e9b6p-2")
    assert orc.or_value2(0.25, 0.75, 1234) == hx("-0x1.0ed85ap-1")
    assert orc.or_value2(hx("-0x1.9p+3"), hx("0x1.9p+1"), 1234) == hx("-0x1.a6cecp-2")
    w = (C.c_float * 3)()
    orc.or_worley2(0.25, 0.75, 1234, w)
    assert w[0] == np.float32(hx("0x1.0f7b3p-1"))
    assert w[1] == np.float32(hx("0x1.2e7808p-1"))
    # kernels_test.cpp:82 (Voronoi value of the nearest cell)
    assert w[2] == np.float32(hx("-0x1.73b024p-1"))
    assert orc.or_basis2(6, hx("-0x1.9p+3"), hx("0x1.9p+1"), 1234) == hx("-0x1.e2e784p-1")


@pytest.mark.parametrize("kind,basis,want", [
    (1, 0, "0x1.8fd9ap+5"),    # ridged
    (2, 0, "-0x1.26132cp+6"),  # billow
    (3, 0, "0x1.58502ap+5"),   # hybrid multifractal
    (0, 1, "-0x1.0f204ep+4"),  # simplex fBm
])
def test_fractal_types_golden(orc, kind, basis, want):
    # kernels_test.cpp:56-68
    f = frac(seed=42, type=kind, basis=basis)
    assert orc.or_fractal2(hx("0x1.004p+9"), hx("-0x1.001p+10"), C.byref(f)) == hx(want)


def test_voronoi_fbm_golden(orc):
    # kernels_test.cpp:89-91
    f = frac(seed=42, basis=6)
    assert orc.or_fractal2(hx("0x1.0008p+9"), hx("-0x1.004p+10"), C.byref(f)) == hx("-0x1.ab66bp+5")


def test_warped_stack_golden(orc):
    # kernels_test.cpp:70-77: two fractals, the second Worley F2-F1 with a 50 m warp
    layers = (N.Fractal * 2)(frac(seed=1), frac(seed=2, basis=5, amplitude=10.0,
                                               warp_amplitude=50.0))
    v = orc.or_eval_fractals(hx("0x1.004p+9"), hx("-0x1.001p+10"), layers, 2)
    assert v == hx("-0x1.38f3cp+4")


def _random_fractals(rng, n):
    out = []
    for _ in range(n):
        out.append(frac(seed=int(rng.integers(0, 2**32)), wavelength=float(rng.uniform(5, 900)),
                        amplitude=float(rng.uniform(0.5, 50)), octaves=int(rng.integers(1, 10)),
                        lacunarity=float(rng.uniform(1.5, 2.5)), gain=float(rng.uniform(0.3, 0.7)),
                        basis=int(rng.integers(0, 7)), type=int(rng.integers(0, 4)),
                        offset=float(rng.uniform(0.5, 1.2)), blend=int(rng.integers(0, 4)),
                        warp_amplitude=float(rng.choice([0.0, rng.uniform(1, 60)])),
                        warp_wavelength=float(rng.uniform(50, 600))))
    return out


def test_oracle_matches_reference_bitwise(orc, ref):
    """Every basis, fractal type, blend and warp: the restatement has the reference's bits."""
    rng = np.random.default_rng(7)
    for trial in range(60):
        fr = _random_fractals(rng, int(rng.integers(1, 4)))
        arr = (N.Fractal * len(fr))(*fr)
        for _ in range(25):
            x, y = (float(np.float32(v)) for v in rng.uniform(-5000, 5000, 2))
            a = orc.or_eval_fractals(x, y, arr, len(fr))
            b = ref.ref_eval_fractals(x, y, arr, len(fr))
            assert np.float32(a).tobytes() == np.float32(b).tobytes(), (trial, x, y, a, b)
        for basis in range(7):
            x, y = (float(np.float32(v)) for v in rng.uniform(-300, 300, 2))
            seed = int(rng.integers(0, 2**32))
            assert orc.or_basis2(basis, x, y, seed) == ref.ref_basis2(basis, x, y, seed)


def test_pcg_matches_reference(orc, ref):
    rng = np.random.default_rng(3)
    for v in rng.integers(0, 2**32, 200, dtype=np.uint64):
        assert orc.or_pcg_hash(int(v)) == ref.ref_pcg_hash(int(v))


def test_noise_layer_kernel_layout_matches_reference(ref):
    """The reference's own MakeNoiseLayerKernel bytes equal this repository's 420-byte layout."""
    fr = [frac(seed=5, octaves=3), frac(enabled=False, seed=9), frac(seed=7, basis=2, blend=1)]
    arr = (N.Fractal * 3)(*fr)
    buf = (C.c_ubyte * 420)()
    ref.ref_make_noise_layer_kernel(arr, 3, buf)
    mine = N.noise_kernel([dict(seed=5, octaves=3), dict(enabled=False, seed=9),
                           dict(seed=7, basis=2, blend=1)])
    assert bytes(buf) == bytes(mine)


def test_perlin_is_zero_on_the_lattice(orc):
    """Gradient noise vanishes at every lattice point (perlin.h:68-70): a dropped or shifted
    fraction would break this."""
    rng = np.random.default_rng(1)
    for ix, iy in rng.integers(-10000, 10000, (50, 2)):
        assert orc.or_perlin2(float(ix), float(iy), int(rng.integers(0, 2**32))) == 0.0


def test_perlin_is_continuous_across_cells(orc):
    """Neighbouring cells meet: the value just left of an edge approaches the value just right
    of it (a wrong corner hash or sign would make a jump)."""
    rng = np.random.default_rng(2)
    for _ in range(200):
        ix, y = int(rng.integers(-500, 500)), float(rng.uniform(-500, 500))
        seed = int(rng.integers(0, 2**32))
        left = orc.or_perlin2(float(np.nextafter(np.float32(ix), np.float32(-1e9))), y, seed)
        right = orc.or_perlin2(float(ix), y, seed)
        assert abs(left - right) < 1e-4


def test_fbm_is_bounded_by_its_amplitude(orc):
    rng = np.random.default_rng(4)
    f = frac(seed=11, amplitude=20.0, wavelength=512.0)
    vals = [orc.or_fractal2(float(x), float(y), C.byref(f)) for x, y in rng.uniform(-4e3, 4e3, (2000, 2))]
    assert max(abs(v) for v in vals) <= 20.0
    assert np.std(vals) > 1.0


def test_fp64_oracle_tracks_fp32(orc):
    """The FP64 oracle differs from the FP32 path by rounding only (relative 1e-5 of the
    amplitude at coordinates of a few km)."""
    rng = np.random.default_rng(5)
    f = frac(seed=1, amplitude=20.0, wavelength=512.0)
    for x, y in rng.uniform(-1100, 1100, (500, 2)):
        x32, y32 = float(np.float32(x)), float(np.float32(y))
        a = orc.or_fractal2(x32, y32, C.byref(f))
        b = orc.or_fractal2_f64(x32, y32, C.byref(f))
        assert abs(a - b) < 20.0 * 1e-5
