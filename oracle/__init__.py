# SPDX-License-Identifier: BSD-3-Clause
"""ORACLE - test infrastructure only.

ctypes loaders for ``liboracle.so`` (this repository's CPU restatement of the reference, in the
reference-equivalent FP32 form and as an FP64 oracle) and ``_ref/libtgref.so`` (the reference's
own code compiled from its sources). Only ``tests/``, ``__graft_entry__.smoke()`` and
``bench.py``'s CPU-baseline leg may import this package; the hip backend never does.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ORACLE_SO = HERE / "liboracle.so"
REF_SO = HERE / "_ref" / "libtgref.so"

_oracle = None
_ref = None


def oracle() -> C.CDLL:
    global _oracle
    if _oracle is None:
        if not ORACLE_SO.exists():
            raise RuntimeError(f"{ORACLE_SO} is missing: run `make -C oracle`")
        _oracle = C.CDLL(str(ORACLE_SO))
        _declare(_oracle)
    return _oracle


def reference() -> C.CDLL | None:
    """The reference's own code, or None where it was not built (the GPU hosts have only what was
    built here)."""
    global _ref
    if _ref is None and REF_SO.exists():
        # Lazy binding: the reference's GPU entry points stay unresolved in this CPU-only build.
        _ref = C.CDLL(str(REF_SO), mode=os.RTLD_LAZY)
        _declare(_ref)
    return _ref


def _ref_case_type():
    from terraingen_amd import native as N

    u32, f64 = C.c_uint32, C.c_double
    return type("RefCase", (C.Structure,), {"_fields_": [
        ("width", u32), ("height", u32), ("pixel", f64), ("center_x", f64), ("center_y", f64),
        ("elevation", C.c_float), ("source_kind", u32), ("fractal_count", u32),
        ("fractals", N.Fractal * 8), ("craters", u32), ("models", u32),
        ("sfd_coefficient", f64), ("sfd_exponent", f64), ("min_d", f64), ("max_d", f64),
        ("lifetime", f64), ("transition_diameter", f64), ("irregular", u32), ("peaks", u32),
        ("terraces", u32), ("terrace_onset", f64), ("terrace_full", f64), ("seed", u32),
        ("threads", u32), ("rim_sigma", f64), ("noise_amplitude", f64),
        ("noise_rim_amplitude", f64), ("wall_streaks", f64), ("ejecta", f64)]})


RefCase = _ref_case_type()


def _ref_boulder_case_type():
    from terraingen_amd import native as N

    u32, f64 = C.c_uint32, C.c_double
    return type("RefBoulderCase", (C.Structure,), {"_fields_": [
        ("width", u32), ("height", u32), ("pixel", f64), ("center_x", f64), ("center_y", f64),
        ("elevation", C.c_float), ("source_kind", u32), ("fractal_count", u32),
        ("fractals", N.Fractal * 8), ("sizes_kind", u32), ("sfd_coefficient", f64),
        ("sfd_exponent", f64), ("min_d", f64), ("max_d", f64), ("ventifact_share", f64),
        ("seed", u32), ("threads", u32)]})


RefBoulderCase = _ref_boulder_case_type()


class RefBoulderSphereCase(C.Structure):
    """``ref_boulder_sphere_case`` of ref_boulders.cpp: a power-law boulder layer on a sphere
    over a longitude-latitude raster, flat at 0 beneath."""
    _fields_ = [("side", C.c_uint32), ("lon", C.c_double), ("lat", C.c_double),
                ("step", C.c_double), ("radius", C.c_double), ("sfd_coefficient", C.c_double),
                ("sfd_exponent", C.c_double), ("min_d", C.c_double), ("max_d", C.c_double),
                ("ventifact_share", C.c_double), ("seed", C.c_uint32), ("threads", C.c_uint32)]


class RefSphereCase(C.Structure):
    """``ref_sphere_case`` of ref_sphere.cpp: a crater layer on a sphere over a lon/lat raster."""

    _fields_ = [
        ("side", C.c_uint32), ("lon", C.c_double), ("lat", C.c_double), ("step", C.c_double),
        ("radius", C.c_double), ("models", C.c_uint32), ("sfd_coefficient", C.c_double),
        ("sfd_exponent", C.c_double), ("min_d", C.c_double), ("max_d", C.c_double),
        ("lifetime", C.c_double), ("irregular", C.c_uint32), ("elongated_share", C.c_double),
        ("seed", C.c_uint32), ("threads", C.c_uint32), ("manual_count", C.c_uint32),
        ("manual", (C.c_double * 4) * 4)]


# ref_placement.cpp's ref_sphere_generation: GenerateCraters on a sphere's cube-sphere cells.
RefSphereGeneration = type("RefSphereGeneration", (C.Structure,), {"_fields_": [
    (n, C.c_double) for n in ("radius", "min_lon", "min_lat", "max_lon", "max_lat",
                              "sfd_coefficient", "sfd_exponent", "min_d", "max_d",
                              "margin_radii", "min_resolved", "lifetime")] + [
    ("seed", C.c_uint32), ("threads", C.c_uint32)]})


def reference_sphere_craters(lib: C.CDLL, **spec) -> np.ndarray:
    """The reference's crater list on a sphere (ref_generate_sphere), in CRATER_DTYPE."""
    from terraingen_amd import native as N

    g = RefSphereGeneration(**spec)
    n = lib.ref_generate_sphere(C.byref(g), None, None, None, None, None, 0)
    if n < 0:
        raise RuntimeError(lib.ref_placement_error().decode())
    cols = [np.zeros(n) for _ in range(4)] + [np.zeros(n, np.uint64)]
    lib.ref_generate_sphere(C.byref(g), *[c.ctypes.data for c in cols], n)
    out = np.zeros(n, N.CRATER_DTYPE)
    for name, col in zip(("x", "y", "diameter", "age_gyr", "id"), cols):
        out[name] = col
    return out


def _declare(lib: C.CDLL) -> None:
    from terraingen_amd import native as N

    f, d, u, sz, vp = C.c_float, C.c_double, C.c_uint32, C.c_size_t, C.c_void_p
    fp = C.POINTER(N.Fractal)
    sigs = {
        "or_pcg_hash": (u, [u]), "ref_pcg_hash": (u, [u]),
        "or_pcg3d": (None, [vp, vp]), "ref_pcg3d": (None, [vp, vp]),
        "or_hash_to_unit_float": (f, [u]),
        "or_perlin2": (f, [f, f, u]), "or_simplex2": (f, [f, f, u]), "or_value2": (f, [f, f, u]),
        "or_worley2": (None, [f, f, u, vp]),
        "or_basis2": (f, [u, f, f, u]), "ref_basis2": (f, [u, f, f, u]),
        "or_fractal2": (f, [f, f, fp]), "ref_fractal2": (f, [f, f, fp]),
        "or_fractal2_f64": (d, [d, d, fp]),
        "or_eval_fractals": (f, [f, f, fp, sz]), "ref_eval_fractals": (f, [f, f, fp, sz]),
        "or_eval_fractals_f64": (d, [d, d, fp, sz]),
        "or_fractal3": (f, [d, d, d, fp]), "ref_fractal3": (f, [d, d, d, fp]),
        "or_basis3": (f, [u, f, f, f, u]), "ref_basis3": (f, [u, f, f, f, u]),
        "or_worley3": (None, [f, f, f, u, vp]),
        "or_eval_fractals3": (f, [d, d, d, fp, sz]), "ref_eval_fractals3": (f, [d, d, d, fp, sz]),
        "or_eval_fractals3_f64": (d, [d, d, d, fp, sz]),
        "or_body_noise_grid": (None, [C.POINTER(N.NoiseLayerKernel), vp, sz,
                                      C.POINTER(N.KernelParams), vp, C.c_uint]),
        "or_body_noise_grid_f64": (None, [C.POINTER(N.NoiseLayerKernel), vp, sz, vp,
                                          C.c_uint]),
        "ref_make_noise_layer_kernel": (None, [fp, sz, vp]),
        "or_noise_grid": (None, [C.POINTER(N.NoiseLayerKernel), vp, vp, sz, sz,
                                 C.POINTER(N.KernelParams), vp, C.c_uint]),
        "or_noise_grid_f64": (None, [C.POINTER(N.NoiseLayerKernel), vp, vp, sz, sz, vp,
                                     C.c_uint]),
        "ref_noise_grid": (None, [vp, vp, vp, sz, sz, vp]),
        "or_apply_crater": (f, [vp, vp, vp, f, f, f]),
        "or_apply_crater_f64": (d, [vp, vp, vp, d, d, d]),
        "or_shape_grid": (None, [vp, vp, vp, vp, vp, vp, vp, vp, sz, sz,
                                 C.POINTER(N.KernelParams), vp, C.c_uint]),
        "or_shape_grid_f64": (None, [vp, vp, vp, vp, vp, vp, vp, sz, sz, u, vp, C.c_uint]),
        "or_shape_grid_body": (None, [vp, vp, vp, vp, vp, vp, vp, sz, sz,
                                      C.POINTER(N.KernelParams), vp, C.c_uint]),
        "or_shape_grid_body_f64": (None, [vp, vp, vp, vp, vp, vp, vp, sz, sz, u, vp, C.c_uint]),
        "ref_apply_crater": (f, [vp, vp, vp, f, f, f]),
        "ref_last_error": (C.c_char_p, []),
        "ref_case_open": (vp, [C.POINTER(RefCase)]),
        "ref_case_close": (None, [vp]),
        "ref_case_evaluate": (C.c_int, [vp, vp]),
        "ref_case_craters": (C.c_long, [vp, vp, vp, vp, vp, vp]),
        "ref_case_batch": (C.c_long, [vp, vp, C.POINTER(C.c_long)]),
        "ref_case_batch_copy": (None, [vp, vp, vp]),
        "or_boulder_grid": (None, [vp, sz, vp, C.c_int, vp]),
        "or_boulder_grid_f64": (None, [vp, sz, vp, C.c_int, vp]),
        "or_boulder_relief_f64": (d, [vp, d, d, d]),
        "or_boulder_body_f64": (None, [vp, sz, vp, d, C.c_int, vp]),
        "or_sphere_offset": (None, [d, d, d, d, d, vp]),
        "or_upsample_grid": (None, [C.POINTER(N.Upsample), vp, vp]),
        "or_kernel_weights": (None, [u, d, vp]),
        "ref_synthetic_grid": (C.c_int, [vp, u, f, d, u, u, u, d, d, d, d, vp]),
        "ref_geotiff_grid": (C.c_int, [C.c_char_p, u, vp, u, u, d, d, d, d, d, vp]),
        "ref_sources_error": (C.c_char_p, []),
        "ref_sources_threads": (None, [u]),
        "ref_geotiff_read": (C.c_long, [C.c_char_p, u, vp, vp, vp, vp]),
        "ref_boulders_open": (vp, [C.POINTER(RefBoulderCase)]),
        "ref_boulders_close": (None, [vp]),
        "ref_boulders_evaluate": (C.c_int, [vp, vp]),
        "ref_boulders_list": (C.c_long, [vp, vp]),
        "ref_boulders_error": (C.c_char_p, []),
        "ref_boulders_sphere": (C.c_long, [C.POINTER(RefBoulderSphereCase), vp, vp, C.c_long]),
        "ref_sphere_open": (vp, [C.POINTER(RefSphereCase)]),
        "ref_sphere_close": (None, [vp]),
        "ref_sphere_sizes": (C.c_long, [vp, C.POINTER(C.c_long)]),
        "ref_sphere_copy": (None, [vp, vp, vp, vp, vp, vp, vp]),
        "ref_sphere_error": (C.c_char_p, []),
        "ref_body_tile": (C.c_int, [vp, sz, d, d, d, u, u, d, vp, vp, vp, vp, vp, vp, C.c_uint,
                                    vp]),
        "ref_plane_tile": (C.c_int, [vp, sz, d, d, d, u, u, vp, vp, vp, vp, vp, C.c_uint, vp]),
        "ref_bench_error": (C.c_char_p, []),
        "ref_generate_sphere": (C.c_long, [C.POINTER(RefSphereGeneration), vp, vp, vp, vp, vp,
                                           C.c_long]),
        "ref_placement_error": (C.c_char_p, []),
        "ref_place_boulders": (C.c_long, [vp, vp, vp, vp, C.c_long]),
        "ref_write_geotiff": (C.c_int, [C.c_char_p, vp, u, u, u, vp, C.c_int, f, vp, sz, vp, sz,
                                        C.c_char_p, C.c_int, u, C.c_int]),
        "ref_write_crater_csv": (C.c_int, [C.c_char_p, vp, vp, vp, vp, vp, vp, vp, sz, vp, vp]),
        "ref_output_error": (C.c_char_p, []),
        "or_datum_grid": (None, [vp, vp, vp, vp, vp, vp, vp, vp, C.POINTER(N.KernelParams), vp,
                                 C.c_uint]),
        "or_body_datum_grid": (None, [vp, vp, vp, vp, vp, vp, vp, vp,
                                      C.POINTER(N.KernelParams), vp, C.c_uint]),
        "ref_datum_grid": (None, [vp, vp, vp, vp, vp, vp, vp, vp, C.POINTER(N.KernelParams),
                                  vp]),
        "ref_body_datum_grid": (None, [vp, vp, vp, vp, vp, vp, vp, vp,
                                       C.POINTER(N.KernelParams), vp]),
    }
    for name, (res, args) in sigs.items():
        fn = getattr(lib, name, None)
        if fn is not None:
            fn.restype = res
            fn.argtypes = args


def ref_write_geotiff(lib: C.CDLL, path, values: np.ndarray, transform, nodata=None,
                      geokeys=None, compression: int = 1, tile_size: int = 256,
                      big_tiff: bool = False) -> None:
    """The reference's WriteGeoTiff of `values` ((bands,) height x width float32)."""
    v = np.ascontiguousarray(values, np.float32)
    bands = 1 if v.ndim == 2 else v.shape[0]
    h, w = v.shape[-2:]
    t = np.ascontiguousarray(transform, np.float64)
    directory, doubles, ascii_ = geokeys if geokeys is not None else ((), (), "")
    dk = np.ascontiguousarray(directory, np.uint16)
    dd = np.ascontiguousarray(doubles, np.float64)
    rc = lib.ref_write_geotiff(str(path).encode(), v.ctypes.data, w, h, bands, t.ctypes.data,
                               int(nodata is not None),
                               float(nodata) if nodata is not None else 0.0,
                               dk.ctypes.data, dk.size, dd.ctypes.data, dd.size,
                               ascii_.encode(), compression, tile_size, int(big_tiff))
    if rc != 0:
        raise RuntimeError(lib.ref_output_error().decode())


def ref_write_crater_csv(lib: C.CDLL, path, craters: np.ndarray, manual=None, replace=None,
                         attributes=None) -> None:
    """The reference's WriteCraterList (CSV) of a CRATER_DTYPE list; `attributes` is a
    structured array with floor_depth, rim_height, depth_ratio, freshness and complex_."""
    cols = [np.ascontiguousarray(craters[k]) for k in ("x", "y", "diameter", "age_gyr", "id")]
    flags = [np.ascontiguousarray(a, np.uint8) if a is not None else None
             for a in (manual, replace)]
    if attributes is not None:
        attrs = np.ascontiguousarray(np.stack([attributes[k] for k in (
            "floor_depth", "rim_height", "depth_ratio", "freshness")], axis=1), np.float64)
        cx = np.ascontiguousarray(attributes["complex_"], np.uint8)
    rc = lib.ref_write_crater_csv(str(path).encode(), *[c.ctypes.data for c in cols],
                                  *[a.ctypes.data if a is not None else None for a in flags],
                                  craters.size,
                                  attrs.ctypes.data if attributes is not None else None,
                                  cx.ctypes.data if attributes is not None else None)
    if rc != 0:
        raise RuntimeError(lib.ref_output_error().decode())


def reference_boulders(lib: C.CDLL, spec) -> np.ndarray:
    """The reference's BoulderLayer::Boulders for a BoulderLayerSpec (ref_place_boulders), in
    PLACED_BOULDER_DTYPE."""
    from terraingen_amd import native as N

    n = lib.ref_place_boulders(C.byref(spec), None, None, None, 0)
    if n < 0:
        raise RuntimeError(lib.ref_placement_error().decode())
    values = np.zeros((max(n, 1), 8))
    ids = np.zeros(max(n, 1), np.uint64)
    models = np.zeros(max(n, 1), np.uint8)
    lib.ref_place_boulders(C.byref(spec), values.ctypes.data, ids.ctypes.data,
                           models.ctypes.data, n)
    out = np.zeros(n, N.PLACED_BOULDER_DTYPE)
    for k, name in enumerate(("x", "y", "diameter", "height", "burial", "yaw", "lean", "wear")):
        out[name] = values[:n, k]
    out["id"], out["model"] = ids[:n], models[:n]
    return out
