# SPDX-License-Identifier: BSD-3-Clause
"""ctypes view of the native libraries: the std430 structs of include/tg_layouts.h and the C ABI of
include/tg_hip.h.

Loading is strict: a missing or unloadable ``libtg_hip.so`` raises, there is no fallback path.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

PKG = Path(__file__).resolve().parent
HIP_SO = PKG / "libtg_hip.so"
HOST_SO = PKG / "libterraingen_amd.so"


def _struct(name: str, fields: list[tuple[str, object]]) -> type:
    return type(name, (C.Structure,), {"_fields_": fields})


u32, i32, f32 = C.c_uint32, C.c_int32, C.c_float

Fractal = _struct("Fractal", [
    ("enabled", u32), ("seed", u32), ("wavelength", f32), ("amplitude", f32), ("octaves", i32),
    ("lacunarity", f32), ("gain", f32), ("basis", u32), ("type", u32), ("offset", f32),
    ("blend", u32), ("warp_amplitude", f32), ("warp_wavelength", f32)])

NoiseLayerKernel = _struct("NoiseLayerKernel", [("fractals", Fractal * 8), ("count", u32)])

KernelParams = _struct("KernelParams", [
    ("width", u32), ("rows", u32), ("no_data", u32), ("no_data_value", f32), ("first_row", u32)])

CraterProfile = _struct("CraterProfile", [
    ("simple", u32), ("rim", u32), ("complex_", u32), ("diffused", u32),
    ("transition_diameter", f32), ("floor_coefficient", f32), ("floor_exponent", f32),
    ("peak_height_coefficient", f32), ("peak_height_exponent", f32),
    ("peak_width_coefficient", f32), ("peak_width_exponent", f32), ("peak_from_volume", u32),
    ("peak_fit_max_diameter", f32), ("ejecta_exponent", f32), ("complex_reach_radii", f32),
    ("table_extent", f32), ("table_cells", u32), ("ages_offset", u32), ("age_count", u32)])

CraterShape = _struct("CraterShape", [("profile", CraterProfile)] + [
    (n, t) for n, t in [
        ("reach_radii", f32), ("profile_reach_radii", f32), ("blend_start", f32),
        ("blend_end", f32), ("blend_shape", u32), ("interior_replacement", f32),
        ("replace_start", f32), ("replace_end", f32), ("mantling", f32),
        ("ejecta_reach_radii", f32), ("rim_sigma", f32), ("rim_inner", f32), ("rim_outer", f32),
        ("rim_strength", f32), ("rim_shape", u32), ("noise_amplitude", f32),
        ("noise_wavelength", f32), ("noise_octaves", i32), ("noise_gain", f32),
        ("noise_rim_amplitude", f32), ("noise_rim_wavelength", f32), ("noise_rim_inner", f32),
        ("noise_rim_outer", f32), ("wall_streaks", f32), ("wall_streak_width", f32),
        ("ejecta", f32), ("ejecta_lineation", f32), ("min_freshness", f32)]])


class CraterDatumKernel(C.Structure):
    """CraterDatumKernel (craters.h:1214-1221): shapes, fit footprint in radii, lattice side,
    size bands of the crater index."""
    _fields_ = [("shape", CraterShape), ("fit_radii", C.c_float), ("lattice_side", C.c_uint32),
                ("band_count", C.c_uint32)]


class CraterBodyDatumKernel(C.Structure):
    """CraterBodyDatumKernel (craters.h:1308-1311)."""
    _fields_ = [("shape", CraterShape)]


assert C.sizeof(CraterDatumKernel) == 200 and C.sizeof(CraterBodyDatumKernel) == 188

RECORD_FIELDS = [
    ("x", f32), ("y", f32), ("diameter", f32), ("resolution", f32), ("fresh_depth", f32),
    ("fresh_rim", f32), ("current_depth", f32), ("current_rim", f32), ("diffusion", f32),
    ("ellipticity", f32), ("heading", f32), ("first_harmonic", u32), ("harmonic_count", u32),
    ("harmonic_offset", u32), ("table_count", u32), ("table_offset", u32),
    ("profile_offset", u32), ("has_datum", u32), ("datum_a", f32), ("datum_b", f32),
    ("datum_c", f32), ("noise_key", u32), ("texture_key", u32), ("mark_count", u32),
    ("mark_offset", u32), ("shift_x", f32), ("shift_y", f32), ("upslope_x", f32),
    ("upslope_y", f32), ("terrace_count", u32), ("terrace_offset", u32), ("peak_count", u32),
    ("peak_offset", u32), ("replace", u32), ("x_low", f32), ("y_low", f32)]
CraterRecord = _struct("CraterRecord", RECORD_FIELDS)
RECORD_DTYPE = np.dtype([(n, np.float32 if t is f32 else np.uint32) for n, t in RECORD_FIELDS])

CraterFrame = _struct("CraterFrame", [(n, f32) for n in (
    "x", "y", "z", "east_x", "east_y", "east_z", "north_x", "north_y", "north_z",
    "sphere_radius", "x_low", "y_low", "z_low")])
FRAME_DTYPE = np.dtype([(n, np.float32) for n, _ in CraterFrame._fields_])

SplitPoint = _struct("SplitPoint", [(n, f32) for n in ("x", "x_low", "y", "y_low", "z", "z_low")])
POINT_DTYPE = np.dtype([(n, np.float32) for n, _ in SplitPoint._fields_])

TgBuf = _struct("TgBuf", [("data", C.c_void_p), ("bytes", C.c_size_t), ("count", C.c_size_t),
                          ("per_cell", C.c_size_t)])

SIZES = {"Fractal": 52, "NoiseLayerKernel": 420, "CraterProfile": 76, "CraterShape": 188,
         "CraterRecord": 144, "CraterFrame": 52, "SplitPoint": 24, "KernelParams": 20}
for _name, _size in SIZES.items():
    assert C.sizeof(globals()[_name]) == _size, (_name, C.sizeof(globals()[_name]))
assert RECORD_DTYPE.itemsize == 144 and FRAME_DTYPE.itemsize == 52 and POINT_DTYPE.itemsize == 24

TG_KERNEL_NOISE_LAYER = 1
TG_KERNEL_CRATERS = 2
TG_KERNEL_BOULDERS = 3
TG_KERNEL_NOISE_BODY = 4
TG_KERNEL_CRATER_DATUMS = 5
TG_KERNEL_CRATER_BODY_DATUMS = 6
NODATA_NONE, NODATA_NAN, NODATA_VALUE = 0, 1, 2

# Every entry point include/tg_hip.h declares, with its ctypes signature.
VP, SZ, CP = C.c_void_p, C.c_size_t, C.c_char_p
Upsample = _struct("Upsample", [
    ("lattice_x", C.c_double), ("lattice_y", C.c_double), ("post", C.c_double),
    ("post_height", C.c_double), ("col0", C.c_int64), ("row0", C.c_int64), ("window_width", C.c_uint32),
    ("window_height", C.c_uint32), ("wrap_columns", C.c_uint32),
    ("lattice_columns", C.c_uint32), ("origin_x", C.c_double), ("origin_y", C.c_double),
    ("pixel", C.c_double), ("width", C.c_uint32), ("height", C.c_uint32),
    ("method", C.c_uint32), ("at_post_spacing", C.c_uint32), ("out_pitch", C.c_uint32),
    ("covered", C.c_uint32), ("cover_col0", C.c_int64), ("cover_row0", C.c_int64),
    ("cover_col1", C.c_int64), ("cover_row1", C.c_int64)])
assert C.sizeof(Upsample) == 144
# ResampleMethod (include/terraingen/raster/resample.h:12-23).
(TG_RESAMPLE_NEAREST, TG_RESAMPLE_BILINEAR, TG_RESAMPLE_CUBIC, TG_RESAMPLE_LANCZOS3,
 TG_RESAMPLE_CUBIC_BSPLINE) = range(5)

HIP_ABI = {
    "tg_hip_open": (C.c_int, [C.c_int, C.POINTER(VP), CP, SZ]),
    "tg_hip_device_name": (CP, [VP]),
    "tg_hip_can_run": (C.c_int, [VP, C.c_int]),
    "tg_hip_close": (C.c_int, [VP]),
    "tg_hip_run": (C.c_int, [VP, C.c_int, VP, SZ, VP, VP, VP, C.POINTER(KernelParams),
                             C.POINTER(TgBuf), SZ, CP, SZ]),
    "tg_hip_noise_plan_create": (C.c_int, [VP, C.POINTER(NoiseLayerKernel), C.POINTER(VP), CP,
                                           SZ]),
    "tg_hip_noise_plan_free": (C.c_int, [VP]),
    "tg_hip_noise_launch": (C.c_int, [VP, VP, VP, VP, VP, C.POINTER(KernelParams), VP, CP, SZ]),
    "tg_hip_noise_fill": (C.c_int, [VP, VP, C.c_float, VP, VP, VP, C.POINTER(KernelParams), VP,
                                    CP, SZ]),
    "tg_hip_crater_set_create": (C.c_int, [VP, C.POINTER(CraterShape), VP, SZ, VP, SZ, VP, SZ,
                                           VP, SZ, VP, SZ, VP, SZ, C.POINTER(VP), CP, SZ]),
    "tg_hip_crater_set_free": (C.c_int, [VP]),
    "tg_hip_crater_launch": (C.c_int, [VP, VP, VP, VP, VP, C.POINTER(KernelParams), VP, CP, SZ]),
    "tg_hip_catalogue_create": (C.c_int, [VP, C.POINTER(CraterShape), VP, SZ, VP, SZ, VP,
                                          C.POINTER(VP), CP, SZ]),
    "tg_hip_catalogue_free": (C.c_int, [VP]),
    "tg_hip_batch_create": (C.c_int, [VP, SZ, SZ, SZ, C.POINTER(VP), CP, SZ]),
    "tg_hip_batch_free": (C.c_int, [VP]),
    "tg_hip_batch_stage": (C.c_int, [VP, VP, SZ, VP, SZ, VP, SZ, CP, SZ]),
    "tg_hip_batch_stage_async": (C.c_int, [VP, VP, SZ, VP, SZ, VP, SZ, CP, SZ]),
    "tg_hip_batch_run": (C.c_int, [VP, VP, C.c_float, VP, VP, SZ, C.c_int, CP, SZ]),
    "tg_hip_batch_wait": (C.c_int, [VP, CP, SZ]),
    "tg_hip_batch_stream": (VP, [VP]),
    "tg_hip_malloc": (C.c_int, [VP, SZ, C.POINTER(VP), CP, SZ]),
    "tg_hip_free": (C.c_int, [VP, VP]),
    "tg_hip_memcpy_h2d": (C.c_int, [VP, VP, VP, SZ, CP, SZ]),
    "tg_hip_memcpy_d2h": (C.c_int, [VP, VP, VP, SZ, CP, SZ]),
    "tg_hip_memset": (C.c_int, [VP, VP, C.c_int, SZ, CP, SZ]),
    "tg_hip_stream_create": (C.c_int, [VP, C.POINTER(VP), CP, SZ]),
    "tg_hip_stream_destroy": (C.c_int, [VP, VP]),
    "tg_hip_stream_synchronize": (C.c_int, [VP, VP, CP, SZ]),
    "tg_hip_event_create": (C.c_int, [VP, C.POINTER(VP), CP, SZ]),
    "tg_hip_event_destroy": (C.c_int, [VP, VP]),
    "tg_hip_event_record": (C.c_int, [VP, VP, VP, CP, SZ]),
    "tg_hip_event_elapsed": (C.c_int, [VP, VP, VP, C.POINTER(C.c_float), CP, SZ]),
    "tg_hip_host_alloc": (C.c_int, [VP, SZ, C.POINTER(VP), CP, SZ]),
    "tg_hip_host_free": (C.c_int, [VP, VP]),
    "tg_hip_memcpy_d2h_async": (C.c_int, [VP, VP, VP, SZ, VP, CP, SZ]),
    "tg_hip_memcpy_h2d_async": (C.c_int, [VP, VP, VP, SZ, VP, CP, SZ]),
    "tg_hip_event_query": (C.c_int, [VP, VP, CP, SZ]),
    "tg_hip_memcpy_d2d_async": (C.c_int, [VP, VP, VP, SZ, VP, CP, SZ]),
    "tg_hip_batch_tune": (C.c_int, [VP, u32, u32, u32, u32, CP, SZ]),
    "tg_hip_batch_surface": (C.c_int, [VP, C.c_double, CP, SZ]),
    "tg_hip_batch_stage_windows": (C.c_int, [VP, VP, SZ, u32, VP, VP, SZ, C.POINTER(SZ), CP,
                                             SZ]),
    "tg_hip_batch_read_bins": (C.c_int, [VP, VP, SZ, VP, SZ, C.POINTER(SZ), C.POINTER(SZ), CP,
                                         SZ]),
    "tg_hip_batch_cull_depth": (C.c_int, [VP, u32, CP, SZ]),
    "tg_hip_batch_run_static": (C.c_int, [VP, VP, SZ, VP, SZ, VP, SZ, VP, C.c_float, VP,
                                          C.c_double, VP, SZ, VP, CP, SZ]),
    "tg_hip_catalogue_frames": (C.c_int, [VP, VP, SZ, CP, SZ]),
    "tg_hip_boulder_catalogue_create": (C.c_int, [VP, VP, SZ, VP, VP, VP, CP, SZ]),
    "tg_hip_upsample": (C.c_int, [VP, C.POINTER(Upsample), VP, VP, VP, CP, SZ]),
    "tg_hip_boulder_catalogue_free": (C.c_int, [VP]),
    "tg_hip_batch_stage_boulders": (C.c_int, [VP, VP, VP, SZ, VP, SZ, C.c_int, CP, SZ]),
    "tg_hip_fill": (C.c_int, [VP, VP, C.c_float, SZ, VP, CP, SZ]),
    "tg_hip_boulder_set_create": (C.c_int, [VP, VP, VP, SZ, VP, SZ, VP, SZ, VP, SZ,
                                            C.POINTER(VP), CP, SZ]),
    "tg_hip_boulder_set_free": (C.c_int, [VP]),
    "tg_hip_crater_boulder_launch": (C.c_int, [VP, VP, VP, VP, VP, VP,
                                               C.POINTER(KernelParams), VP, CP, SZ]),
    "tg_hip_boulder_launch": (C.c_int, [VP, VP, VP, VP, VP, C.POINTER(KernelParams), VP, CP,
                                        SZ]),
    "tg_hip_valu_probe": (C.c_int, [VP, C.c_int, u32, u32, C.POINTER(C.c_float),
                                    C.POINTER(C.c_double), CP, SZ]),
    "tg_hip_store_probe": (C.c_int, [VP, C.c_int, u32, u32, VP, SZ, C.POINTER(C.c_float), CP,
                                     SZ]),
}
TG_WALK_TOGETHER, TG_WALK_PER_LANE, TG_WALK_CULLED = 0, 1, 2

_hip = None


def hip_library() -> C.CDLL:
    """The kernels' library; raises if it is missing (build it with ``python -m
    terraingen_amd.build`` or ``__graft_entry__.build()``)."""
    global _hip
    if _hip is None:
        if not HIP_SO.exists():
            raise RuntimeError(f"{HIP_SO} is missing: build the hip backend first")
        lib = C.CDLL(str(HIP_SO))
        for name, (res, args) in HIP_ABI.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _hip = lib
    return _hip


class HipError(RuntimeError):
    """An error the hip backend reported, with the runtime's message."""


def _check(code: int, err: C.Array) -> None:
    if code != 0:
        raise HipError(err.value.decode(errors="replace") or f"error {code}")


def _ptr(a: np.ndarray) -> int:
    return a.ctypes.data


class HipDevice:
    """One GPU opened through ``tg_hip_open``; the same object per ordinal for the process."""

    def __init__(self, ordinal: int = 0):
        lib = hip_library()
        err = C.create_string_buffer(512)
        handle = VP()
        _check(lib.tg_hip_open(ordinal, C.byref(handle), err, 512), err)
        self.lib = lib
        self.handle = handle
        self.ordinal = ordinal

    def call(self, name: str, *args) -> None:
        """Calls entry point ``name`` with ``args`` plus the error buffer; raises on failure."""
        err = C.create_string_buffer(512)
        _check(getattr(self.lib, name)(*args, err, 512), err)

    def synchronize(self, stream=None) -> None:
        self.call("tg_hip_stream_synchronize", self.handle, stream)

    @property
    def name(self) -> str:
        return self.lib.tg_hip_device_name(self.handle).decode()

    def run(self, kernel_id: int, kernel: C.Structure, heights: np.ndarray, xs: np.ndarray,
            ys: np.ndarray, params: KernelParams, extra: list[tuple[np.ndarray, int]] = ()) -> None:
        """``tg_hip_run`` on host arrays; ``heights`` (float32, C order) is updated in place."""
        assert heights.dtype == np.float32 and heights.flags.c_contiguous
        assert xs.dtype == np.float32 and ys.dtype == np.float32
        bufs = (TgBuf * max(1, len(extra)))()
        keep = []
        for i, (arr, per_cell) in enumerate(extra):
            arr = np.ascontiguousarray(arr)
            keep.append(arr)
            count = arr.shape[0] if arr.ndim > 0 else 1
            bufs[i] = TgBuf(arr.ctypes.data, arr.nbytes, count, per_cell)
        err = C.create_string_buffer(512)
        code = self.lib.tg_hip_run(self.handle, kernel_id, C.byref(kernel), C.sizeof(kernel),
                                   _ptr(heights), _ptr(xs), _ptr(ys), C.byref(params), bufs,
                                   len(extra), err, 512)
        _check(code, err)


def noise_kernel(fractals: list[dict]) -> NoiseLayerKernel:
    """A NoiseLayerKernel of the enabled ``fractals`` (MakeNoiseLayerKernel, noise_layer.h:41-50),
    each a dict of Fractal fields with the reference's defaults."""
    k = NoiseLayerKernel()
    for i in range(8):
        # Unused slots hold default-constructed fractals, as the reference's FractalStack does.
        k.fractals[i] = make_fractal()
    n = 0
    for f in fractals:
        fr = make_fractal(**f)
        if fr.enabled and n < 8:
            k.fractals[n] = fr
            n += 1
    k.count = n
    return k


def make_fractal(enabled=True, seed=0, wavelength=1000.0, amplitude=100.0, octaves=8,
                 lacunarity=2.0, gain=0.5, basis=0, type=0, offset=1.0, blend=0,
                 warp_amplitude=0.0, warp_wavelength=500.0) -> Fractal:
    """A Fractal with the reference's member defaults (fractals.h:29-44)."""
    return Fractal(int(bool(enabled)), seed, wavelength, amplitude, octaves, lacunarity, gain,
                   basis, type, offset, blend, warp_amplitude, warp_wavelength)


class DeviceArray:
    """Device memory from ``tg_hip_malloc``, freed with the object."""

    def __init__(self, dev: HipDevice, nbytes: int):
        self.dev = dev
        self.nbytes = int(nbytes)
        ptr = VP()
        dev.call("tg_hip_malloc", dev.handle, max(1, self.nbytes), C.byref(ptr))
        self.ptr = ptr.value

    @classmethod
    def from_host(cls, dev: HipDevice, arr: np.ndarray) -> "DeviceArray":
        arr = np.ascontiguousarray(arr)
        d = cls(dev, arr.nbytes)
        d.upload(arr)
        return d

    def upload(self, arr: np.ndarray) -> None:
        arr = np.ascontiguousarray(arr)
        assert arr.nbytes <= self.nbytes
        self.dev.call("tg_hip_memcpy_h2d", self.dev.handle, self.ptr, arr.ctypes.data, arr.nbytes)

    def download(self, out: np.ndarray, offset: int = 0) -> np.ndarray:
        """Copies `out.nbytes` bytes starting `offset` bytes into the array into `out`."""
        assert out.flags.c_contiguous and 0 <= offset and offset + out.nbytes <= self.nbytes
        self.dev.call("tg_hip_memcpy_d2h", self.dev.handle, out.ctypes.data, self.ptr + offset,
                      out.nbytes)
        return out

    def zero(self) -> None:
        self.dev.call("tg_hip_memset", self.dev.handle, self.ptr, 0, self.nbytes)

    def __del__(self):
        if getattr(self, "ptr", None):
            self.dev.lib.tg_hip_free(self.dev.handle, self.ptr)
            self.ptr = None


class Event:
    def __init__(self, dev: HipDevice):
        self.dev = dev
        e = VP()
        dev.call("tg_hip_event_create", dev.handle, C.byref(e))
        self.handle = e.value

    def record(self, stream=None) -> None:
        self.dev.call("tg_hip_event_record", self.dev.handle, self.handle, stream)

    def elapsed_ms(self, stop: "Event") -> float:
        ms = C.c_float()
        self.dev.call("tg_hip_event_elapsed", self.dev.handle, self.handle, stop.handle,
                      C.byref(ms))
        return ms.value

    def __del__(self):
        if getattr(self, "handle", None):
            self.dev.lib.tg_hip_event_destroy(self.dev.handle, self.handle)
            self.handle = None


# ---- The host library (include/terraingen_amd.h) ----------------------------------------------

Grid = _struct("Grid", [("width", u32), ("height", u32), ("origin_x", C.c_double),
                        ("origin_y", C.c_double), ("pixel", C.c_double)])

HOST_ABI = {
    "tgx_free": (None, [VP]),
    "tgx_outline_max_scale": (C.c_double, [C.POINTER(CraterRecord), VP]),
    "tgx_bin_plane": (C.c_int, [C.POINTER(Grid), VP, VP, VP, SZ, u32, C.POINTER(C.POINTER(u32)),
                                C.POINTER(SZ), C.POINTER(C.POINTER(u32)), C.POINTER(SZ), CP, SZ]),
}

_host = None


def host_library() -> C.CDLL:
    """The host library; raises if it is missing."""
    global _host
    if _host is None:
        if not HOST_SO.exists():
            raise RuntimeError(f"{HOST_SO} is missing: build it first")
        lib = C.CDLL(str(HOST_SO))
        for name, (res, args) in HOST_ABI.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _host = lib
    return _host


def north_up_grid(width: int, height: int, pixel: float, center=(0.0, 0.0)) -> Grid:
    """The grid FlatSource builds (src/sources/flat_source.cpp:23-25)."""
    return Grid(width, height, center[0] - 0.5 * width * pixel, center[1] + 0.5 * height * pixel,
                pixel)


def grid_centers(g: Grid) -> tuple[np.ndarray, np.ndarray]:
    """FP64 cell-centre x per column and y per row (raster.h:187-190)."""
    xs = g.origin_x + (np.arange(g.width) + 0.5) * g.pixel + 0.0
    ys = g.origin_y + (np.arange(g.height) + 0.5) * -g.pixel
    return xs, ys


def split(v: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
    """SplitCoordinate (crater_kernel.h:35-38): the float and the float of what it leaves."""
    hi = v.astype(np.float32)
    lo = (v - hi.astype(np.float64)).astype(np.float32)
    return hi, lo


def bin_plane(g: Grid, cx: np.ndarray, cy: np.ndarray, reach: np.ndarray, side: int = 16):
    """``tgx_bin_plane``: (tiles, entries) as uint32 arrays."""
    lib = host_library()
    cx, cy, reach = (np.ascontiguousarray(a, np.float64) for a in (cx, cy, reach))
    tp, ep = C.POINTER(u32)(), C.POINTER(u32)()
    tn, en = SZ(), SZ()
    err = C.create_string_buffer(512)
    code = lib.tgx_bin_plane(C.byref(g), cx.ctypes.data, cy.ctypes.data, reach.ctypes.data,
                             cx.size, side, C.byref(tp), C.byref(tn), C.byref(ep), C.byref(en),
                             err, 512)
    _check(code, err)
    try:
        tiles = np.ctypeslib.as_array(tp, (tn.value,)).copy()
        entries = np.ctypeslib.as_array(ep, (en.value,)).copy()
    finally:
        lib.tgx_free(tp)
        lib.tgx_free(ep)
    return tiles, entries


def record_reach(records: np.ndarray, data: np.ndarray, reach_radii: float) -> np.ndarray:
    """Each record's footprint half-width, MaxScale x reach radii x radius (prepared.h:31-33)."""
    lib = host_library()
    out = np.empty(records.size)
    rp = records.ctypes.data
    for i in range(records.size):
        rec = CraterRecord.from_address(rp + i * 144)
        out[i] = lib.tgx_outline_max_scale(C.byref(rec), data.ctypes.data) * reach_radii * (
            0.5 * float(records["diameter"][i]))
    return out


CRATER_DTYPE = np.dtype([("x", np.float64), ("y", np.float64), ("diameter", np.float64),
                         ("age_gyr", np.float64), ("id", np.uint64)])

Generation = _struct("Generation", [(n, C.c_double) for n in (
    "min_x", "min_y", "max_x", "max_y", "sfd_coefficient", "sfd_exponent", "min_diameter",
    "max_diameter", "margin_radii", "cell_size_factor", "min_resolved_diameter",
    "lifetime_gyr")] + [("seed", u32), ("model", u32), ("sphere_radius", C.c_double),
                        ("threads", u32), ("pad", u32)])

HOST_ABI["tgx_philox4x32"] = (None, [VP, VP, VP])
HOST_ABI["tgx_set_host_threads"] = (None, [u32])
HOST_ABI["tgx_generate_craters"] = (C.c_int, [C.POINTER(Generation), C.POINTER(VP),
                                              C.POINTER(SZ), CP, SZ])
HOST_ABI["tgx_generate_craters_regions"] = (C.c_int, [C.POINTER(Generation), VP, SZ,
                                                      C.POINTER(VP), C.POINTER(SZ), CP, SZ])


def generate_craters(spec: Generation, more_regions=None) -> np.ndarray:
    """``tgx_generate_craters``: a structured array of x, y, diameter, age_gyr, id; with
    `more_regions` (an (n, 4) array of boxes, on a sphere), ``tgx_generate_craters_regions``."""
    lib = host_library()
    ptr, n = VP(), SZ()
    err = C.create_string_buffer(512)
    if more_regions is not None and len(more_regions):
        boxes = np.ascontiguousarray(more_regions, np.float64).reshape(-1, 4)
        _check(lib.tgx_generate_craters_regions(C.byref(spec), boxes.ctypes.data,
                                                boxes.shape[0], C.byref(ptr), C.byref(n),
                                                err, 512), err)
    else:
        _check(lib.tgx_generate_craters(C.byref(spec), C.byref(ptr), C.byref(n), err, 512),
               err)
    try:
        out = _owned_copy(ptr, n.value * CRATER_DTYPE.itemsize).view(CRATER_DTYPE)
    finally:
        lib.tgx_free(ptr)
    return out


class GeoTiffSpec(C.Structure):
    _fields_ = [("path", C.c_char_p), ("width", u32), ("height", u32), ("bands", u32),
                ("compression", u32), ("tile_size", u32), ("big_tiff", u32),
                ("has_nodata", u32), ("nodata", C.c_float), ("transform", C.c_double * 6),
                ("geokeys", VP), ("geokey_count", SZ), ("geo_doubles", VP),
                ("geo_double_count", SZ), ("geo_ascii", C.c_char_p), ("threads", u32),
                ("pad", u32)]


CRATER_ATTRIBUTES_DTYPE = np.dtype([("floor_depth", np.float64), ("rim_height", np.float64),
                                    ("depth_ratio", np.float64), ("freshness", np.float64),
                                    ("complex_", np.uint8), ("pad", np.uint8, 7)])
assert CRATER_ATTRIBUTES_DTYPE.itemsize == 40
TIFF_COMPRESSION = {"none": 0, "lzw": 1, "deflate": 2}

HOST_ABI["tgx_write_geotiff"] = (C.c_int, [C.POINTER(GeoTiffSpec), VP, CP, SZ])
HOST_ABI["tgx_write_crater_csv"] = (C.c_int, [C.c_char_p, VP, SZ, VP, VP, VP, CP, SZ])


def north_up_transform(origin_x: float, origin_y: float, pixel: float) -> tuple:
    """GeoTransform::NorthUp in GDAL order: the north-west corner, square cells, rows south."""
    return (origin_x, pixel, 0.0, origin_y, 0.0, -pixel)


def sphere_geokeys(radius: float) -> tuple:
    """GeographicSphereKeys (src/raster/geotiff/geotiff.cpp:103-120): geographic longitude and
    latitude in degrees on a user-defined sphere of `radius` meters, pixels as areas, as
    (directory, doubles, ascii)."""
    if not radius > 0.0:
        raise ValueError("A sphere's radius must be positive")
    keys = [(1024, 0, 2), (1025, 0, 1), (2048, 0, 32767), (2050, 0, 32767), (2054, 0, 9102),
            (2056, 0, 32767), (2057, 34736, 0), (2058, 34736, 1)]
    directory = [1, 1, 0, len(keys)]
    for key, location, value in keys:
        directory += [key, location, 1, value]
    return directory, [radius, radius], ""


def write_geotiff(path, values: np.ndarray, transform=(0.0, 1.0, 0.0, 0.0, 0.0, 1.0),
                  nodata=None, geokeys=None, compression: str = "lzw", tile_size: int = 256,
                  big_tiff: bool = False, threads: int = 0) -> None:
    """``tgx_write_geotiff``: `values` (height x width, or bands x height x width, float32) as
    the reference's WriteGeoTiff writes them, with a GDAL-order `transform`, an optional no-data
    value and projection keys (directory, doubles, ascii)."""
    v = np.ascontiguousarray(values, np.float32)
    if v.ndim not in (2, 3):
        raise ValueError("values must be height x width or bands x height x width")
    directory, doubles, text = geokeys if geokeys is not None else ((), (), "")
    dk = np.ascontiguousarray(directory, np.uint16)
    dd = np.ascontiguousarray(doubles, np.float64)
    spec = GeoTiffSpec(path=str(path).encode(), width=v.shape[-1], height=v.shape[-2],
                       bands=1 if v.ndim == 2 else v.shape[0],
                       compression=TIFF_COMPRESSION[compression], tile_size=tile_size,
                       big_tiff=int(big_tiff), has_nodata=int(nodata is not None),
                       nodata=float(nodata) if nodata is not None else 0.0,
                       transform=(C.c_double * 6)(*transform),
                       geokeys=dk.ctypes.data if dk.size else None, geokey_count=dk.size,
                       geo_doubles=dd.ctypes.data if dd.size else None,
                       geo_double_count=dd.size, geo_ascii=text.encode() if text else None,
                       threads=threads)
    err = C.create_string_buffer(512)
    _check(host_library().tgx_write_geotiff(C.byref(spec), v.ctypes.data, err, 512), err)


def write_crater_csv(path, craters: np.ndarray, manual=None, replace=None,
                     attributes=None) -> None:
    """``tgx_write_crater_csv``: a CRATER_DTYPE list as the reference's crater list CSV, with
    optional manual/replace flags and CRATER_ATTRIBUTES_DTYPE attributes."""
    c = np.ascontiguousarray(craters, CRATER_DTYPE)
    flags = [np.ascontiguousarray(a, np.uint8) if a is not None else None
             for a in (manual, replace)]
    attrs = (np.ascontiguousarray(attributes, CRATER_ATTRIBUTES_DTYPE)
             if attributes is not None else None)
    for a in (*flags, attrs):
        if a is not None and a.size != c.size:
            raise ValueError("flags and attributes must have one entry per crater")
    err = C.create_string_buffer(512)
    _check(host_library().tgx_write_crater_csv(
        str(path).encode(), c.ctypes.data, c.size,
        *[a.ctypes.data if a is not None else None for a in flags],
        attrs.ctypes.data if attrs is not None else None, err, 512), err)


def generate_sphere_craters(radius: float, region: tuple, sfd_coefficient: float,
                            sfd_exponent: float, min_diameter: float, max_diameter: float,
                            margin_radii: float = 0.0, min_resolved_diameter: float = 0.0,
                            lifetime_gyr: float = 1.0, seed: int = 1, model: int = 0,
                            threads: int = 0) -> np.ndarray:
    """``tgx_generate_craters`` on a sphere of `radius` meters over a longitude-latitude
    `region` (min_lon, min_lat, max_lon, max_lat, degrees), or the union over several (an (n, 4)
    array): the reference's cube-sphere placement; x and y of the list are longitude and
    latitude."""
    boxes = np.asarray(region, np.float64).reshape(-1, 4)
    lon0, lat0, lon1, lat1 = (float(v) for v in boxes[0])
    return generate_craters(Generation(
        min_x=lon0, min_y=lat0, max_x=lon1, max_y=lat1, sfd_coefficient=sfd_coefficient,
        sfd_exponent=sfd_exponent, min_diameter=min_diameter, max_diameter=max_diameter,
        margin_radii=margin_radii, cell_size_factor=16.0,
        min_resolved_diameter=min_resolved_diameter, lifetime_gyr=lifetime_gyr, seed=seed,
        model=model, sphere_radius=radius, threads=threads), boxes[1:])


CraterModels = _struct("CraterModels", [("profile", u32), ("complex_", u32),
                                        ("transition_diameter", C.c_double),
                                        ("lifetime_gyr", C.c_double), ("model", u32),
                                        ("pad", u32)])
# tgx_crater_models.model / tgx_generation.model: 1 selects the reference's DefaultCraterModels.
MODELS_DEFAULT = 1
Apply = _struct("Apply", [("irregular", u32), ("peaks", u32), ("terraces", u32)] + [
    (n, C.c_double) for n in ("rim_sigma", "noise_amplitude", "noise_rim_amplitude",
                              "wall_streaks", "ejecta", "terrace_onset", "terrace_full")])

HOST_ABI["tgx_margin_radii"] = (C.c_double, [C.POINTER(CraterModels), C.POINTER(Apply)])
HOST_ABI["tgx_prepare_craters"] = (C.c_int, [
    VP, SZ, C.POINTER(CraterModels), C.POINTER(Apply), u32, C.c_double, C.POINTER(CraterShape),
    C.POINTER(VP), C.POINTER(SZ), C.POINTER(VP), C.POINTER(SZ), C.POINTER(VP), CP, SZ])


def _owned_copy(ptr: VP, nbytes: int) -> np.ndarray:
    """One copy of `nbytes` at a library-allocated pointer into a new byte array."""
    out = np.empty(nbytes, np.uint8)
    if nbytes:
        C.memmove(out.ctypes.data, ptr.value, nbytes)
    return out


def prepare_craters(craters: np.ndarray, models, apply, layer_seed: int, spacing: float):
    """``tgx_prepare_craters``: (kernel, records, data, footprint[n, 4]: x, y, reach, D)."""
    lib = host_library()
    craters = np.ascontiguousarray(craters, CRATER_DTYPE)
    kernel = CraterShape()
    rp, dp, fp = VP(), VP(), VP()
    rn, dn = SZ(), SZ()
    err = C.create_string_buffer(512)
    _check(lib.tgx_prepare_craters(craters.ctypes.data, craters.size, C.byref(models),
                                   C.byref(apply), layer_seed, spacing, C.byref(kernel),
                                   C.byref(rp), C.byref(rn), C.byref(dp), C.byref(dn),
                                   C.byref(fp), err, 512), err)
    try:
        records = _owned_copy(rp, rn.value * 144).view(RECORD_DTYPE)
        data = _owned_copy(dp, dn.value * 4).view(np.float32)
        foot = _owned_copy(fp, rn.value * 32).view(np.float64).reshape(-1, 4)
    finally:
        for p in (rp, dp, fp):
            lib.tgx_free(p)
    return kernel, records, data, foot


def margin_radii(models, apply) -> float:
    return host_library().tgx_margin_radii(C.byref(models), C.byref(apply))


TILE_DESC_DTYPE = np.dtype([("origin_x", np.float64), ("origin_y", np.float64),
                            ("pixel", np.float64), ("width", np.uint32), ("height", np.uint32),
                            ("out_offset", np.uint64), ("bin_first", np.uint32),
                            ("bins_across", np.uint32)])
BIN_DTYPE = np.dtype([("tile", np.uint32), ("first", np.uint32), ("count", np.uint32),
                      ("pad", np.uint32)])
TILE_DTYPE = np.dtype([("origin_x", np.float64), ("origin_y", np.float64), ("pixel", np.float64),
                       ("width", np.uint32), ("height", np.uint32)])
assert TILE_DESC_DTYPE.itemsize == 48 and BIN_DTYPE.itemsize == 16 and TILE_DTYPE.itemsize == 32

HOST_ABI["tgx_batch_bin"] = (C.c_int, [VP, VP, VP, VP, SZ, VP, SZ, C.c_uint, C.POINTER(VP),
                                       C.POINTER(VP), C.POINTER(SZ), C.POINTER(VP), C.POINTER(SZ),
                                       CP, SZ])


def batch_bin(foot: np.ndarray, tiles: np.ndarray, threads: int = 0):
    """``tgx_batch_bin``: (tile descriptors, bins, entries) for a catalogue footprint
    (x, y, reach, D per crater, apply order) and a structured array of tiles."""
    lib = host_library()
    foot = np.ascontiguousarray(foot, np.float64)
    cx, cy, reach, d = (np.ascontiguousarray(foot[:, i]) for i in range(4))
    tiles = np.ascontiguousarray(tiles, TILE_DTYPE)
    dp, bp, ep = VP(), VP(), VP()
    bn, en = SZ(), SZ()
    err = C.create_string_buffer(512)
    _check(lib.tgx_batch_bin(cx.ctypes.data, cy.ctypes.data, reach.ctypes.data, d.ctypes.data,
                             cx.size, tiles.ctypes.data, tiles.size, threads, C.byref(dp),
                             C.byref(bp), C.byref(bn), C.byref(ep), C.byref(en), err, 512), err)
    try:
        descs = _owned_copy(dp, tiles.size * 48).view(TILE_DESC_DTYPE)
        bins = _owned_copy(bp, bn.value * 16).view(BIN_DTYPE)
        entries = np.frombuffer(C.string_at(ep.value, max(en.value, 1) * 4), np.uint32)[
            :en.value].copy()
    finally:
        for p in (dp, bp, ep):
            lib.tgx_free(p)
    return descs, bins, entries


HOST_ABI["tgx_batch_pair_ops"] = (C.c_int, [VP, SZ, VP, SZ, VP, SZ, VP, VP, VP, VP, SZ, C.c_uint,
                                            C.POINTER(C.c_uint64), C.POINTER(C.c_uint64),
                                            C.POINTER(C.c_double), CP, SZ])


def batch_pair_ops(descs, bins, entries, foot: np.ndarray, cost: np.ndarray, threads: int = 0):
    """``tgx_batch_pair_ops``: (pairs listed, pairs inside the footprint, summed cost of those)
    for a binned batch, with `cost` one value per catalogue crater."""
    lib = host_library()
    foot = np.ascontiguousarray(foot, np.float64)
    cx, cy, reach = (np.ascontiguousarray(foot[:, i]) for i in range(3))
    cost = np.ascontiguousarray(cost, np.float64)
    if cost.size != cx.size:
        raise ValueError("one cost per catalogue crater")
    descs = np.ascontiguousarray(descs, TILE_DESC_DTYPE)
    bins = np.ascontiguousarray(bins, BIN_DTYPE)
    entries = np.ascontiguousarray(entries, np.uint32)
    listed, inside, ops = C.c_uint64(), C.c_uint64(), C.c_double()
    err = C.create_string_buffer(512)
    _check(lib.tgx_batch_pair_ops(descs.ctypes.data, descs.size, bins.ctypes.data, bins.size,
                                  entries.ctypes.data, entries.size, cx.ctypes.data,
                                  cy.ctypes.data, reach.ctypes.data, cost.ctypes.data, cx.size,
                                  threads, C.byref(listed), C.byref(inside), C.byref(ops), err,
                                  512), err)
    return int(listed.value), int(inside.value), float(ops.value)


HOST_ABI["tgx_index_create"] = (C.c_int, [VP, VP, VP, VP, SZ, C.POINTER(VP), CP, SZ])
HOST_ABI["tgx_index_free"] = (None, [VP])
HOST_ABI["tgx_sphere_index_create"] = (C.c_int, [VP, VP, VP, VP, SZ, C.c_double, C.POINTER(VP),
                                                 CP, SZ])
HOST_ABI["tgx_index_count"] = (SZ, [VP])
HOST_ABI["tgx_index_bin"] = (C.c_int, [VP, VP, SZ, C.c_uint, C.POINTER(VP), C.POINTER(VP),
                                       C.POINTER(SZ), C.POINTER(VP), C.POINTER(SZ), CP, SZ])
HOST_ABI["tgx_index_windows"] = (C.c_int, [VP, VP, SZ, u32, C.c_uint, C.POINTER(VP),
                                           C.POINTER(VP), C.POINTER(VP), C.POINTER(SZ), CP, SZ])
WINDOW_DTYPE = np.dtype([("crater", np.uint32), ("col0", np.uint32), ("col1", np.uint32),
                         ("row0", np.uint32), ("row1", np.uint32)])


def _take_batch(lib, tiles_n, dp, bp, bn, ep, en):
    try:
        descs = _owned_copy(dp, tiles_n * 48).view(TILE_DESC_DTYPE)
        bins = _owned_copy(bp, bn.value * 16).view(BIN_DTYPE)
        entries = np.frombuffer(C.string_at(ep.value, max(en.value, 1) * 4), np.uint32)[
            :en.value].copy()
    finally:
        for p in (dp, bp, ep):
            lib.tgx_free(p)
    return descs, bins, entries


HOST_ABI["tgx_sphere_positions"] = (C.c_int, [C.c_double, C.c_double, C.c_double, u32, u32,
                                              C.c_double, VP, CP, SZ])
HOST_ABI["tgx_sphere_split_points"] = (C.c_int, [C.c_double, C.c_double, C.c_double, u32, u32,
                                                 C.c_double, VP, CP, SZ])
HOST_ABI["tgx_sphere_frames"] = (C.c_int, [VP, VP, SZ, C.c_double, VP, CP, SZ])
HOST_ABI["tgx_sphere_bin"] = (C.c_int, [VP, VP, VP, VP, SZ, C.c_double, VP, SZ, C.c_uint,
                                        C.POINTER(VP), C.POINTER(VP), C.POINTER(SZ),
                                        C.POINTER(VP), C.POINTER(SZ), CP, SZ])


def sphere_positions(g: Grid, radius: float) -> np.ndarray:
    """``tgx_sphere_positions``: the body-fixed FP64 centres of a longitude-latitude grid's cells,
    (height * width, 3)."""
    out = np.empty((g.height * g.width, 3), np.float64)
    err = C.create_string_buffer(512)
    _check(host_library().tgx_sphere_positions(g.origin_x, g.origin_y, g.pixel, g.width,
                                               g.height, radius, out.ctypes.data, err, 512), err)
    return out


def sphere_split_points(g: Grid, radius: float) -> np.ndarray:
    """``tgx_sphere_split_points``: MakeSplitPoint of each cell centre, POINT_DTYPE."""
    out = np.empty(g.height * g.width, POINT_DTYPE)
    err = C.create_string_buffer(512)
    _check(host_library().tgx_sphere_split_points(g.origin_x, g.origin_y, g.pixel, g.width,
                                                  g.height, radius, out.ctypes.data, err, 512),
           err)
    return out


def sphere_frames(lon: np.ndarray, lat: np.ndarray, radius: float) -> np.ndarray:
    """``tgx_sphere_frames``: MakeCraterFrame of craters at (lon, lat) degrees, FRAME_DTYPE."""
    lon, lat = (np.ascontiguousarray(a, np.float64) for a in (lon, lat))
    out = np.empty(max(lon.size, 1), FRAME_DTYPE)
    err = C.create_string_buffer(512)
    _check(host_library().tgx_sphere_frames(lon.ctypes.data, lat.ctypes.data, lon.size, radius,
                                            out.ctypes.data, err, 512), err)
    return out[:lon.size]


def sphere_bin(foot: np.ndarray, radius: float, tiles: np.ndarray, threads: int = 0):
    """``tgx_sphere_bin``: (tile descriptors, bins, entries) for a catalogue footprint (lon, lat,
    reach in meters of arc, D per crater, apply order) and tiles in degrees."""
    lib = host_library()
    foot = np.ascontiguousarray(foot, np.float64)
    lon, lat, reach, d = (np.ascontiguousarray(foot[:, i]) for i in range(4))
    tiles = np.ascontiguousarray(tiles, TILE_DTYPE)
    dp, bp, ep = VP(), VP(), VP()
    bn, en = SZ(), SZ()
    err = C.create_string_buffer(512)
    _check(lib.tgx_sphere_bin(lon.ctypes.data, lat.ctypes.data, reach.ctypes.data, d.ctypes.data,
                              lon.size, radius, tiles.ctypes.data, tiles.size, threads,
                              C.byref(dp), C.byref(bp), C.byref(bn), C.byref(ep), C.byref(en),
                              err, 512), err)
    return _take_batch(lib, tiles.size, dp, bp, bn, ep, en)


HOST_ABI["tgx_sphere_pair_ops"] = (C.c_int, [VP, SZ, VP, SZ, VP, SZ, VP, VP, VP, VP, SZ,
                                             C.c_double, C.c_uint, C.POINTER(C.c_uint64),
                                             C.POINTER(C.c_uint64), C.POINTER(C.c_double), CP,
                                             SZ])


def sphere_pair_ops(descs, bins, entries, foot: np.ndarray, cost: np.ndarray, radius: float,
                    threads: int = 0):
    """``tgx_sphere_pair_ops``: (pairs listed, pairs within reach along the sphere, summed cost
    of those) for a batch binned by ``sphere_bin``."""
    lib = host_library()
    foot = np.ascontiguousarray(foot, np.float64)
    lon, lat, reach = (np.ascontiguousarray(foot[:, i]) for i in range(3))
    cost = np.ascontiguousarray(cost, np.float64)
    if cost.size != lon.size:
        raise ValueError("one cost per catalogue crater")
    descs = np.ascontiguousarray(descs, TILE_DESC_DTYPE)
    bins = np.ascontiguousarray(bins, BIN_DTYPE)
    entries = np.ascontiguousarray(entries, np.uint32)
    listed, inside, ops = C.c_uint64(), C.c_uint64(), C.c_double()
    err = C.create_string_buffer(512)
    _check(lib.tgx_sphere_pair_ops(descs.ctypes.data, descs.size, bins.ctypes.data, bins.size,
                                   entries.ctypes.data, entries.size, lon.ctypes.data,
                                   lat.ctypes.data, reach.ctypes.data, cost.ctypes.data,
                                   lon.size, radius, threads, C.byref(listed), C.byref(inside),
                                   C.byref(ops), err, 512), err)
    return int(listed.value), int(inside.value), float(ops.value)


class CatalogueIndex:
    """``tgx_index``: a catalogue footprint (x, y, reach, D per crater, apply order) indexed once
    for binning many batches of tiles."""

    def __init__(self, foot: np.ndarray, radius: float = 0.0):
        """On a sphere of `radius` meters (``tgx_sphere_index_create``) the footprint's x and y
        are longitude and latitude in degrees and its reach is in meters of arc."""
        lib = host_library()
        foot = np.ascontiguousarray(foot, np.float64)
        cols = [np.ascontiguousarray(foot[:, i]) for i in range(4)]
        h = VP()
        err = C.create_string_buffer(512)
        if radius > 0.0:
            _check(lib.tgx_sphere_index_create(*(c.ctypes.data for c in cols), cols[0].size,
                                               radius, C.byref(h), err, 512), err)
        else:
            _check(lib.tgx_index_create(*(c.ctypes.data for c in cols), cols[0].size,
                                        C.byref(h), err, 512), err)
        self.handle = h.value
        self.count = cols[0].size
        self.radius = radius

    def bin(self, tiles: np.ndarray, threads: int = 0):
        """(tile descriptors, bins, entries) of a structured array of tiles."""
        lib = host_library()
        tiles = np.ascontiguousarray(tiles, TILE_DTYPE)
        dp, bp, ep = VP(), VP(), VP()
        bn, en = SZ(), SZ()
        err = C.create_string_buffer(512)
        _check(lib.tgx_index_bin(self.handle, tiles.ctypes.data, tiles.size, threads,
                                 C.byref(dp), C.byref(bp), C.byref(bn), C.byref(ep),
                                 C.byref(en), err, 512), err)
        return _take_batch(lib, tiles.size, dp, bp, bn, ep, en)

    def windows(self, tiles: np.ndarray, bin_side: int = 16, threads: int = 0):
        """``tgx_index_windows``: (descs for bins of `bin_side` cells, per-tile (first, count) of
        windows, WINDOW_DTYPE windows), for binning on the device."""
        lib = host_library()
        tiles = np.ascontiguousarray(tiles, TILE_DTYPE)
        dp, tp, wp = VP(), VP(), VP()
        wn = SZ()
        err = C.create_string_buffer(512)
        _check(lib.tgx_index_windows(self.handle, tiles.ctypes.data, tiles.size, bin_side,
                                     threads, C.byref(dp), C.byref(tp), C.byref(wp),
                                     C.byref(wn), err, 512), err)
        try:
            descs = _owned_copy(dp, tiles.size * 48).view(TILE_DESC_DTYPE)
            tile_windows = _owned_copy(tp, tiles.size * 8).view(np.uint32)
            windows = _owned_copy(wp, wn.value * 20).view(WINDOW_DTYPE)
        finally:
            for p in (dp, tp, wp):
                lib.tgx_free(p)
        return descs, tile_windows, windows

    def __del__(self):
        if getattr(self, "handle", None):
            host_library().tgx_index_free(self.handle)
            self.handle = None


# ---- The tile server (tgx_server_*) -----------------------------------------------------------

ServerConfig = _struct("ServerConfig", [("slots", u32), ("max_batch_tiles", u32),
                                        ("max_batch_cells", C.c_uint64),
                                        ("pool_bytes", C.c_uint64), ("bin_threads", u32),
                                        ("poll_microseconds", u32)])
TicketInfo = _struct("TicketInfo", [("submitted", C.c_double), ("admitted", C.c_double),
                                    ("launched", C.c_double), ("done", C.c_double),
                                    ("batch", u32), ("batch_tiles", u32)])
ServerStats = _struct("ServerStats", [("batches", C.c_uint64), ("tiles", C.c_uint64),
                                      ("cells", C.c_uint64), ("failed", C.c_uint64),
                                      ("max_in_flight", u32), ("pad", u32),
                                      ("pool_bytes", C.c_uint64),
                                      ("pool_peak_bytes", C.c_uint64),
                                      ("pending", C.c_uint64)])
TileDesc = _struct("TileDesc", [("origin_x", C.c_double), ("origin_y", C.c_double),
                                ("pixel", C.c_double), ("width", u32), ("height", u32)])

HOST_ABI["tgx_server_create"] = (C.c_int, [VP, VP, C.c_float, VP, VP, C.POINTER(ServerConfig),
                                           C.POINTER(VP), CP, SZ])
HOST_ABI["tgx_server_boulders"] = (C.c_int, [VP, VP, VP, CP, SZ])
HOST_ABI["tgx_server_submit"] = (C.c_int, [VP, C.POINTER(TileDesc), VP, C.POINTER(C.c_uint64),
                                           CP, SZ])
HOST_ABI["tgx_server_wait"] = (C.c_int, [VP, C.c_uint64, C.c_double, C.POINTER(TicketInfo), CP,
                                         SZ])
HOST_ABI["tgx_server_stats_get"] = (C.c_int, [VP, C.POINTER(ServerStats)])
HOST_ABI["tgx_server_free"] = (C.c_int, [VP])
HOST_ABI["tgx_hip_unavailable_reason"] = (C.c_char_p, [])

TG_NOT_READY = 4


# ---- The backend registry and RunKernel (backend.cpp) ------------------------------------------

BACKENDS = ("auto", "cpu", "vulkan", "gpu", "cuda", "metal", "hip")  # the enum's order
HOST_ABI["tgx_backend_named"] = (C.c_int, [CP, C.POINTER(C.c_int)])
HOST_ABI["tgx_backend_name"] = (SZ, [C.c_int, CP, SZ])
HOST_ABI["tgx_backends_report"] = (SZ, [CP, SZ])
HOST_ABI["tgx_default_backend"] = (C.c_int, [C.POINTER(C.c_int), CP, SZ])
HOST_ABI["tgx_set_default_backend"] = (None, [C.c_int])
HOST_ABI["tgx_resolve_backend"] = (C.c_int, [C.c_int, C.POINTER(C.c_int), CP, SZ])
HOST_ABI["tgx_run_kernel"] = (C.c_int, [C.c_int, C.c_int, VP, SZ, C.POINTER(Grid), VP,
                                        C.POINTER(C.c_float), C.c_double, C.c_double, VP, SZ, SZ,
                                        C.POINTER(C.c_int), CP, SZ])


def backend_named(name: str):
    """BackendNamed: the backend's enum value, or None for an unknown name."""
    b = C.c_int()
    code = host_library().tgx_backend_named(name.encode(), C.byref(b))
    return b.value if code == 0 else None


def backend_name(backend: int) -> str:
    out = C.create_string_buffer(64)
    host_library().tgx_backend_name(backend, out, 64)
    return out.value.decode()


def backends_report() -> str:
    lib = host_library()
    n = lib.tgx_backends_report(None, 0)
    out = C.create_string_buffer(n + 1)
    lib.tgx_backends_report(out, n + 1)
    return out.value.decode()


def default_backend() -> int:
    b = C.c_int()
    err = C.create_string_buffer(512)
    code = host_library().tgx_default_backend(C.byref(b), err, 512)
    if code:
        raise ValueError(err.value.decode())
    return b.value


def set_default_backend(backend: int) -> None:
    host_library().tgx_set_default_backend(backend)


def resolve_backend(requested: int) -> int:
    b = C.c_int()
    err = C.create_string_buffer(1024)
    code = host_library().tgx_resolve_backend(requested, C.byref(b), err, 1024)
    if code:
        raise RuntimeError(err.value.decode())
    return b.value


def run_kernel(backend: int, program: int, kernel, grid: Grid, values: np.ndarray,
               origin=(0.0, 0.0), extra=(), no_data=None, max_strip_bytes: int = 0) -> bool:
    """RunKernel: runs `program` (a TG_KERNEL_* id, 0 for none) with the ctypes kernel struct over
    `values` (float32, height x width, updated in place). `extra` holds (array, per_cell) pairs."""
    assert values.dtype == np.float32 and values.flags.c_contiguous
    assert values.size == grid.width * grid.height
    keep = [np.ascontiguousarray(a) for a, _ in extra]
    bufs = (TgBuf * max(1, len(keep)))()
    for i, (a, (_, per_cell)) in enumerate(zip(keep, extra)):
        bufs[i] = TgBuf(a.ctypes.data, a.nbytes, a.shape[0] if a.ndim else 1, per_cell)
    ran = C.c_int()
    err = C.create_string_buffer(1024)
    nd = C.c_float(no_data) if no_data is not None else None
    code = host_library().tgx_run_kernel(
        backend, program, C.cast(C.byref(kernel), VP) if kernel is not None else None,
        C.sizeof(kernel) if kernel is not None else 0, C.byref(grid), values.ctypes.data,
        C.byref(nd) if nd is not None else None, origin[0], origin[1], C.cast(bufs, VP),
        len(keep), max_strip_bytes, C.byref(ran), err, 1024)
    if code:
        raise RuntimeError(err.value.decode())
    return bool(ran.value)


# ---- Boulders (tgx_prepare_boulders) -----------------------------------------------------------

BOULDER_DTYPE = np.dtype([("x", np.float64), ("y", np.float64), ("diameter", np.float64),
                          ("exposed", np.float64), ("facets", np.float64, 6),
                          ("facet_count", np.uint32), ("model", np.uint32)])
BOULDER_RECORD_DTYPE = np.dtype([("x", np.float32), ("y", np.float32), ("x_low", np.float32),
                                 ("y_low", np.float32), ("radius", np.float32),
                                 ("exposed", np.float32), ("reach", np.float32),
                                 ("weight", np.float32), ("facet", np.float32, 8),
                                 ("facet_count", np.uint32), ("pad", np.uint32, 3)])
assert BOULDER_DTYPE.itemsize == 88 and BOULDER_RECORD_DTYPE.itemsize == 80
BoulderKernel = _struct("BoulderKernel", [("width", f32), ("count", u32), ("pad", u32 * 2)])
HOST_ABI["tgx_prepare_boulders"] = (C.c_int, [VP, SZ, C.c_double, C.c_int,
                                              C.POINTER(BoulderKernel), C.POINTER(VP),
                                              C.POINTER(SZ), C.POINTER(VP), CP, SZ])


BoulderLayerSpec = _struct("BoulderLayerSpec", [(n, C.c_double) for n in (
    "min_x", "min_y", "max_x", "max_y", "sphere_radius", "spacing", "sfd_coefficient",
    "sfd_exponent", "min_d", "max_d", "min_height_ratio", "max_height_ratio", "min_burial",
    "max_burial", "max_lean_degrees")] + [("seed", u32), ("threads", u32)])
PLACED_BOULDER_DTYPE = np.dtype([(n, np.float64) for n in (
    "x", "y", "diameter", "height", "burial", "yaw", "lean", "wear")] +
    [("id", np.uint64), ("model", np.uint32), ("pad", np.uint32)])
assert PLACED_BOULDER_DTYPE.itemsize == 80
HOST_ABI["tgx_place_boulders"] = (C.c_int, [C.POINTER(BoulderLayerSpec), C.POINTER(VP),
                                            C.POINTER(SZ), CP, SZ])
# BoulderPose's defaults (boulder_options.h:247-252): lean up to 0.26 rad.
BOULDER_POSE_DEFAULTS = dict(min_height_ratio=0.4, max_height_ratio=0.6, min_burial=0.1,
                             max_burial=0.4, max_lean_degrees=0.26 * 180.0 / np.pi)


def place_boulders(region: tuple, spacing: float, sfd_coefficient: float, sfd_exponent: float,
                   min_d: float, max_d: float, seed: int = 1, sphere_radius: float = 0.0,
                   threads: int = 0, **pose) -> np.ndarray:
    """``tgx_place_boulders``: the boulders a power-law BoulderLayer lists over `region`
    (min_x, min_y, max_x, max_y; degrees on a sphere) for a grid of `spacing` meters, as the
    reference lists them, in PLACED_BOULDER_DTYPE."""
    spec = BoulderLayerSpec(*region, sphere_radius, spacing, sfd_coefficient, sfd_exponent,
                            min_d, max_d, **(BOULDER_POSE_DEFAULTS | pose), seed=seed,
                            threads=threads)
    lib = host_library()
    ptr, n = VP(), SZ()
    err = C.create_string_buffer(512)
    _check(lib.tgx_place_boulders(C.byref(spec), C.byref(ptr), C.byref(n), err, 512), err)
    try:
        return _owned_copy(ptr, n.value * 80).view(PLACED_BOULDER_DTYPE)
    finally:
        lib.tgx_free(ptr)


HOST_ABI["tgx_place_boulders_boxes"] = (C.c_int, [C.POINTER(BoulderLayerSpec), VP, SZ,
                                                  C.POINTER(VP), C.POINTER(SZ), CP, SZ])


def place_boulders_boxes(boxes: np.ndarray, spacing: float, sfd_coefficient: float,
                         sfd_exponent: float, min_d: float, max_d: float, seed: int = 1,
                         sphere_radius: float = 0.0, threads: int = 0, **pose) -> np.ndarray:
    """``tgx_place_boulders_boxes``: the boulders a power-law BoulderLayer lists over any of
    `boxes` ([n, 4]: min_x, min_y, max_x, max_y) for `spacing`, each once."""
    spec = BoulderLayerSpec(0.0, 0.0, 0.0, 0.0, sphere_radius, spacing, sfd_coefficient,
                            sfd_exponent, min_d, max_d, **(BOULDER_POSE_DEFAULTS | pose),
                            seed=seed, threads=threads)
    boxes = np.ascontiguousarray(boxes, np.float64).reshape(-1, 4)
    lib = host_library()
    ptr, n = VP(), SZ()
    err = C.create_string_buffer(512)
    _check(lib.tgx_place_boulders_boxes(C.byref(spec), boxes.ctypes.data, boxes.shape[0],
                                        C.byref(ptr), C.byref(n), err, 512), err)
    try:
        return _owned_copy(ptr, n.value * 80).view(PLACED_BOULDER_DTYPE)
    finally:
        lib.tgx_free(ptr)


GeoTiffInfo = _struct("GeoTiffInfo", [(n, u32) for n in (
    "width", "height", "bands", "bits_per_sample", "sample_format", "compression", "predictor",
    "tiled", "chunk_width", "chunk_height", "big_tiff", "big_endian", "georeferenced",
    "geographic", "pixel_is_point", "has_no_data")] + [("no_data", C.c_double),
                                                       ("transform", C.c_double * 6)])
assert C.sizeof(GeoTiffInfo) == 120
HOST_ABI["tgx_geotiff_open"] = (C.c_int, [C.c_char_p, C.POINTER(VP), CP, SZ])
HOST_ABI["tgx_geotiff_info_get"] = (C.c_int, [VP, u32, C.POINTER(GeoTiffInfo),
                                              C.POINTER(C.c_double), C.POINTER(C.c_double)])
HOST_ABI["tgx_geotiff_read"] = (C.c_int, [VP, u32, C.c_uint64, C.c_uint64, u32, u32, C.c_double,
                                          C.c_double, VP, CP, SZ])
HOST_ABI["tgx_geotiff_close"] = (None, [VP])


class GeoTiffFile:
    """``tgx_geotiff_*``: a GeoTIFF's layout and windows of one band as floats."""

    def __init__(self, path):
        lib = host_library()
        h = VP()
        err = C.create_string_buffer(512)
        _check(lib.tgx_geotiff_open(str(path).encode(), C.byref(h), err, 512), err)
        self.handle = h.value

    def info(self, band: int = 0):
        """(GeoTiffInfo, the band's scale, its offset)."""
        info = GeoTiffInfo()
        scale, offset = C.c_double(), C.c_double()
        if host_library().tgx_geotiff_info_get(self.handle, band, C.byref(info), C.byref(scale),
                                               C.byref(offset)):
            raise ValueError(f"band {band} is missing")
        return info, scale.value, offset.value

    def read(self, band: int, col0: int, row0: int, width: int, height: int,
             scale: float = 1.0, offset: float = 0.0) -> np.ndarray:
        out = np.zeros((height, width), np.float32)
        err = C.create_string_buffer(512)
        _check(host_library().tgx_geotiff_read(self.handle, band, col0, row0, width, height,
                                                scale, offset, out.ctypes.data, err, 512), err)
        return out

    def __del__(self):
        if getattr(self, "handle", None):
            host_library().tgx_geotiff_close(self.handle)
            self.handle = None


def relief_boulders(placed: np.ndarray) -> np.ndarray:
    """BOULDER_DTYPE of placed boulders for the dome relief: centre, diameter, exposed height
    (Boulder::ExposedHeight, height times one less the burial), model flag, no facets."""
    out = np.zeros(placed.size, BOULDER_DTYPE)
    out["x"], out["y"], out["diameter"] = placed["x"], placed["y"], placed["diameter"]
    out["exposed"] = placed["height"] * (1.0 - placed["burial"])
    out["model"] = placed["model"]
    return out


def prepare_boulders(boulders: np.ndarray, spacing: float, dome: bool = True):
    """``tgx_prepare_boulders``: (kernel block, records, footprint[n, 4]: x, y, reach, D)."""
    lib = host_library()
    boulders = np.ascontiguousarray(boulders, BOULDER_DTYPE)
    kernel = BoulderKernel()
    rp, fp = VP(), VP()
    n = SZ()
    err = C.create_string_buffer(512)
    _check(lib.tgx_prepare_boulders(boulders.ctypes.data, boulders.size, spacing, int(dome),
                                    C.byref(kernel), C.byref(rp), C.byref(n), C.byref(fp),
                                    err, 512), err)
    try:
        records = np.frombuffer(C.string_at(rp.value, max(n.value, 1) * 80),
                                BOULDER_RECORD_DTYPE)[:n.value].copy()
        foot = np.frombuffer(C.string_at(fp.value, max(n.value, 1) * 32), np.float64)[
            :4 * n.value].reshape(n.value, 4).copy()
    finally:
        lib.tgx_free(rp)
        lib.tgx_free(fp)
    return kernel, records, foot
