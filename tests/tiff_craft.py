# SPDX-License-Identifier: BSD-3-Clause
"""Small GeoTIFF files for the reader's tests, written here field by field (TIFF 6.0, BigTIFF,
GeoTIFF 1.1 and GDAL's no-data and metadata tags), so each layout a DEM may come in is covered:
either byte order, classic or BigTIFF, strips or tiles, interleaved or planar samples, integer
and float samples of 8 to 64 bits, no compression, Deflate or PackBits, the horizontal and the
floating-point predictors."""
from __future__ import annotations

import struct
import zlib
from dataclasses import dataclass, field

import numpy as np

SHORT, LONG, DOUBLE, ASCII, LONG8 = 3, 4, 12, 2, 16


@dataclass
class Spec:
    values: np.ndarray                 # (bands, height, width)
    big_endian: bool = False
    big_tiff: bool = False
    tile: int = 0                      # 0: strips
    rows_per_strip: int = 8
    planar: bool = False
    compression: int = 1               # 1 none, 8 Deflate, 32773 PackBits
    predictor: int = 1
    no_data: str | None = None
    metadata: str | None = None
    transform: tuple = (500.0, 2.0, 0.0, 800.0, 0.0, -2.0)  # GDAL order
    matrix: bool = False               # ModelTransformation rather than scale + tiepoint
    geographic: bool = False
    pixel_is_point: bool = False
    extra: dict = field(default_factory=dict)


def _format(dtype: np.dtype) -> int:
    return 3 if dtype.kind == "f" else (2 if dtype.kind == "i" else 1)


def packbits(data: bytes) -> bytes:
    """PackBits with literal runs only, and a repeat run wherever three bytes repeat."""
    out = bytearray()
    i = 0
    while i < len(data):
        run = 1
        while i + run < len(data) and run < 128 and data[i + run] == data[i]:
            run += 1
        if run >= 3:
            out += struct.pack("b", 1 - run) + data[i:i + 1]
            i += run
            continue
        j = i
        while j < len(data) and j - i < 128 and not (
                j + 2 < len(data) and data[j] == data[j + 1] == data[j + 2]):
            j += 1
        out += struct.pack("b", j - i - 1) + data[i:j]
        i = j
    return bytes(out)


def _encode_chunk(block: np.ndarray, spec: Spec, order: str) -> bytes:
    """One chunk (rows, cols, samples) of the image, predicted and compressed."""
    rows, cols, samples = block.shape
    a = block.astype(block.dtype.newbyteorder(order), copy=True)
    if spec.predictor == 2:
        flat = a.reshape(rows, cols * samples).astype(block.dtype)
        diff = flat.copy()
        with np.errstate(over="ignore"):
            diff[:, samples:] = flat[:, samples:] - flat[:, :-samples]
        raw = diff.astype(block.dtype.newbyteorder(order)).tobytes()
    elif spec.predictor == 3:
        size = block.dtype.itemsize
        big = block.astype(block.dtype.newbyteorder(">")).reshape(rows, cols * samples)
        out = bytearray()
        for r in range(rows):
            b = np.frombuffer(big[r].tobytes(), np.uint8).reshape(cols * samples, size)
            planes = b.T.reshape(-1).copy()  # most significant byte's plane first
            with np.errstate(over="ignore"):
                planes[samples:] = planes[samples:] - planes[:-samples]
            out += planes.tobytes()
        raw = bytes(out)
    else:
        raw = a.tobytes()
    if spec.compression == 8:
        return zlib.compress(raw, 6)
    if spec.compression == 32773:
        return packbits(raw)
    return raw


def write(path, spec: Spec) -> None:
    order = ">" if spec.big_endian else "<"
    bands, height, width = spec.values.shape
    dtype = spec.values.dtype
    chunks = []
    planes = range(bands) if spec.planar else [None]
    cw = spec.tile if spec.tile else width
    ch = spec.tile if spec.tile else spec.rows_per_strip
    for band in planes:
        img = spec.values[band:band + 1] if band is not None else spec.values
        img = np.moveaxis(img, 0, -1)  # (height, width, samples)
        for y0 in range(0, height, ch):
            for x0 in range(0, width, cw):
                if spec.tile:
                    block = np.zeros((ch, cw, img.shape[2]), dtype)
                    part = img[y0:y0 + ch, x0:x0 + cw]
                    block[:part.shape[0], :part.shape[1]] = part
                else:
                    block = img[y0:y0 + ch]
                chunks.append(_encode_chunk(block, spec, order))
    word = 8 if spec.big_tiff else 4
    header = (b"MM" if spec.big_endian else b"II") + struct.pack(
        order + "H", 43 if spec.big_tiff else 42)
    header += struct.pack(order + "HHQ", 8, 0, 0) if spec.big_tiff else struct.pack(order + "I", 0)
    data = bytearray(header)
    offsets, counts = [], []
    for c in chunks:
        if len(data) % 2:
            data += b"\0"
        offsets.append(len(data))
        counts.append(len(c))
        data += c
    t = spec.transform
    tags = {256: (LONG, [width]), 257: (LONG, [height]),
            258: (SHORT, [dtype.itemsize * 8] * bands), 259: (SHORT, [spec.compression]),
            262: (SHORT, [1]), 277: (SHORT, [bands]), 284: (SHORT, [2 if spec.planar else 1]),
            339: (SHORT, [_format(dtype)] * bands)}
    if spec.predictor != 1:
        tags[317] = (SHORT, [spec.predictor])
    off_type = LONG8 if spec.big_tiff else LONG
    if spec.tile:
        tags[322], tags[323] = (LONG, [spec.tile]), (LONG, [spec.tile])
        tags[324], tags[325] = (off_type, offsets), (off_type, counts)
    else:
        tags[278] = (LONG, [spec.rows_per_strip])
        tags[273], tags[279] = (off_type, offsets), (off_type, counts)
    if spec.matrix:
        tags[34264] = (DOUBLE, [t[1], t[2], 0.0, t[0], t[4], t[5], 0.0, t[3], 0.0, 0.0, 0.0, 0.0,
                                0.0, 0.0, 0.0, 1.0])
    else:
        tags[33550] = (DOUBLE, [t[1], -t[5], 0.0])
        tags[33922] = (DOUBLE, [0.0, 0.0, 0.0, t[0], t[3], 0.0])
    keys = [1, 1, 0, 2, 1024, 0, 1, 2 if spec.geographic else 1, 1025, 0, 1,
            2 if spec.pixel_is_point else 1]
    tags[34735] = (SHORT, keys)
    if spec.no_data is not None:
        tags[42113] = (ASCII, spec.no_data)
    if spec.metadata is not None:
        tags[42112] = (ASCII, spec.metadata)
    tags.update(spec.extra)
    if len(data) % 2:
        data += b"\0"
    ifd_at = len(data)
    entry = 4 + 2 * word
    table = (8 if spec.big_tiff else 2) + len(tags) * entry + word
    values = bytearray()
    ifd = bytearray(struct.pack(order + ("Q" if spec.big_tiff else "H"), len(tags)))
    for tag in sorted(tags):
        typ, val = tags[tag]
        if typ == ASCII:
            payload, count = val.encode() + b"\0", len(val) + 1
        else:
            fmt = {SHORT: "H", LONG: "I", DOUBLE: "d", LONG8: "Q"}[typ]
            payload, count = struct.pack(order + fmt * len(val), *val), len(val)
        ifd += struct.pack(order + "HH", tag, typ)
        ifd += struct.pack(order + ("Q" if spec.big_tiff else "I"), count)
        if len(payload) <= word:
            ifd += payload + b"\0" * (word - len(payload))
        else:
            if len(values) % 2:
                values += b"\0"
            ifd += struct.pack(order + ("Q" if spec.big_tiff else "I"),
                               ifd_at + table + len(values))
            values += payload
    ifd += b"\0" * word
    assert len(ifd) == table
    data += ifd + values
    first = struct.pack(order + ("Q" if spec.big_tiff else "I"), ifd_at)
    data[8 if spec.big_tiff else 4:(16 if spec.big_tiff else 8)] = first
    with open(path, "wb") as f:
        f.write(bytes(data))


def cases() -> dict:
    """The reader's test files: name -> Spec."""
    rng = np.random.default_rng(11)
    ramp = (np.arange(37 * 53).reshape(37, 53) % 251).astype(np.float64)
    terrain = rng.normal(0.0, 300.0, (37, 53)).cumsum(axis=1)
    out = {
        "int16_be_strips_deflate_pred2": Spec(
            terrain.astype(np.int16)[None], big_endian=True, rows_per_strip=7, compression=8,
            predictor=2, no_data="-32768",
            metadata='<GDALMetadata><Item name="SCALE" sample="0" role="scale">0.5</Item>'
                     '<Item name="OFFSET" sample="0" role="offset">100</Item></GDALMetadata>'),
        "uint8_tiles_packbits": Spec(ramp.astype(np.uint8)[None], tile=16, compression=32773),
        "float64_planar_bigtiff_point": Spec(
            np.stack([terrain, -terrain * 0.25]), big_tiff=True, planar=True, matrix=True,
            pixel_is_point=True, transform=(-120.5, 0.25, 0.0, 45.0, 0.0, -0.125)),
        "uint16_be_chunky_tiles_pred2": Spec(
            np.stack([ramp * 200, ramp * 3, 65535 - ramp * 100]).astype(np.uint16),
            big_endian=True, tile=32, compression=8, predictor=2),
        "float32_be_pred3_deflate": Spec(
            terrain.astype(np.float32)[None], big_endian=True, tile=16, compression=8,
            predictor=3, no_data="nan"),
        "int32_geographic_wrap": Spec(
            (terrain[:, :48] * 10).astype(np.int32)[None], geographic=True, rows_per_strip=5,
            transform=(-180.0, 7.5, 0.0, 80.0, 0.0, -4.0), no_data="-999999"),
    }
    out["int16_be_strips_deflate_pred2"].values[0, 3:6, 10:14] = -32768
    out["int32_geographic_wrap"].values[0, 20, 0:3] = -999999
    return out
