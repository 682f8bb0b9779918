# SPDX-License-Identifier: BSD-3-Clause
"""A served stream of tile requests from a moving camera, driven through the tile server.

Requests arrive as a Poisson process. At each arrival, the camera has flown along a straight
ground track (on a sphere, a great circle) at constant speed and altitude. The requested tile is the quadtree tile, at the level
of detail the camera's distance calls for, that holds a ground point drawn around the camera.

The level-of-detail rule (a screen-space-error rule of this module):
* level 10 (the finest) covers tiles of ``finest_side`` metres; each coarser level doubles the
  side;
* the level for a point at slant range rho from the camera is ``10 - floor(log2(rho / altitude))``,
  clamped to 0..10, so the tile's ground size grows in proportion to its distance;
* the tile's edge is drawn from ``edges`` (cells, 2^k + 1), so its pixel size is
  side / (edge - 1).
"""
from __future__ import annotations

import threading
import time
from dataclasses import dataclass

import numpy as np


@dataclass
class CameraStream:
    rate: float = 2000.0            # requests per second
    duration: float = 10.0          # seconds
    path_length: float = 50_000.0   # metres of ground track, flown along +x from x = 0, y = 0
    altitude: float = 200.0         # metres
    finest_side: float = 128.0      # ground side of a level-10 tile, metres
    levels: int = 10
    edges: tuple = (129, 257, 513, 1025)
    min_ground: float = 20.0        # ground distances from the camera, log-uniform, metres
    max_ground: float = 20_000.0
    seed: int = 5

    def requests(self) -> np.ndarray:
        """Arrival time (s from the start), level and tile of every request, in arrival order."""
        rng = np.random.default_rng(self.seed)
        n = rng.poisson(self.rate * self.duration)
        t = np.sort(rng.uniform(0.0, self.duration, n))
        cam_x = self.path_length * t / self.duration
        ground = np.exp(rng.uniform(np.log(self.min_ground), np.log(self.max_ground), n))
        bearing = rng.uniform(0.0, 2.0 * np.pi, n)
        px = cam_x + ground * np.cos(bearing)
        py = ground * np.sin(bearing)
        rho = np.hypot(ground, self.altitude)
        level = np.clip(self.levels - np.floor(np.log2(rho / self.altitude)), 0,
                        self.levels).astype(np.int64)
        side = self.finest_side * np.exp2(self.levels - level)
        edge = np.asarray(self.edges)[rng.integers(0, len(self.edges), n)]
        out = np.zeros(n, dtype=[("t", np.float64), ("level", np.int32), ("origin_x", np.float64),
                                 ("origin_y", np.float64), ("pixel", np.float64),
                                 ("width", np.uint32), ("height", np.uint32)])
        out["t"] = t
        out["level"] = level
        out["origin_x"] = np.floor(px / side) * side
        out["origin_y"] = (np.floor(py / side) + 1.0) * side
        out["pixel"] = side / (edge - 1)
        out["width"] = edge
        out["height"] = edge
        return out


def destination(lon, lat, bearing, arc):
    """The point `arc` radians along the great circle leaving (lon, lat) (degrees) at `bearing`
    (radians clockwise from north), in degrees."""
    phi, lam = np.radians(lat), np.radians(lon)
    sin_phi2 = np.sin(phi) * np.cos(arc) + np.cos(phi) * np.sin(arc) * np.cos(bearing)
    phi2 = np.arcsin(np.clip(sin_phi2, -1.0, 1.0))
    lam2 = lam + np.arctan2(np.sin(bearing) * np.sin(arc) * np.cos(phi),
                            np.cos(arc) - np.sin(phi) * sin_phi2)
    lon2 = (np.degrees(lam2) + 180.0) % 360.0 - 180.0
    return lon2, np.degrees(phi2)


@dataclass
class GreatCircleStream(CameraStream):
    """The stream on a sphere: the camera flies `path_length` meters along the great circle that
    leaves (`start_lon`, `start_lat`) at `heading` (degrees clockwise from north), and tiles are
    longitude-latitude quadtree tiles whose side in degrees is the level's ground side over the
    sphere's meters per degree."""
    radius: float = 1737400.0
    start_lon: float = 20.0
    start_lat: float = 5.0
    heading: float = 60.0

    def unit_length(self) -> float:
        return self.radius * np.pi / 180.0

    def track(self, along: np.ndarray, across: np.ndarray | float = 0.0):
        """Longitude and latitude `along` meters down the track and `across` meters to its right."""
        lon, lat = destination(self.start_lon, self.start_lat, np.radians(self.heading),
                               np.asarray(along) / self.radius)
        if np.all(np.asarray(across) == 0.0):
            return lon, lat
        # The track's bearing there: back to the start, turned half a circle.
        back = np.arctan2(
            np.sin(np.radians(self.start_lon - lon)) * np.cos(np.radians(self.start_lat)),
            np.cos(np.radians(lat)) * np.sin(np.radians(self.start_lat)) -
            np.sin(np.radians(lat)) * np.cos(np.radians(self.start_lat)) *
            np.cos(np.radians(self.start_lon - lon)))
        forward = np.where(np.asarray(along) > 0.0, back + np.pi, np.radians(self.heading))
        return destination(lon, lat, forward + 0.5 * np.pi, np.asarray(across) / self.radius)

    def requests(self) -> np.ndarray:
        rng = np.random.default_rng(self.seed)
        n = rng.poisson(self.rate * self.duration)
        t = np.sort(rng.uniform(0.0, self.duration, n))
        cam_lon, cam_lat = self.track(self.path_length * t / self.duration)
        ground = np.exp(rng.uniform(np.log(self.min_ground), np.log(self.max_ground), n))
        bearing = rng.uniform(0.0, 2.0 * np.pi, n)
        plon, plat = destination(cam_lon, cam_lat, bearing, ground / self.radius)
        rho = np.hypot(ground, self.altitude)
        level = np.clip(self.levels - np.floor(np.log2(rho / self.altitude)), 0,
                        self.levels).astype(np.int64)
        side = self.finest_side * np.exp2(self.levels - level) / self.unit_length()
        edge = np.asarray(self.edges)[rng.integers(0, len(self.edges), n)]
        out = np.zeros(n, dtype=[("t", np.float64), ("level", np.int32), ("origin_x", np.float64),
                                 ("origin_y", np.float64), ("pixel", np.float64),
                                 ("width", np.uint32), ("height", np.uint32)])
        out["t"] = t
        out["level"] = level
        out["origin_x"] = np.floor(plon / side) * side
        out["origin_y"] = (np.floor(plat / side) + 1.0) * side
        out["pixel"] = side / (edge - 1)
        out["width"] = edge
        out["height"] = edge
        return out

    def swath_craters(self, half_width: float, coefficient: float = 0.05,
                      exponent: float = 2.0, min_diameter: float = 2.0,
                      max_diameter: float = 200.0, margin: float = 2000.0,
                      seed: int = 5) -> np.ndarray:
        """A synthetic catalogue along the track (CRATER_DTYPE, longitude and latitude in x and
        y): the power law N(>=D) = coefficient D^-exponent per m^2 over the strip `half_width`
        meters each side of the track and `margin` meters beyond its ends, centres uniform in
        (along, across), diameters by inverse transform, ages uniform over 1 Gyr."""
        from . import native as N

        rng = np.random.default_rng(seed)
        length = self.path_length + 2.0 * margin
        lo, hi = min_diameter ** -exponent, max_diameter ** -exponent
        count = rng.poisson(coefficient * length * 2.0 * half_width * (lo - hi))
        along = rng.uniform(-margin, self.path_length + margin, count)
        across = rng.uniform(-half_width, half_width, count)
        out = np.zeros(count, N.CRATER_DTYPE)
        out["x"], out["y"] = self.track(along, across)
        out["diameter"] = (lo - rng.uniform(0.0, 1.0, count) * (lo - hi)) ** (-1.0 / exponent)
        out["age_gyr"] = rng.uniform(0.0, 1.0, count)
        out["id"] = np.arange(1, count + 1, dtype=np.uint64)
        return out


    def placed_swath(self, half_width: float, coefficient: float, exponent: float,
                     min_diameter: float, max_diameter: float, margin_radii: float = 0.0,
                     min_resolved_diameter: float = 0.0, lifetime_gyr: float = 1.0,
                     margin: float = 2000.0, seed: int = 5, segment: float = 250.0,
                     threads: int = 0) -> np.ndarray:
        """The catalogue along the track as the reference places it (CRATER_DTYPE, longitude and
        latitude in x and y): GenerateCraters on the sphere's cube-sphere cells over the
        longitude-latitude box of each `segment` meters of the strip `half_width` meters each
        side of the track (and `margin` beyond its ends), in one generation over the union of the
        boxes (each cell drawn once, a crater kept when any box keeps it), keeping the craters
        whose centres lie in the strip. The density is N(>=D) = coefficient D^-exponent per m^2
        with ages uniform over `lifetime_gyr`."""
        from . import native as N

        boxes = []
        for a0 in np.arange(-margin, self.path_length + margin, segment):
            along = np.array([a0, a0, a0 + segment, a0 + segment])
            lon, lat = self.track(along, np.array([-half_width, half_width] * 2))
            boxes.append((lon.min(), lat.min(), lon.max(), lat.max()))
        craters = N.generate_sphere_craters(
            self.radius, np.array(boxes), coefficient, exponent, min_diameter, max_diameter,
            margin_radii=margin_radii, min_resolved_diameter=min_resolved_diameter,
            lifetime_gyr=lifetime_gyr, seed=seed, threads=threads)
        # Along-track and cross-track position against the track's great circle.
        start = np.array(unit_vector(self.start_lon, self.start_lat))
        ahead = np.array(unit_vector(*self.track(np.array([1000.0]))))[:, 0]
        normal = np.cross(start, ahead)
        normal /= np.linalg.norm(normal)
        p = np.stack(unit_vector(craters["x"], craters["y"]), axis=1)
        across = np.arcsin(np.clip(p @ normal, -1.0, 1.0)) * self.radius
        forward = np.cross(normal, start)
        along = np.arctan2(p @ forward, p @ start) * self.radius
        keep = (np.abs(across) <= half_width) & (along >= -margin) & (
            along <= self.path_length + margin)
        return craters[keep]


def served_boulders(stream, requests: np.ndarray, coefficient: float, exponent: float,
                    min_diameter: float, max_diameter: float, seed: int = 5,
                    threads: int = 0) -> tuple[np.ndarray, dict]:
    """The boulders the stream's tiles draw, as the reference's BoulderLayer lists them for each
    tile (ReliefBoulders over the tile's box grown by half a sample, at its spacing,
    boulder_relief.cpp:287-291): for each spacing that resolves a boulder (below half the largest
    diameter), one placement over the union of that spacing's tile boxes (tgx_place_boulders_boxes,
    each box grown by the spacing and 3 m, which only adds boulders that reach no tile), and the
    union over spacings, each boulder once. A boulder resolved at a spacing is listed with the same
    draws at every finer one, so the union holds every tile's own list. Returns the relief list
    (BOULDER_DTYPE; longitude and latitude on a sphere) and the placement's figures."""
    from . import native as N

    sphere = isinstance(stream, GreatCircleStream)
    unit = stream.unit_length() if sphere else 1.0
    spacing = requests["pixel"] * unit
    parts, stats = [], {}
    for sp in np.unique(spacing):
        if not sp < 0.5 * max_diameter:
            continue
        sel = requests[spacing == sp]
        w, h = sel["width"] * sel["pixel"], sel["height"] * sel["pixel"]
        boxes = np.unique(np.stack([sel["origin_x"], sel["origin_y"] - h, sel["origin_x"] + w,
                                    sel["origin_y"]], axis=1), axis=0)
        pad = np.full(len(boxes), (sp + 3.0) / unit)
        if sphere:
            lat = np.maximum(np.abs(boxes[:, 1]), np.abs(boxes[:, 3]))
            pad_x = pad / np.cos(np.radians(np.minimum(lat + pad, 89.0)))
        else:
            pad_x = pad
        boxes = boxes + np.stack([-pad_x, -pad, pad_x, pad], axis=1)
        placed = N.place_boulders_boxes(boxes, float(sp), coefficient, exponent, min_diameter,
                                        max_diameter, seed=seed,
                                        sphere_radius=stream.radius if sphere else 0.0,
                                        threads=threads)
        stats[f"{sp:g} m"] = {"tiles": int(len(boxes)), "boulders": int(placed.size)}
        parts.append(placed)
    placed = np.concatenate(parts) if parts else np.zeros(0, N.PLACED_BOULDER_DTYPE)
    _, first = np.unique(placed["id"], return_index=True)
    placed = placed[np.sort(first)]
    return N.relief_boulders(placed), stats


def unit_vector(lon, lat):
    """The unit vector(s) at longitude and latitude in degrees, as (x, y, z)."""
    lam, phi = np.radians(lon), np.radians(lat)
    return np.cos(phi) * np.cos(lam), np.cos(phi) * np.sin(lam), np.sin(phi)


def serve(server, requests: np.ndarray, keep=None) -> dict:
    """Submits `requests` at their arrival times (a submitting thread) while another thread waits
    for them in order. Returns per-request (latency, batch, batch size) arrays and the wall time.
    `keep(i, heights)` may look at each result; by default results are dropped once done."""
    n = requests.size
    tickets = np.zeros(n, np.uint64)
    submitted = threading.Semaphore(0)
    lat = np.zeros(n)
    queue_wait = np.zeros(n)
    batch = np.zeros(n, np.int64)
    batch_tiles = np.zeros(n, np.int64)
    errors = []
    start = time.perf_counter() + 0.05

    def submitter():
        try:
            for i in range(n):
                due = start + float(requests["t"][i])
                while True:
                    left = due - time.perf_counter()
                    if left <= 0.0:
                        break
                    if left > 0.0015:
                        time.sleep(left - 0.001)
                r = requests[i]
                tickets[i] = server.submit(float(r["origin_x"]), float(r["origin_y"]),
                                           float(r["pixel"]), int(r["width"]), int(r["height"]))
                submitted.release()
        except Exception as e:  # noqa: BLE001 - reported to the caller
            errors.append(e)
            for _ in range(n):
                submitted.release()

    th = threading.Thread(target=submitter)
    th.start()
    for i in range(n):
        submitted.acquire()
        if errors:
            break
        heights, info = server.wait(int(tickets[i]))
        lat[i] = info["done"] - info["submitted"]
        queue_wait[i] = info["admitted"] - info["submitted"]
        batch[i] = info["batch"]
        batch_tiles[i] = info["batch_tiles"]
        if keep is not None:
            keep(i, heights)
    th.join()
    if errors:
        raise errors[0]
    wall = time.perf_counter() - start
    return {"latency_s": lat, "queue_s": queue_wait, "batch": batch, "batch_tiles": batch_tiles,
            "wall_s": wall}
