"""Host side of sphrayderv's ray tracing over resident tables.

sphrayderv (src/sphrayderv.c) reads a spec file, a station list and an event
file, traces every pick's ray from the event back to the station through the
station's travel-time table, and writes the inversion's derivative files.
This module keeps the reference's set-up arithmetic -- the grid in radians
and geocentric colatitude, the coarse node coordinates, station and event
positions, the arrival times relative to the origin time -- so that the rays
handed to the device (`fdgpu_rays_trace`) start from the same doubles the
reference traces from, and it parses the reference's input files.

Units after set-up (sphrayderv.c:527-545): longitudes in radians, latitudes
as geocentric colatitudes in radians, depths in km positive down.
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np

# sphrayderv.c:95-97: float literals held in doubles (rearth stays a float)
HPI = float(np.float32(1.570796))
DEGRAD = float(np.float32(0.017453292))
REARTH = 6371.0


def glat(hlat: float) -> float:
    """Geographic -> geocentric latitude (lib/common/geographic_method.c:7-16),
    with the function's float constants."""
    halfpi, polfac, elfac = HPI, float(np.float32(0.010632)), float(np.float32(0.993277))
    if halfpi - abs(hlat) >= 0.05:
        return math.atan(elfac * math.sin(hlat) / math.cos(hlat))
    return hlat / elfac - abs(polfac) * (1 if hlat >= 0 else -1)


def glatinv(hlat: float) -> float:
    """Geocentric -> geographic latitude (lib/common/geographic_method.c:22-35)."""
    if 1.570796 - abs(hlat) >= 0.05:
        return math.atan(math.sin(hlat) / math.cos(hlat) / 0.993277)
    return (hlat + 0.010632) * 0.993277 if hlat > 0 else (hlat - 0.010632) * 0.993277


@dataclasses.dataclass
class RayGrid:
    """The fine grid of the tables and the coarse inversion grid, in
    sphrayderv's units (the fields of fdgpu_ray_grid)."""
    nx: int
    ny: int
    nz: int
    x0: float
    y0: float
    z0: float
    df: float
    dq: float
    h: float
    gx: np.ndarray      # float32 coarse node coordinates
    gy: np.ndarray
    gz: np.ndarray
    ido1d: int = 0
    max_segments: int = 0

    @property
    def nxc(self) -> int:
        return len(self.gx)

    @property
    def nyc(self) -> int:
        return len(self.gy)

    @property
    def nzc(self) -> int:
        return len(self.gz)

    def segment_limit(self) -> int:
        return self.max_segments or 8 * (self.nx + self.ny + self.nz)


def grid_from_spec(nxc: int, nyc: int, nzc: int, igridx, igridy, igridz, x0_deg: float, y0_deg: float,
                   z0: float, h: float, dq_deg: float = 0.0, df_deg: float = 0.0, ido1d: int = 0) -> RayGrid:
    """sphrayderv's grid set-up from its spec values (sphrayderv.c:483-545,
    814-826): fine dimensions 1 + the sum of the igrid entries, origin to
    radians and geocentric colatitude, dq/df from h when below 1e-4, coarse
    node coordinates accumulated in float."""
    igridx, igridy, igridz = (np.asarray(a, dtype=np.int64) for a in (igridx, igridy, igridz))
    nx = 1 + int(igridx[:nxc - 1].sum())
    ny = 1 + int(igridy[:nyc - 1].sum())
    nz = 1 + int(igridz[:nzc - 1].sum())
    y0 = HPI - glat(y0_deg * DEGRAD)
    x0 = x0_deg * DEGRAD
    dq = dq_deg * DEGRAD
    df = df_deg * DEGRAD
    if abs(dq) < 0.0001:
        dq = h / REARTH
    if abs(df) < 0.0001:
        df = abs(h / (REARTH * math.sin(y0)))

    def coarse(origin, step, ig, n):
        g = np.zeros(n, dtype=np.float32)
        g[0] = np.float32(origin)
        for i in range(1, n):
            g[i] = np.float32(float(g[i - 1]) + step * int(ig[i - 1]))
        return g
    return RayGrid(nx, ny, nz, x0, y0, float(z0), df, dq, float(h), coarse(x0, df, igridx, nxc),
                   coarse(y0, dq, igridy, nyc), coarse(z0, h, igridz, nzc), ido1d)


@dataclasses.dataclass
class Station:
    name: str
    lon: float      # stlon: radians, held as float (sphrayderv.c:874)
    colat: float    # stlat: geocentric colatitude, float
    z: float        # stz: km (-elevation)
    tcor: tuple = (0.0, 0.0)


def station(name: str, lat_deg: float, lon_deg: float, elev_m: int, tcor=(0.0, 0.0)) -> Station:
    """A station-list line's position as sphrayderv stores it
    (sphrayderv.c:867-876): float longitude and colatitude in radians and
    the depth -(elevation/1000) in float."""
    slat, slon = float(np.float32(lat_deg)), float(np.float32(lon_deg))
    return Station(name, float(np.float32(slon * DEGRAD)), float(np.float32(HPI - glat(slat * DEGRAD))),
                   float(-np.float32(np.float32(int(elev_m)) / np.float32(1000.0))),
                   (float(np.float32(tcor[0])), float(np.float32(tcor[1]))))


def event_position(lat_deg: float, lon_deg: float, depth: float) -> tuple[float, float, float]:
    """An event header's position (sphrayderv.c:949-951): the depth is read
    as a float, latitude and longitude as doubles."""
    return lon_deg * DEGRAD, HPI - glat(lat_deg * DEGRAD), float(np.float32(depth))


def epoch(year: int, day: int, hour: int, minute: int, sec: float) -> float:
    """htoe (lib/common/time_process.c:85-109): seconds since 1970."""
    def leap(y):
        return (y % 4 == 0 and y % 100 != 0) or y % 400 == 0
    days = 0
    if year > 1970:
        for y in range(1970, year):
            days += 365 + leap(y)
    elif year < 1970:
        for y in range(year, 1970):
            days -= 365 + leap(y)
    days += day - 1
    return float(days * 86400) + 3600.0 * hour + 60.0 * minute + sec
