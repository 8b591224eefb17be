"""Synthetic velocity models and station sets for the sphfd workloads.

All models are written on the reference's spherical grid: node (i, j, k) sits
at longitude x0 + i*df, geocentric colatitude y0' + j*dq and radius
z0r - k*h (src/sphfd.c:151-156, 437-455).  Par values are chosen so that the
grid is centred on a given geographic point with dq = df = 0 in the par file,
i.e. the reference derives dq = h/6371 and df = h/(6371 sin y0') itself.
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np

PI = 3.141592654          # the reference's constants (src/sphfd.c:144-150)
HPI = 1.570796327
REARTH = 6371.0
DEGRAD = PI / 180.0


def glath(hlat: float, depth_km: float) -> tuple[float, float]:
    """Geodetic latitude (rad) and ellipsoidal depth (km) -> geocentric
    latitude (rad) and radius (km), WGS84 (src/sphfd.c:5836-5893)."""
    ap = 6378137.0
    esq = 0.00669437978616733379
    hm = -depth_km * 1000.0
    s = math.sin(hlat)
    anu = ap / math.sqrt(1.0 - esq * s * s)
    xp = (anu + hm) * math.cos(hlat)
    zp = ((1.0 - esq) * anu + hm) * s
    return math.atan2(zp, xp), math.sqrt(xp * xp + zp * zp) / 1000.0


@dataclasses.dataclass(frozen=True)
class GridSpec:
    nx: int
    ny: int
    nz: int
    h: float
    x0: float      # geographic longitude of the west edge (deg)
    y0: float      # geodetic latitude of the north edge (deg)
    z0: float = 0.0

    @property
    def n(self) -> int:
        return self.nx * self.ny * self.nz

    def dq_deg(self) -> float:
        return self.h / REARTH / DEGRAD

    def df_deg(self) -> float:
        colat = HPI - glath(self.y0 * DEGRAD, self.z0)[0]
        return abs(self.h / (REARTH * math.sin(colat))) / DEGRAD

    def node_lonlat(self, i: float, j: float) -> tuple[float, float]:
        """Approximate geographic lon/lat (deg) of a (possibly fractional) node."""
        return self.x0 + i * self.df_deg(), self.y0 - j * self.dq_deg()


def centred_grid(nx: int, ny: int, nz: int, h: float, lat: float = 23.5,
                 lon: float = 121.0, z0: float = 0.0) -> GridSpec:
    """A grid whose centre node is near (lat, lon); corner values rounded to
    1e-6 deg so that the par file states them exactly."""
    dq = h / REARTH / DEGRAD
    y0 = round(lat + 0.5 * (ny - 1) * dq, 6)
    colat = HPI - glath(y0 * DEGRAD, z0)[0]
    df = abs(h / (REARTH * math.sin(colat))) / DEGRAD
    x0 = round(lon - 0.5 * (nx - 1) * df, 6)
    return GridSpec(nx, ny, nz, float(h), x0, y0, float(z0))


def gradient_velocity(g: GridSpec, v_top: float = 5.0, v_bot: float = 8.0,
                      zspan: float = 64.0) -> np.ndarray:
    """1D linear gradient v(z) = v_top + (v_bot - v_top) * depth / zspan,
    continued linearly below zspan.  Depth is k*h below the grid top."""
    depth = np.arange(g.nz, dtype=np.float64) * g.h
    v = v_top + (v_bot - v_top) * depth / zspan
    return np.broadcast_to(v[:, None, None], (g.nz, g.ny, g.nx)).astype(np.float32)


def checkerboard(g: GridSpec, cell_km: float = 20.0, amp: float = 0.05) -> np.ndarray:
    """Multiplicative +-amp checkerboard of cubic cells of cell_km."""
    k, j, i = np.meshgrid(np.arange(g.nz), np.arange(g.ny), np.arange(g.nx), indexing="ij")
    c = (np.floor(i * g.h / cell_km) + np.floor(j * g.h / cell_km) + np.floor(k * g.h / cell_km)) % 2
    return (1.0 + amp * np.where(c == 0, 1.0, -1.0)).astype(np.float32)


def gaussian_anomalies(g: GridSpec, rng: np.random.Generator, count: int = 3,
                       amp: float = 0.08, sigma=(10.0, 30.0)) -> np.ndarray:
    """Multiplicative field 1 + sum of +-amp Gaussian blobs with sigma in km."""
    k, j, i = np.meshgrid(np.arange(g.nz) * g.h, np.arange(g.ny) * g.h,
                          np.arange(g.nx) * g.h, indexing="ij")
    f = np.ones((g.nz, g.ny, g.nx), dtype=np.float64)
    for _ in range(count):
        cx, cy, cz = rng.uniform(0, g.nx * g.h), rng.uniform(0, g.ny * g.h), rng.uniform(0, g.nz * g.h)
        s = rng.uniform(*sigma)
        sign = 1.0 if rng.random() < 0.5 else -1.0
        f += sign * amp * np.exp(-((i - cx) ** 2 + (j - cy) ** 2 + (k - cz) ** 2) / (2 * s * s))
    return f.astype(np.float32)


def layered_perturbation(g: GridSpec, rng: np.random.Generator, amp: float = 0.04,
                         min_thick: float = 2.0, max_thick: float = 12.0) -> np.ndarray:
    """Multiplicative 1D field of random layers, each scaled by U(-amp, amp)."""
    f = np.ones(g.nz, dtype=np.float64)
    z = 0.0
    depth = np.arange(g.nz) * g.h
    while z < g.nz * g.h:
        t = rng.uniform(min_thick, max_thick)
        f[(depth >= z) & (depth < z + t)] = 1.0 + rng.uniform(-amp, amp)
        z += t
    return np.broadcast_to(f[:, None, None], (g.nz, g.ny, g.nx)).astype(np.float32)


@dataclasses.dataclass(frozen=True)
class Station:
    name: str
    lon: float      # fxs (deg)
    lat: float      # fys, geodetic (deg)
    depth: float    # fzs, km, positive down (elevation negative)


def surface_node_station(g: GridSpec, rng: np.random.Generator, name: str = "S000",
                         margin: int = 0) -> Station:
    """A station at the geographic position of a random top-surface node."""
    i = int(rng.integers(margin, g.nx - margin))
    j = int(rng.integers(margin, g.ny - margin))
    lon, lat = g.node_lonlat(i, j)
    return Station(name, round(lon, 5), round(lat, 5), round(g.z0, 3))


def random_surface_stations(g: GridSpec, rng: np.random.Generator, count: int,
                            prefix: str = "S") -> list[Station]:
    """Stations uniform-random over the top surface (continuous positions)."""
    out = []
    for s in range(count):
        fi = rng.uniform(0, g.nx - 1)
        fj = rng.uniform(0, g.ny - 1)
        lon, lat = g.node_lonlat(fi, fj)
        out.append(Station(f"{prefix}{s:03d}", round(lon, 5), round(lat, 5), round(g.z0, 3)))
    return out
