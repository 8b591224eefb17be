"""Synthetic workloads of the sphfd path (seeded; written in the reference's
layout when files are needed).

  C1  one station, 65 x 65 x 33 at 1 km, linear gradient Vp 5.0 -> 8.0 km/s
      over 0-64 km, source at a seeded surface node
  C2  48 stations x (P, S) = 96 tables of 161 x 161 x 81 at 1 km, gradient x
      +-5 % checkerboard of 20 km cells, Vs = Vp / 1.73
  C3  256 station solves whose boxes are clipped to per-station radii
      r ~ U(40, 200) km: windows (r+1) x (r+1) x (min(r,100)+1) of one
      401 x 401 x 101 model (gradient x three +-8 % Gaussian anomalies of
      sigma 10-30 km), each window its own table
  C4  a Poisson stream of requests, sizes 81^3 to 257 x 257 x 129, each with
      its own velocity model (the gradient times +-3 % Gaussian blobs, as an
      inversion update would make)
  C5  1000 stations x (P, S) on 321 x 321 x 161, gradient x random layers
      (+-4 %), tables resident across inversion iterations

Grids are centred at 23.5N 121E.  Station par values follow make_par_file's
format (fxs/fys at 1e-5 deg, fzs at 1e-3 km; reverse=50).
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np

from . import synth


def table_par(g: synth.GridSpec, st: synth.Station, reverse: int = 50) -> dict:
    return dict(nx=g.nx, ny=g.ny, nz=g.nz, h=g.h, x0=g.x0, y0=g.y0, z0=g.z0,
                fxs=st.lon, fys=st.lat, fzs=st.depth, reverse=reverse)


# ------------------------------------------------------------------ C1

def c1(seed: int = 0):
    """(grid, velocity, station) of C1: Vp(z) = 5 + 3 z / 64 km."""
    g = synth.centred_grid(65, 65, 33, 1.0)
    v = synth.gradient_velocity(g, 5.0, 8.0, 64.0)
    st = synth.surface_node_station(g, np.random.default_rng(seed), "C1", margin=4)
    return g, np.ascontiguousarray(v, dtype=np.float32), st


def gradient_time_flat(g: synth.GridSpec, geom: dict, v0: float, grad: float) -> np.ndarray:
    """FP64 travel times from the source in a medium whose velocity grows
    linearly with depth below the grid top, v = v0 + grad * d, taking the
    straight chord between source and node for the distance and depths
    measured along the radius (flat-layer approximation of the spherical
    grid): t = arccosh(1 + grad^2 L^2 / (2 v_s v_n)) / grad."""
    k, j, i = np.meshgrid(np.arange(g.nz), np.arange(g.ny), np.arange(g.nx), indexing="ij")
    r = geom["z0r"] - geom["h"] * k
    q = geom["y0"] + geom["dq"] * j
    f = geom["x0"] + geom["df"] * i
    P = np.stack([r * np.cos(q), r * np.sin(q) * np.cos(f), r * np.sin(q) * np.sin(f)])
    rs = geom["fzss"]
    S = np.array([rs * math.cos(geom["fyss"]), rs * math.sin(geom["fyss"]) * math.cos(geom["fxss"]),
                  rs * math.sin(geom["fyss"]) * math.sin(geom["fxss"])])
    L = np.linalg.norm(P - S[:, None, None, None], axis=0)
    vs = v0 + grad * (geom["z0r"] - rs)
    vn = v0 + grad * (geom["z0r"] - r)
    return np.arccosh(1.0 + grad * grad * L * L / (2.0 * vs * vn)) / grad


# ------------------------------------------------------------------ C2

def c2(seed: int = 2):
    """(grid, Vp, Vs, stations) of C2."""
    g = synth.centred_grid(161, 161, 81, 1.0)
    vp = synth.gradient_velocity(g, 5.0, 8.0, 64.0) * synth.checkerboard(g, 20.0, 0.05)
    vs = (vp / np.float32(1.73)).astype(np.float32)
    stations = synth.random_surface_stations(g, np.random.default_rng(seed), 48)
    return g, vp.astype(np.float32), vs, stations


# ------------------------------------------------------------------ C3

@dataclasses.dataclass
class WindowTable:
    par: dict
    origin: tuple        # model node of table node (0, 0, 0)
    radius: float


def c3(count: int = 256, seed: int = 3):
    """(model grid, Vp, [WindowTable]) of C3."""
    rng = np.random.default_rng(seed)
    G = synth.centred_grid(401, 401, 101, 1.0)
    v = synth.gradient_velocity(G, 5.0, 8.0, 64.0) * synth.gaussian_anomalies(G, rng, 3, 0.08, (10.0, 30.0))
    v = np.ascontiguousarray(v, dtype=np.float32)
    dq, df = G.dq_deg(), G.df_deg()
    out = []
    for s in range(count):
        r = float(rng.uniform(40.0, 200.0))
        n_h = int(r) + 1
        n_z = min(int(r), 100) + 1
        # station anywhere on the surface; its box is clipped to the model
        ci, cj = float(rng.uniform(0, G.nx - 1)), float(rng.uniform(0, G.ny - 1))
        ox = int(min(max(round(ci - n_h / 2), 0), G.nx - 1))
        oy = int(min(max(round(cj - n_h / 2), 0), G.ny - 1))
        nx = min(n_h, G.nx - ox)
        ny = min(n_h, G.ny - oy)
        nz = n_z
        # the station sits inside its window, two nodes from a clipped edge (the
        # geodetic-to-geocentric conversion moves it by a fraction of a node)
        ci = min(max(ci, ox + 2.0), ox + nx - 3.0)
        cj = min(max(cj, oy + 2.0), oy + ny - 3.0)
        x0 = round(G.x0 + ox * df, 6)
        y0 = round(G.y0 - oy * dq, 6)
        lon = round(G.x0 + ci * df, 5)
        lat = round(G.y0 - cj * dq, 5)
        par = dict(nx=nx, ny=ny, nz=nz, h=1.0, x0=x0, y0=y0, z0=0.0, dq=dq, df=df,
                   fxs=lon, fys=lat, fzs=0.0, reverse=50)
        out.append(WindowTable(par, (ox, oy, 0), r))
    return G, v, out


# ------------------------------------------------------------------ C4

@dataclasses.dataclass
class Request:
    t_arrival: float
    par: dict
    shape: tuple         # (nz, ny, nx)
    seed: int


def c4_requests(count: int, rate: float, seed: int = 4) -> list[Request]:
    """A Poisson stream: exponential inter-arrival times of mean 1/rate s,
    sizes from 81^3 to 257 x 257 x 129 (nx = ny and nz interpolated)."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for k in range(count):
        t += float(rng.exponential(1.0 / rate))
        s = float(rng.uniform(0.0, 1.0))
        nxy = int(round(81 + s * (257 - 81)))
        nz = int(round(81 + s * (129 - 81)))
        g = synth.centred_grid(nxy, nxy, nz, 1.0)
        st = synth.random_surface_stations(g, rng, 1, prefix=f"R{k:04d}")[0]
        out.append(Request(t, table_par(g, st), (nz, nxy, nxy), int(rng.integers(1 << 31))))
    return out


def c4_velocity(req: Request) -> np.ndarray:
    """The request's velocities: gradient times 1 + three +-3 % separable
    Gaussian blobs (sigma 10-30 km) -- cheap outer products, so the stream
    can be generated at submission time."""
    nz, ny, nx = req.shape
    rng = np.random.default_rng(req.seed)
    z = np.arange(nz, dtype=np.float32)
    y = np.arange(ny, dtype=np.float32)
    x = np.arange(nx, dtype=np.float32)
    v = (5.0 + 3.0 * z / 64.0).astype(np.float32)[:, None, None] * np.ones((1, ny, nx), np.float32)
    f = np.ones((nz, ny, nx), np.float32)
    for _ in range(3):
        cz, cy, cx = rng.uniform(0, nz), rng.uniform(0, ny), rng.uniform(0, nx)
        sg = float(rng.uniform(10.0, 30.0))
        a = np.float32(0.03 if rng.random() < 0.5 else -0.03)
        gz = np.exp(-((z - cz) ** 2) / (2 * sg * sg)).astype(np.float32)
        gy = np.exp(-((y - cy) ** 2) / (2 * sg * sg)).astype(np.float32)
        gx = np.exp(-((x - cx) ** 2) / (2 * sg * sg)).astype(np.float32)
        f += a * gz[:, None, None] * gy[None, :, None] * gx[None, None, :]
    return np.ascontiguousarray(v * f, dtype=np.float32)


# ------------------------------------------------------------------ C5

def c5(stations: int = 1000, seed: int = 5):
    """(grid, Vp, Vs, stations) of C5."""
    rng = np.random.default_rng(seed)
    g = synth.centred_grid(321, 321, 161, 1.0)
    vp = synth.gradient_velocity(g, 5.0, 8.0, 64.0) * synth.layered_perturbation(g, rng, 0.04, 2.0, 12.0)
    vp = np.ascontiguousarray(vp, dtype=np.float32)
    vs = (vp / np.float32(1.73)).astype(np.float32)
    sts = synth.random_surface_stations(g, rng, stations)
    return g, vp, vs, sts


# ------------------------------------------------------------------ c2f

def c2f_coarse(seed: int = 6, nxc: int = 31, nyc: int = 31, nzc: int = 7, step: int = 10, h: float = 1.0):
    """A coarse inversion model for c2f's interpolation (src/c2f.c): nxc x nyc x
    nzc nodes every `step` fine nodes (31 x 31 x 7 at step 10 gives the
    301 x 301 x 61 fine grid, the largest the reference's arrays hold,
    include/common/parameter.h), Vp 5.0 -> 8.0 km/s over 0-64 km times a
    seeded +-4 % perturbation per node, Vs = Vp / 1.73 times +-2 %.  Returns
    (nxc, nyc, nzc, igridx, igridy, igridz, x0, y0, z0, h, velocities P then S)."""
    rng = np.random.default_rng(seed)
    z = np.arange(nzc, dtype=np.float64) * step * h
    vp = (5.0 + 3.0 * np.clip(z / 64.0, 0.0, None))[:, None, None] * np.ones((nzc, nyc, nxc))
    vp = vp * (1.0 + 0.04 * rng.uniform(-1.0, 1.0, size=vp.shape))
    vs = vp / 1.73 * (1.0 + 0.02 * rng.uniform(-1.0, 1.0, size=vp.shape))
    vel = np.concatenate([vp.ravel(), vs.ravel()]).astype(np.float32)
    ig = lambda n: np.full(n - 1, step, dtype=np.int32)  # noqa: E731
    return nxc, nyc, nzc, ig(nxc), ig(nyc), ig(nzc), 0.0, 0.0, 0.0, h, vel


def c2f_spec_text(nxc, nyc, nzc, igridx, igridy, igridz, x0, y0, z0, h, oldvfil="coarse.mod",
                  finevel="fine.mod") -> str:
    """A c2f spec file (keys of src/c2f.c:81-83, 170-257); c2f reads nxc
    values from each igrid line, the last one unused."""
    def row(name, arr):
        return name + " " + " ".join(str(int(v)) for v in list(arr) + [0]) + "\n"
    return (f"nxc {nxc}\nnyc {nyc}\nnzc {nzc}\nh {h!r}\noldvfil {oldvfil}\ntgrdfil grid.txt\n"
            f"finevel {finevel}\nx0 {x0!r}\ny0 {y0!r}\nz0 {z0!r}\n"
            + row("igridx", igridx) + row("igridy", igridy) + row("igridz", igridz))


# ------------------------------------------------------------------ location

def loc_params(g: synth.GridSpec, **kw) -> dict:
    """sphfdloc's search grid for the tables of grid g, set up as the program
    does (src/sphfdloc.c:382-393: radians, geocentric colatitude, spacings
    from h when dq and df are 0), with its parameter defaults (111-128)."""
    degrad = float(np.float32(0.017453292))     # sphfdloc.c:26, a float constant
    hpi = float(np.float32(1.570796))
    colat = hpi - synth.glath(g.y0 * degrad, g.z0)[0]
    d = dict(nx=g.nx, ny=g.ny, nz=g.nz, x0=g.x0 * degrad, y0=colat, z0=g.z0, h=g.h, dq=g.h / 6371.0,
             df=abs(g.h / (6371.0 * math.sin(colat))), kmin=2, ndiv=20, ndiv2=20, ivs=1, vpvs=1.78, nthres=8,
             resthres=0.5, resthrep=5.0, stdmax=15.0)
    d.update(kw)
    return d


def loc_events(g: synth.GridSpec, time_at, ntables: int, count: int = 256, seed: int = 7,
               picks: tuple = (24, 40)) -> list:
    """Synthetic events for the tables 0..ntables-1 (P and S of each station in
    turn): a source node below layer 2, between picks[0] and picks[1] distinct
    station phases, each arrival the table time at the source plus an origin
    offset and 0.05 s noise; about one pick in twenty late by 2-6 s (an
    outlier) and one in thirty flagged unusable.  time_at(q, k, j, i) gives
    table q's time at a node."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        i, j, k = int(rng.integers(0, g.nx)), int(rng.integers(0, g.ny)), int(rng.integers(2, g.nz - 1))
        n = int(rng.integers(picks[0], picks[1] + 1))
        qs = rng.choice(ntables, size=n, replace=False)
        t0 = rng.uniform(0.0, 5.0)
        obs = np.array([time_at(int(q), k, j, i) + t0 + rng.normal(0.0, 0.05) for q in qs], dtype=np.float32)
        late = rng.random(n) < 0.05
        obs[late] += rng.uniform(2.0, 6.0, int(late.sum())).astype(np.float32)
        good = (rng.random(n) >= 1.0 / 30).astype(np.int32)
        rwt = rng.choice(np.array([0.05, 0.1, 0.2], dtype=np.float32), size=n)
        out.append(dict(qs=qs, obstime=obs, pwt=(np.float32(1.0) / (rwt * rwt)).astype(np.float32),
                        phase="".join("P" if q % 2 == 0 else "S" for q in qs), isgood=good, src=(i, j, k)))
    return out
