"""Seeded random sphfd cases for parity tests (small grids the oracle and the
compiled reference finish in well under a second).

Each case draws a grid size, spacing and location, one of four velocity
families (gradient x checkerboard, gradient x Gaussian anomalies, two-layer
with a strong contrast that triggers head waves and reverse passes, gradient x
random layers), a source anywhere in the model (often at the surface), and
sweep options the reference exposes (NCUBE, reverse, headpref, growth rates,
maxoff).
"""
from __future__ import annotations

import numpy as np

from fdtomoc_amd import synth
from oracle_bind import Par


def make_case(seed: int, nmin: int = 6, nmax: int = 26):
    rng = np.random.default_rng(seed)
    nx, ny, nz = [int(rng.integers(nmin, nmax)) for _ in range(3)]
    h = float(rng.choice([0.5, 1.0, 2.0]))
    g = synth.centred_grid(nx, ny, nz, h, lat=float(rng.uniform(-60, 60)), lon=float(rng.uniform(-170, 170)))
    kind = rng.integers(0, 4)
    if kind == 0:
        v = synth.gradient_velocity(g, 5.0, 8.0, 64.0) * synth.checkerboard(g, cell_km=float(rng.uniform(2, 8)), amp=0.05)
    elif kind == 1:
        v = synth.gradient_velocity(g, float(rng.uniform(3, 6)), float(rng.uniform(6, 9)), float(rng.uniform(10, 40))) \
            * synth.gaussian_anomalies(g, rng, 3, 0.08, (2, 6))
    elif kind == 2:
        d = np.arange(nz) * h
        zb = rng.uniform(0.2, 0.8) * nz * h
        v = np.where(d < zb, rng.uniform(3, 5), rng.uniform(6, 8)).astype(np.float32)[:, None, None] \
            * np.ones((nz, ny, nx), np.float32)
    else:
        v = synth.gradient_velocity(g) * synth.layered_perturbation(g, rng, 0.15, h, 4 * h)
    v = np.ascontiguousarray(v, dtype=np.float32)
    fi, fj = rng.uniform(0, nx - 1), rng.uniform(0, ny - 1)
    fk = rng.uniform(0, nz - 1) if rng.random() < 0.5 else 0.0
    lon, lat = g.node_lonlat(fi, fj)
    par = Par(nx=nx, ny=ny, nz=nz, h=h, x0=g.x0, y0=g.y0, z0=g.z0, fxs=round(lon, 5), fys=round(lat, 5),
              fzs=round(fk * h, 3), ncube=int(rng.integers(1, 4)), reverse=int(rng.choice([0, 1, 3, 50])),
              headpref=int(rng.choice([0, 6, 6, 6])))
    if rng.random() < 0.15:
        par.kminus = 2
    if rng.random() < 0.15:
        par.iplus = 2
    if rng.random() < 0.1:
        par.maxoff = float(rng.uniform(2, 10)) * h
    return par, v


def par_dict(par: Par) -> dict:
    """fdgpu_par fields of an oracle Par."""
    keys = ["nx", "ny", "nz", "h", "dq", "df", "x0", "y0", "z0", "iminus", "iplus", "jminus", "jplus", "kminus",
            "kplus", "floatsrc", "srctype", "srcwall", "fxs", "fys", "fzs", "xs", "ys", "zs", "ncube", "reverse",
            "headpref", "parse", "fill", "maxoff", "invh", "invq", "invf", "invx0", "invy0", "invz0"]
    return {k: getattr(par, k) for k in keys}
