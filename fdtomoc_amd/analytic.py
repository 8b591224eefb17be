"""Analytic FP64 travel times on sphfd's spherical grid, for stating the finite-
difference error of the solver per workload.

  homogeneous_time   straight chord / v: in a uniform medium the first
                     arrival runs along the straight line between source and
                     node (the reference initialises its source cube with these
                     Cartesian chords, src/sphfd.c:727-749)
  gradient_time      a spherically symmetric medium whose velocity grows
                     linearly with depth below a reference radius R,
                     v(r) = v0 + g (R - r) = A - g r, by ray theory: with the
                     ray parameter p = r sin(i) / v(r) constant along a ray
                     and eta(r) = r / v(r),
                         Delta(p) = int p dr / (r sqrt(eta^2 - p^2)),
                         T(p)     = int eta^2 dr / (r sqrt(eta^2 - p^2)),
                     over the ray's path between the source radius and the
                     node radius, through the turning radius r_t (eta(r_t) = p,
                     r_t = p A / (1 + p g)) when the ray turns.  Here
                     d eta / dr = (v + g r) / v^2 > 0, so each (radius,
                     epicentral angle) is reached by exactly one ray: a direct
                     one (no turning point between the two radii) up to the
                     angle of the ray that is horizontal at the deeper end,
                     a turning one beyond it.  The integrals are evaluated with
                     r = r_t + u^2 (which removes the inverse-square-root
                     singularity at the turning point) by Gauss-Legendre
                     quadrature, and the angle equation is solved for r_t by
                     bisection, vectorised over nodes.

Positions follow the reference's grid: node (i, j, k) at longitude
x0 + i df, geocentric colatitude y0 + j dq and radius z0r - k h
(src/sphfd.c:151-156); the source at (fxss, fyss, fzss) as sphfd derives it
(src/sphfd.c:437-506).  All arithmetic is FP64.
"""
from __future__ import annotations

import math

import numpy as np

from . import synth

_GL = {n: np.polynomial.legendre.leggauss(n) for n in (24, 64)}


def geometry(par: dict) -> dict:
    """A table's grid and source in radians and km, as sphfd derives them
    (src/sphfd.c:415-509): geocentric colatitude and top radius of the grid
    from glath, dq and df from h when the par values leave them 0, the
    source's longitude, geocentric colatitude and radius."""
    D = synth.DEGRAD
    lat0, z0r = synth.glath(par["y0"] * D, par.get("z0", 0.0))
    y0 = synth.HPI - lat0
    h = par["h"]
    dq = par.get("dq", 0.0) * D or h / synth.REARTH
    df = par.get("df", 0.0) * D or abs(h / (synth.REARTH * math.sin(y0)))
    lats, rs = synth.glath(par["fys"] * D, par["fzs"])
    return {"h": h, "dq": dq, "df": df, "x0": par["x0"] * D, "y0": y0, "z0r": z0r,
            "fxss": par["fxs"] * D, "fyss": synth.HPI - lats, "fzss": rs}


def node_positions(geom: dict, nx: int, ny: int, nz: int, sel: np.ndarray | None = None):
    """(radius, colatitude, longitude) of the nodes (all, or the flat indices
    `sel` in file order: x fastest, then y, then z)."""
    if sel is None:
        sel = np.arange(nx * ny * nz)
    k, rem = np.divmod(np.asarray(sel, dtype=np.int64), nx * ny)
    j, i = np.divmod(rem, nx)
    return geom["z0r"] - geom["h"] * k, geom["y0"] + geom["dq"] * j, geom["x0"] + geom["df"] * i


def _unit(colat, lon):
    s = np.sin(colat)
    return np.stack(np.broadcast_arrays(np.cos(colat), s * np.cos(lon), s * np.sin(lon)))


def angle_between(colat, lon, colat_s: float, lon_s: float):
    """Epicentral angle between points (colat, lon) and one point (colat_s,
    lon_s), from the chord of the unit vectors (stable near 0)."""
    u = _unit(np.asarray(colat, np.float64), np.asarray(lon, np.float64))
    s = _unit(np.float64(colat_s), np.float64(lon_s)).reshape(3, *([1] * (u.ndim - 1)))
    c = np.sqrt(np.sum((u - s) ** 2, axis=0))
    return 2.0 * np.arcsin(np.minimum(c / 2.0, 1.0))


def chord(geom: dict, r, colat, lon):
    """Straight-line distance (km) from the source to points (r, colat, lon)."""
    p = r * _unit(colat, lon)
    s = geom["fzss"] * _unit(np.float64(geom["fyss"]), np.float64(geom["fxss"]))
    d = p - s.reshape(3, *([1] * (p.ndim - 1)))
    return np.sqrt(np.sum(d * d, axis=0))


def homogeneous_time(geom: dict, v: float, r, colat, lon):
    return chord(geom, r, colat, lon) / v


def _segment(A, g, rt, lo, hi, npts=64):
    """(Delta, T) of the ray with turning radius rt between radii lo <= hi
    (rt <= lo), integrated in u = sqrt(r - rt) with npts Gauss-Legendre points."""
    _GL_X, _GL_W = _GL[npts]
    p = rt / (A - g * rt)
    u0 = np.sqrt(np.maximum(lo - rt, 0.0))
    u1 = np.sqrt(np.maximum(hi - rt, 0.0))
    half = 0.5 * (u1 - u0)
    mid = 0.5 * (u1 + u0)
    u = mid[..., None] + half[..., None] * _GL_X
    r = rt[..., None] + u * u
    eta = r / (A - g * r)
    # eta^2 - p^2 = (eta - p)(eta + p), with eta - p = A (r - rt) / ((A - g r)(A - g rt)) exactly
    diff = A * (u * u) / ((A - g * r) * (A - g * rt[..., None]))
    den = r * np.sqrt(diff * (eta + p[..., None]))
    jac = 2.0 * u
    dd = np.sum(_GL_W * (p[..., None] * jac / den), axis=-1) * half
    tt = np.sum(_GL_W * (eta * eta * jac / den), axis=-1) * half
    return dd, tt


def gradient_time(geom: dict, v0: float, g: float, R: float, r, colat, lon, iters: int = 56):
    """First-arrival time (s) in v(r) = v0 + g (R - r) from the source of
    `geom` to the points (r, colat, lon); g > 0 (km/s per km)."""
    A = v0 + g * R
    r = np.asarray(r, dtype=np.float64)
    rs = float(geom["fzss"])
    delta = angle_between(colat, lon, geom["fyss"], geom["fxss"])
    r, delta = np.broadcast_arrays(r, delta)
    lo = np.minimum(r, rs)
    hi = np.maximum(r, rs)
    # the ray horizontal at the deeper end separates direct and turning rays
    dmax, _ = _segment(A, g, lo.copy(), lo, hi)
    direct = delta <= dmax
    out = np.empty(r.shape, dtype=np.float64)
    # direct rays: rt in (0, lo], Delta increasing in rt
    d_lo, d_hi, d_t = lo[direct], hi[direct], delta[direct]
    a, b = np.zeros(d_lo.shape), d_lo.copy()
    for _ in range(iters):
        m = 0.5 * (a + b)
        dd, _ = _segment(A, g, m, d_lo, d_hi, 24)
        far = dd > d_t
        b = np.where(far, m, b)
        a = np.where(far, a, m)
    out[direct] = _segment(A, g, 0.5 * (a + b), d_lo, d_hi)[1]
    # turning rays: rt below lo, Delta = (rt -> lo) + (rt -> hi), growing as rt decreases
    t_lo, t_hi, t_t = lo[~direct], hi[~direct], delta[~direct]
    a, b = np.maximum(t_lo - 3000.0, 1.0), t_lo.copy()
    for _ in range(iters):
        m = 0.5 * (a + b)
        dd = _segment(A, g, m, m, t_lo, 24)[0] + _segment(A, g, m, m, t_hi, 24)[0]
        far = dd > t_t
        a = np.where(far, m, a)
        b = np.where(far, b, m)
    m = 0.5 * (a + b)
    out[~direct] = _segment(A, g, m, m, t_lo)[1] + _segment(A, g, m, m, t_hi)[1]
    return out


def flat_gradient_time(v_s: float, v_n, g: float, L):
    """Closed form in a flat medium with v = v0 + g z (circular rays):
    t = arccosh(1 + g^2 L^2 / (2 v_s v_n)) / g, L the straight distance."""
    return np.arccosh(1.0 + g * g * L * L / (2.0 * v_s * v_n)) / g


def table_error(table: np.ndarray, geom: dict, exact_fn, sample: int = 0, seed: int = 0) -> dict:
    """Max and mean |T_table - T_exact| over the nodes (or a seeded sample of
    `sample` nodes), excluding nodes the solver never reached."""
    nz, ny, nx = table.shape
    n = table.size
    sel = np.arange(n) if not sample or sample >= n else np.sort(
        np.random.default_rng(seed).choice(n, size=sample, replace=False))
    r, q, f = node_positions(geom, nx, ny, nz, sel)
    exact = exact_fn(r, q, f)
    t = table.reshape(-1)[sel].astype(np.float64)
    ok = t < 1e9
    err = np.abs(t[ok] - exact[ok])
    return {"max": float(err.max()), "mean": float(err.mean()), "nodes": int(ok.sum())}
