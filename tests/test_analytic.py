"""The analytic FP64 travel times (fdtomoc_amd/analytic.py) that state the
solver's finite-difference error, pinned by what the mathematics fixes:

  * a vanishing gradient gives the straight-chord time of a uniform medium;
  * a vertical ray has the closed form ln(v_n / v_s) / g;
  * far from the centre (R = 1e7 km) the sphere is flat and the times approach
    the circular-ray closed form arccosh(1 + g^2 L^2 / (2 v_s v_n)) / g;
  * source and receiver can be exchanged;
  * at the Earth's radius, rays traced numerically through the same medium
    (RK4 on the ray equations in the great-circle plane, take-off angle found
    by bisection) arrive at the same times;
  * the grid geometry equals the oracle's (sphfd.c:415-509).
"""
import math

import numpy as np
import pytest

import oracle_bind as ob
from fdtomoc_amd import analytic as an


def test_uniform_limit_is_the_chord_time():
    rng = np.random.default_rng(0)
    geom = dict(fzss=6375.0, fyss=1.0, fxss=2.0)
    r = 6375.0 - rng.uniform(0, 64, 500)
    q = 1.0 + rng.uniform(-0.01, 0.01, 500)
    f = 2.0 + rng.uniform(-0.01, 0.01, 500)
    t = an.gradient_time(geom, 5.0, 1e-9, 6375.0, r, q, f)
    assert np.abs(t - an.homogeneous_time(geom, 5.0, r, q, f)).max() < 1e-6


def test_vertical_ray_closed_form():
    R, g = 6371.0, 3.0 / 64.0
    geom = dict(fzss=R, fyss=1.0, fxss=2.0)
    rv = R - np.array([1.0, 10.0, 50.0, 64.0])
    t = an.gradient_time(geom, 5.0, g, R, rv, np.ones(4), np.full(4, 2.0))
    assert np.abs(t - np.log((5.0 + g * (R - rv)) / 5.0) / g).max() < 1e-8


def test_flat_limit():
    rng = np.random.default_rng(1)
    R, g = 1e7, 3.0 / 64.0
    geom = dict(fzss=R, fyss=1.0, fxss=2.0)
    d = rng.uniform(0, 64, 500)
    r = R - d
    q = 1.0 + rng.uniform(0, 300, 500) / R
    f = np.full(500, 2.0)
    t = an.gradient_time(geom, 5.0, g, R, r, q, f)
    flat = an.flat_gradient_time(5.0, 5.0 + g * d, g, an.chord(geom, r, q, f))
    assert (np.abs(t - flat) / flat).max() < 1e-5


def test_reciprocity():
    R, g = 6371.0, 3.0 / 64.0
    a = an.gradient_time(dict(fzss=R - 30.0, fyss=1.0, fxss=2.0), 5.0, g, R, np.array([R - 2.0]),
                         np.array([1.02]), np.array([2.01]))
    b = an.gradient_time(dict(fzss=R - 2.0, fyss=1.02, fxss=2.01), 5.0, g, R, np.array([R - 30.0]),
                         np.array([1.0]), np.array([2.0]))
    assert abs(a[0] - b[0]) < 1e-9


def _shoot(A, g, rs, rn, delta, nstep=4000):
    """Travel time of the ray from (radius rs, angle 0) to (rn, delta) in
    v(r) = A - g r, traced with RK4 in the ray's plane: dX/ds = P, dP/ds =
    u grad u, dT/ds = u^2 (u = 1/v, s the ray's sigma parameter)."""
    def f(s):
        x, y, px, py = s
        r = math.hypot(x, y)
        v = A - g * r
        u = 1.0 / v
        du = g / (v * v)
        return np.array([px, py, u * du * x / r, u * du * y / r]), u * u

    def trace(ang):
        v = A - g * rs
        s = np.array([0.0, rs, math.sin(ang) / v, -math.cos(ang) / v])
        h = (math.hypot(rn * math.sin(delta), rn * math.cos(delta) - rs) * 1.5 + 1) / nstep * v
        T = 0.0
        for _ in range(nstep * 20):
            prev, prevT = s.copy(), T
            k1, t1 = f(s)
            k2, t2 = f(s + 0.5 * h * k1)
            k3, t3 = f(s + 0.5 * h * k2)
            k4, t4 = f(s + h * k3)
            s = s + h / 6 * (k1 + 2 * k2 + 2 * k3 + k4)
            T += h / 6 * (t1 + 2 * t2 + 2 * t3 + t4)
            a = math.atan2(s[0], s[1])
            if a >= delta:
                a0 = math.atan2(prev[0], prev[1])
                w = (delta - a0) / (a - a0)
                return (1 - w) * math.hypot(prev[0], prev[1]) + w * math.hypot(s[0], s[1]), (1 - w) * prevT + w * T
        return None, None

    lo, hi = 1e-6, 1.2   # take-off angle from the downward vertical; downgoing rays
    for _ in range(50):
        m = 0.5 * (lo + hi)
        r, _ = trace(m)
        if r is None or r < rn:
            lo = m
        else:
            hi = m
    return trace(0.5 * (lo + hi))[1]


@pytest.mark.parametrize("depth,offset", [(10.0, 50.0), (40.0, 150.0), (64.0, 20.0)])
def test_matches_traced_rays(depth, offset):
    R, g, v0 = 6371.0, 3.0 / 64.0, 5.0
    geom = dict(fzss=R, fyss=1.0, fxss=2.0)
    t = an.gradient_time(geom, v0, g, R, np.array([R - depth]), np.array([1.0 + offset / R]), np.array([2.0]))[0]
    assert abs(t - _shoot(v0 + g * R, g, R, R - depth, offset / R)) < 1e-6


def test_geometry_matches_oracle():
    from fdtomoc_amd import workloads
    g, v, st = workloads.c1(0)
    par = workloads.table_par(g, st)
    mine = an.geometry(par)
    ref = ob.geometry(ob.Par(**par))
    for k in ("h", "dq", "df", "x0", "y0", "z0r", "fxss", "fyss", "fzss"):
        assert mine[k] == pytest.approx(ref[k], rel=1e-14, abs=1e-14), k
