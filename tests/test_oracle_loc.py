"""Pins of the oracle's sphfdloc location (oracle/loc_oracle.c) by what the
mathematics fixes: the reference program is not built here (about 60 GB of
static arrays, text-only output), so find_time is checked against tables that
trilinear interpolation reproduces exactly, and the search against synthetic
tables whose source is known (src/sphfdloc.c:752-1130, 1305-1345)."""
import numpy as np
import pytest

import oracle_bind as ob


def _grid(nx, ny, nz, **kw):
    d = dict(nx=nx, ny=ny, nz=nz, x0=0.0, y0=0.0, z0=0.0, df=1.0, dq=1.0, h=1.0)
    d.update(kw)
    return ob.loc_grid(**d)


def test_find_time_reproduces_trilinear_tables():
    """A table in span{1, x, y, z, xy, xz, yz, xyz} of the node indices is
    reproduced exactly at nodes, inside cells and beyond the clamped faces
    (the last cell's polynomial extends); binary spacings keep every operand
    exact, so any wrong corner, weight or index fails."""
    nx, ny, nz = 7, 5, 6
    g = _grid(nx, ny, nz, x0=-1.5, y0=0.25, z0=3.0, df=0.25, dq=0.5, h=2.0)
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")

    def f(a, b, c):
        return 3.0 * a - 2.0 * b + 5.0 * c + 0.5 * a * b - 0.25 * b * c + 1.0 * a * c + 0.125 * a * b * c + 7.0

    t = f(i, j, k).astype(np.float32)
    assert np.array_equal(t.astype(np.float64), f(i, j, k))
    pts = [(0, 0, 0), (6, 4, 5), (2.5, 1.25, 3.75), (0.125, 3.5, 0.5), (5.75, 0.0, 4.5), (-0.5, 2.0, 1.0),
           (7.5, 4.5, 6.5), (-1.0, -0.5, -0.25), (3.0, 6.25, 2.5)]
    for a, b, c in pts:
        x, y, z = g.x0 + a * g.df, g.y0 + b * g.dq, g.z0 + c * g.h
        assert ob.loc_find_time(g, t, x, y, z) == f(a, b, c), (a, b, c)


def _synthetic(seed, nx=21, ny=17, nz=12, src=(9.0, 7.0, 6.0), nsta=12, v=1.0, zmax=0.5):
    """Tables of straight-ray times (index units, speed v) from stations above
    the top face down to depth zmax, and the exact arrival times from src."""
    rng = np.random.default_rng(seed)
    sta = np.stack([rng.uniform(0, nx - 1, nsta), rng.uniform(0, ny - 1, nsta), rng.uniform(-2.0, zmax, nsta)], 1)
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    tabs = [(np.sqrt((i - s[0]) ** 2 + (j - s[1]) ** 2 + (k - s[2]) ** 2) / v).astype(np.float32) for s in sta]
    obs = np.array([np.sqrt(((np.array(src) - s) ** 2).sum()) / v for s in sta], dtype=np.float32)
    return sta, tabs, obs


def test_search_finds_a_node_source():
    """Exact times from a node: the node search returns it with deviation 0;
    no refinement point beats 0, so the hypocentre is the node itself, and the
    computed times are the table values there."""
    src = (9, 7, 6)
    _, tabs, obs = _synthetic(1, src=src)
    n = len(tabs)
    g = _grid(21, 17, 12)
    r = ob.locate(g, tabs, obs, np.ones(n), "P" * n, np.ones(n))
    assert r["node"] == src and r["stdmin"] == 0.0 and r["avrmin"] == 0.0
    assert (r["exm"], r["eym"], r["ezm"]) == tuple(float(c) for c in src)
    assert r["status"] == 1 and r["passes"] == 1 and r["ngood"] == n
    assert np.array_equal(r["tps"], np.array([t[src[2], src[1], src[0]] for t in tabs]))
    assert not r["res"].any()


def test_search_refines_toward_a_subgrid_source():
    """A source between nodes: the refinements move the hypocentre from the
    node to within a small fraction of a cell of the true position, and never
    increase the deviation found at the node (stations spread through the
    volume, as in boreholes, so that depth is resolved)."""
    src = (9.37, 7.61, 5.28)
    _, tabs, obs = _synthetic(2, src=src, nsta=16, zmax=11.0)
    n = len(tabs)
    g = _grid(21, 17, 12)
    r = ob.locate(g, tabs, obs, np.ones(n), "P" * n, np.ones(n))
    e = np.array([r["exm"], r["eym"], r["ezm"]])
    assert np.abs(e - np.array(src)).max() < 0.08, e
    assert r["node"] == (9, 8, 5)
    g0 = _grid(21, 17, 12, ndiv=1, ndiv2=1)   # the node search alone (refinement boxes of nodes)
    r0 = ob.locate(g0, tabs, obs, np.ones(n), "P" * n, np.ones(n))
    assert r["stdmin"] <= r0["stdmin"]


def test_first_of_tied_minima_wins():
    """Stations mirrored about x = 10 make the deviation symmetric; a source at
    the mirror of node 4 is located at node 4, the first in scan order (the
    strict comparison of sphfdloc.c:793)."""
    nx, ny, nz = 21, 9, 8
    rng = np.random.default_rng(3)
    half = np.stack([rng.uniform(0, 9, 5), rng.uniform(0, ny - 1, 5), rng.uniform(-1, 0, 5)], 1)
    sta = np.concatenate([half, half * [-1, 1, 1] + [nx - 1, 0, 0]])
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    tabs = [np.sqrt((i - s[0]) ** 2 + (j - s[1]) ** 2 + (k - s[2]) ** 2).astype(np.float32) for s in sta]
    src = np.array([16.0, 4.0, 5.0])
    obs = np.array([np.sqrt(((src - s) ** 2).sum()) for s in sta], dtype=np.float32)
    # the same times with the stations swapped pairwise are the mirrored source's
    n = len(tabs)
    r = ob.locate(_grid(nx, ny, nz, ndiv=1, ndiv2=1), tabs, obs, np.ones(n), "P" * n, np.ones(n))
    assert r["stdmin"] == 0.0
    tabs_m = tabs[5:] + tabs[:5]
    r2 = ob.locate(_grid(nx, ny, nz, ndiv=1, ndiv2=1), tabs_m, obs, np.ones(n), "P" * n, np.ones(n))
    assert r2["node"] == (4, 4, 5) and r["node"] == (16, 4, 5)
    # both mirrors fit exactly when each pick's table is the average of the pair
    sym = [((a.astype(np.float64) + b) / 2).astype(np.float32) for a, b in zip(tabs, tabs_m)]
    obs_s = np.array([t[5, 4, 16] for t in sym], dtype=np.float32)
    assert np.array_equal(obs_s, np.array([t[5, 4, 4] for t in sym]))
    r3 = ob.locate(_grid(nx, ny, nz, ndiv=1, ndiv2=1), sym, obs_s, np.ones(n), "P" * n, np.ones(n))
    assert r3["node"] == (4, 4, 5) and r3["stdmin"] == 0.0


def test_outlier_removed_and_flagged_pick_restored():
    """A pick 5 s late is dropped (its residual exceeds 0.5 s and 5 %) and the
    event is located again without it; a consistent pick flagged unusable on
    input is brought back by the one retry (sphfdloc.c:1080-1126)."""
    src = (9, 7, 6)
    _, tabs, obs = _synthetic(4, src=src, nsta=14)
    n = len(tabs)
    g = _grid(21, 17, 12)
    bad = obs.copy()
    bad[3] += 5.0
    r = ob.locate(g, tabs, bad, np.ones(n), "P" * n, np.ones(n))
    assert r["isgood"][3] == 0 and r["isgood"].sum() == n - 1 and r["ngood"] == n - 1
    assert r["passes"] >= 2 and r["node"] == src and r["stdmin"] == 0.0
    assert r["res"][3] == pytest.approx(5.0, abs=1e-5)
    flags = np.ones(n, dtype=np.int32)
    flags[5] = 0
    r = ob.locate(g, tabs, obs, np.ones(n), "P" * n, flags)
    assert r["isgood"].all() and r["passes"] == 2 and r["ngood"] == n and r["node"] == src


def test_status_thresholds_and_s_scaling():
    """Too few usable picks gives status 0, a deviation above stdmax status 2;
    with ivs = 0 an S pick's time is vpvs times its table, so S times made
    that way locate the source and the same picks read as unscaled do not."""
    src = (9, 7, 6)
    _, tabs, obs = _synthetic(5, src=src, nsta=10)
    n = len(tabs)
    assert ob.locate(_grid(21, 17, 12, nthres=11), tabs, obs, np.ones(n), "P" * n, np.ones(n))["status"] == 0
    noisy = obs + np.random.default_rng(6).normal(0, 0.05, n).astype(np.float32)
    assert ob.locate(_grid(21, 17, 12, stdmax=1e-6), tabs, noisy, np.ones(n), "P" * n, np.ones(n))["status"] == 2
    vpvs = np.float32(1.78)
    ph = "PSPSPSPSPS"
    sobs = np.array([o * vpvs if c == "S" else o for o, c in zip(obs, ph)], dtype=np.float32)
    r = ob.locate(_grid(21, 17, 12, ivs=0), tabs, sobs, np.ones(n), ph, np.ones(n))
    assert r["node"] == src and r["stdmin"] < 1e-5
    r = ob.locate(_grid(21, 17, 12, ivs=1), tabs, sobs, np.ones(n), ph, np.ones(n))
    assert r["stdmin"] > 0.1


def test_deviation_formula_closed_form():
    """Constant tables give every trial point the same residuals r = obs - c,
    so the first node wins and the refinements (ndiv = 1: points on nodes,
    binary spacings) cannot beat it; the deviation and the mean residual then
    equal sphfdloc's weighted variance (sphfdloc.c:786-799) evaluated
    independently in float64, with every pick counted in the divisor on the
    first pass: sqrt(|sum w r^2 - (sum w r)^2 / sum w| / ((n - 4) sum w / n))
    and sum w r / sum w."""
    rng = np.random.default_rng(17)
    nx, ny, nz, n = 4, 3, 5, 11
    c = rng.uniform(2.0, 20.0, n)
    tabs = [np.full((nz, ny, nx), np.float32(v)) for v in c]
    r = rng.normal(0.0, 0.2, n)
    obs = (np.float32(c) + np.float32(r)).astype(np.float32)
    w = rng.choice([1.0, 4.0, 25.0, 100.0], n).astype(np.float32)
    g = _grid(nx, ny, nz, ndiv=1, ndiv2=1, kmin=1, resthres=100.0)
    out = ob.locate(g, tabs, obs, w, "P" * n, np.ones(n))
    rr = obs.astype(np.float64) - np.float32(c).astype(np.float64)
    wd = w.astype(np.float64)
    sw, swr, swr2 = wd.sum(), (wd * rr).sum(), (wd * rr * rr).sum()
    dev = np.sqrt(abs(swr2 - swr * swr / sw) / ((n - 4) * sw / n))
    assert out["node"] == (0, 0, 1) and out["passes"] == 1
    assert out["stdmin"] == pytest.approx(dev, rel=2e-5)
    assert out["avrmin"] == pytest.approx(swr / sw, rel=2e-5, abs=1e-7)
    assert np.allclose(out["res"], rr, atol=1e-6)
    # one pick flagged unusable and 50 s off (so the retry does not restore
    # it): absent from the sums but counted in the first pass's divisor
    # (sphfdloc.c:683), which is the only pass
    obs2 = obs.copy()
    obs2[3] += np.float32(50.0)
    flags = np.ones(n, dtype=np.int32)
    flags[3] = 0
    g2 = _grid(nx, ny, nz, ndiv=1, ndiv2=1, kmin=1, resthres=1.0, resthrep=0.01)
    out = ob.locate(g2, tabs, obs2, w, "P" * n, flags)
    keep = np.arange(n) != 3
    sw, swr, swr2 = wd[keep].sum(), (wd * rr)[keep].sum(), (wd * rr * rr)[keep].sum()
    dev_n = np.sqrt(abs(swr2 - swr * swr / sw) / ((n - 4) * sw / n))
    dev_n1 = np.sqrt(abs(swr2 - swr * swr / sw) / ((n - 5) * sw / (n - 1)))
    assert out["passes"] == 1 and out["isgood"][3] == 0 and out["ngood"] == n - 1
    assert out["stdmin"] == pytest.approx(dev_n, rel=2e-5)
    assert abs(dev_n - dev_n1) > 1e-3 * dev_n


def test_refinements_land_exactly_on_a_lattice_source():
    """Tables linear in the node indices (small integer gradients, exact in
    float and reproduced exactly by find_time), binary spacings and
    divisions (ndiv = ndiv2 = 4: steps 1/4 and 1/16), and a source on the
    second refinement's lattice: the residuals vanish exactly at the source
    and nowhere else, so the node search, the first box (the best node +- 1
    in quarter steps) and the second box (+- one quarter in sixteenth
    steps) must land on it exactly, with deviation 0."""
    rng = np.random.default_rng(23)
    nx, ny, nz, n = 12, 10, 9, 10
    grads = rng.integers(-3, 4, size=(n, 3)).astype(np.float64)
    grads[:3] = np.eye(3) * 2 + 1          # well spread directions
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    tabs = [(gv[0] * i + gv[1] * j + gv[2] * k + 40.0).astype(np.float32) for gv in grads]
    for src in ((5.3125, 4.5625, 4.1875), (2.75, 6.0625, 3.9375), (8.0625, 2.25, 5.5)):
        obs = np.array([gv @ np.array(src) + 40.0 for gv in grads], dtype=np.float32)
        assert np.array_equal(obs.astype(np.float64), np.array([gv @ np.array(src) + 40.0 for gv in grads]))
        g = _grid(nx, ny, nz, ndiv=4, ndiv2=4, resthres=100.0)
        out = ob.locate(g, tabs, obs, np.ones(n), "P" * n, np.ones(n))
        assert (out["exm"], out["eym"], out["ezm"]) == src, (src, out["exm"], out["eym"], out["ezm"])
        assert out["stdmin"] == 0.0
