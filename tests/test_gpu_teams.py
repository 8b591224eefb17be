"""GPU: wavefront passes shared by teams of workgroups (fdgpu_set_team).

Sixteen of C3's windows (the eight largest, 3.3-4.0 M nodes, and eight of
median size) in one launch with every pass as a wavefront: alone (no teams),
with fixed team sizes of 2, 5 and 16 workgroups, and with the planned sizes
at two loads.  Every run must give the lone run's tables bit for bit, with
the same pass counts, no table handed back to a face-by-face solve, and the
teams it asked for; the largest window is also checked against the oracle's
recorded table.  Teams never change a result: a round's nodes are
independent, whichever workgroup evaluates them."""
from __future__ import annotations

import numpy as np
import pytest

import config_cases as cc
from test_gpu_configs import GOLDEN, _check

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    from fdtomoc_amd import fdgpu
    c = fdgpu.Context(0)
    yield c
    c.close()


@pytest.fixture(scope="module")
def windows():
    G, v, wins, check = cc.c3()
    sizes = np.array([w.par["nx"] * w.par["ny"] * w.par["nz"] for w in wins])
    order = np.argsort(-sizes, kind="stable")
    pick = [int(q) for q in list(order[:8]) + list(order[120:128])]
    return v, wins, pick, int(order[0])


def _solve(ctx, v, wins, pick, team):
    from fdtomoc_amd import fdgpu
    m = ctx.model(v)
    ctx.set_wave(3)
    ctx.set_team(*team)
    before = ctx.wave_again()
    try:
        ids = ctx.solve([fdgpu.TableSpec(wins[q].par, m, wins[q].origin) for q in pick], fdgpu.BATCHED)
    finally:
        ctx.set_wave(0)
        ctx.set_team(0, 0)
    again = ctx.wave_again() - before
    out = []
    for q, tid in zip(pick, ids):
        shape = (wins[q].par["nz"], wins[q].par["ny"], wins[q].par["nx"])
        out.append((ctx.read(tid, shape), ctx.stats(tid)))
        ctx.release(tid)
    ctx.release_model(m)
    return out, again


def test_teams_give_the_lone_tables(ctx, windows):
    v, wins, pick, largest = windows
    lone, again = _solve(ctx, v, wins, pick, (1, 0))
    assert again == 0
    assert all(s["team"] == 1 and s["wave_passes"] == s["passes"] for _, s in lone)
    expected = cc.Expected(GOLDEN)
    k = pick.index(largest)
    _check(expected, f"c3_w{largest}", wins[largest].par, cc.c3_window(v, wins[largest]), lone[k][0], lone[k][1])
    for team in [(2, 0), (5, 0), (16, 0), (0, 0), (0, 150)]:
        got, again = _solve(ctx, v, wins, pick, team)
        teams = [s["team"] for _, s in got]
        print(f"team {team}: team sizes {teams}, fallbacks {again}", flush=True)
        assert again == 0 and all(s["wave_fail"] == 0 for _, s in got)
        if team[0] >= 2:
            # every table of at least 2^20 nodes (all sixteen) in a team of that size
            assert teams == [team[0]] * len(pick)
        else:
            assert teams[0] > 1
        for (t, s), (t0, s0), q in zip(got, lone, pick):
            assert s["error"] & ~1 == 0
            assert s["passes"] == s0["passes"] and s["wave_passes"] == s["passes"]
            assert int((t.view(np.uint32) != t0.view(np.uint32)).sum()) == 0, f"window {q}, team {team}"
