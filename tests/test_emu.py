"""The kernel's bulk-synchronous code, run by host threads (tests/native),
against the oracle: checks the LEVEL/CHANGE schedules, the carried-register
resolution and the radix sort on a machine without a GPU."""
import os
import subprocess

import numpy as np
import pytest

import oracle_bind as ob
from cases import make_case

NATIVE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "native")


@pytest.fixture(scope="module")
def emu():
    subprocess.run(["make", "-s", "-C", NATIVE], check=True)
    import emu_bind
    return emu_bind


@pytest.mark.parametrize("seed", range(60))
def test_emulated_kernel_matches_oracle(emu, seed):
    par, v = make_case(seed)
    s = ob.slowness(v)
    t_ref, st = ob.solve(par, s)
    # thread counts from one lane to more lanes than most rounds have ready nodes
    nthreads = [1, 2, 3, 4, 8, 16, 24][seed % 7]
    # the per-face schedule knob must not change a bit: 0 (CHANGE never gives
    # up), 1 (most reverse-pass faces handed to LEVEL), 8 (default)
    budget = [0, 1, 8][seed % 3]
    t, est = emu.solve(par, s, nthreads, change_budget=budget)
    assert est["error"] == 0
    assert est["passes"] == st["passes"] and est["restarts"] == st["restarts"]
    assert int((t.view(np.uint32) != t_ref.view(np.uint32)).sum()) == 0


@pytest.mark.parametrize("seed", range(24))
def test_global_scratch_schedule_state_matches_oracle(emu, seed):
    """Faces too large for LDS keep the LEVEL pending counters and successor
    masks and the CHANGE queue marks in global scratch (on the device: faces
    above 52,000 nodes on the wide build, 26,800 on the narrow one).  Here
    every face takes that path, on the same cases, bit for bit."""
    par, v = make_case(seed + 300)
    s = ob.slowness(v)
    t_ref, st = ob.solve(par, s)
    nthreads = [1, 3, 8, 24][seed % 4]
    budget = [0, 1, 8][seed % 3]
    t, est = emu.solve(par, s, nthreads, change_budget=budget, global_state=True)
    assert est["error"] == 0
    assert est["passes"] == st["passes"] and est["restarts"] == st["restarts"]
    assert est["faces_level"] > 0
    assert int((t.view(np.uint32) != t_ref.view(np.uint32)).sum()) == 0


def test_schedules_exercised(emu):
    """Across a handful of cases both schedules run, carried registers are
    resolved by deferral, and with the budget switched off CHANGE faces
    converge in few rounds; a budget of 1 hands faces to LEVEL."""
    tot, tot1 = {}, {}
    for seed in range(100, 130):
        par, v = make_case(seed)
        _, est = emu.solve(par, ob.slowness(v), 4, change_budget=0)
        for k, x in est.items():
            tot[k] = tot.get(k, 0) + x
        _, est = emu.solve(par, ob.slowness(v), 4, change_budget=1)
        for k, x in est.items():
            tot1[k] = tot1.get(k, 0) + x
    assert tot["faces_level"] > 0 and tot["faces_change"] > 0
    assert tot["defers"] > 0
    assert tot["rounds_change"] < tot["faces_change"] * 64
    assert tot1["fallbacks"] > tot["fallbacks"]


@pytest.mark.parametrize("seed", range(24))
def test_wavefront_passes_match_oracle(emu, seed):
    """Passes run as one dataflow wavefront over the table (the first pass, or
    every pass), with 1 to 24 host threads standing in for the lanes: the same
    grid bit for bit, the same pass count, and the wavefront actually ran."""
    par, v = make_case(seed + 500)
    s = ob.slowness(v)
    t_ref, st = ob.solve(par, s)
    nthreads = [1, 2, 5, 8][seed % 4]
    wave = [1, 3][seed % 2]
    t, est = emu.solve(par, s, nthreads, change_budget=[0, 1, 8][seed % 3], wave=wave,
                       global_state=bool(seed % 4 == 3))
    assert est["error"] == 0
    assert est["passes"] == st["passes"] and est["restarts"] == st["restarts"]
    assert est["wave_passes"] == (st["passes"] if wave == 3 else 1)
    assert int((t.view(np.uint32) != t_ref.view(np.uint32)).sum()) == 0


def test_wavefront_carried_registers_exercised(emu):
    """Across a handful of cases the wavefront meets nodes whose 2D trials read
    the carried t[0]/t[4] of their sequential predecessor, resolves them by face
    scans, and needs far fewer rounds than the face-by-face schedule has."""
    tot = {}
    for seed in range(700, 708):
        par, v = make_case(seed)
        s = ob.slowness(v)
        _, est = emu.solve(par, s, 8, wave=3)
        _, est0 = emu.solve(par, s, 8, wave=0, change_budget=0)
        for k, x in est.items():
            tot[k] = tot.get(k, 0) + x
        tot["rounds_face"] = tot.get("rounds_face", 0) + est0["rounds_level"] + est0["rounds_change"]
    assert tot["wave_carry"] > 0 and tot["wave_scans"] >= tot["wave_carry"]
    assert tot["wave_seeds"] > 0
    assert tot["rounds_wave"] < tot["rounds_face"]


@pytest.mark.parametrize("seed", range(16))
def test_team_wavefront_passes_match_oracle(emu, seed):
    """Wavefront passes shared by a team of 2-4 workgroups (each of 1-5 host
    threads, meeting at the team barrier the device uses): the same grid bit
    for bit as the oracle, the same passes, every evaluation accounted for."""
    par, v = make_case(seed + 900)
    s = ob.slowness(v)
    t_ref, st = ob.solve(par, s)
    team = [2, 3, 4][seed % 3]
    nthreads = [1, 2, 5][seed % 3 if seed % 2 else (seed + 1) % 3]
    wave = [1, 3][seed % 2]
    t, est = emu.solve(par, s, nthreads, change_budget=[0, 8][seed % 2], wave=wave, team=team,
                       global_state=bool(seed % 5 == 4))
    assert est["error"] == 0
    assert est["team"] == team
    assert est["passes"] == st["passes"] and est["restarts"] == st["restarts"]
    assert est["wave_passes"] == (st["passes"] if wave == 3 else 1)
    assert int((t.view(np.uint32) != t_ref.view(np.uint32)).sum()) == 0


def test_team_carried_registers_and_counts(emu):
    """Teams resolve the carried-register requests their members file and
    report the same seeds, requests and head waves as one workgroup does."""
    tot = {}
    for seed in range(700, 706):
        par, v = make_case(seed)
        s = ob.slowness(v)
        t1, e1 = emu.solve(par, s, 4, wave=3)
        t3, e3 = emu.solve(par, s, 4, wave=3, team=3)
        assert int((t1.view(np.uint32) != t3.view(np.uint32)).sum()) == 0
        for k in ("wave_seeds", "wave_carry", "passes", "restarts", "rounds_wave", "evals") + \
                tuple(f"headw{i}" for i in range(7)):
            assert e1[k] == e3[k], k
        tot["wave_carry"] = tot.get("wave_carry", 0) + e3["wave_carry"]
    assert tot["wave_carry"] > 0


@pytest.mark.parametrize("team", [2, 4])
def test_team_member_that_never_starts(emu, team):
    """A team whose last member never starts: the others give the pass up
    after the team's wait limit (20 ms here), leave their barriers, and the
    leader reports the failed wavefront (error bit 7, reason 7) instead of
    waiting for ever; the host would then solve the table face by face."""
    import time as _time
    par, v = make_case(901)
    s = ob.slowness(v)
    t0 = _time.monotonic()
    _, est = emu.solve(par, s, 2, wave=3, team=team, missing_member=True)
    assert _time.monotonic() - t0 < 60
    assert est["error"] & 128
    assert est["wave_fail"] == 7
