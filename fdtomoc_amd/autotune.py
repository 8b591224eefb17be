"""Autotuner of the sweep's schedules.

Three knobs, none with an effect on the results.  The wavefront flags
(`fdgpu_set_wave`): every pass face by face (flags 0), the first pass as one
dataflow wavefront over the table (flags 1), or every pass as a wavefront
(flags 3); passes run face by face use the per-face schedule below.  Teams
(`fdgpu_set_team`): the largest tables' wavefront passes shared by several
workgroups, planned per table by the library.

Reverse-pass faces first try the CHANGE schedule (evaluate the whole face at
once, then re-evaluate what changed); when changes cascade through a face, the
LEVEL schedule (each node once, in dependency order) is cheaper, and the face
is handed over after `change_budget` evaluations per node.  Which budget is
fastest depends on the workload: many reverse passes with small changes favour
CHANGE (C2-like batches), a few passes with large changes favour LEVEL
(C3-like windows).  The results are bit-identical for every budget, so the
tuner only times: one batched solve per candidate on the caller's tables (or a
sample of them), keeping the fastest.

    from fdtomoc_amd import autotune
    best, timings = autotune.tune_change_budget(ctx, specs)
    (wave, budget), timings = autotune.tune_schedule(ctx, specs)
"""
from __future__ import annotations

import time

from . import fdgpu

CANDIDATES = (0, 2, 4, 8, 16, 32)
# (size, share) of fdgpu_set_team: no teams, and the planner at 150 % load
TEAM_CANDIDATES = ((1, 0), (0, 150))


def tune_change_budget(ctx: "fdgpu.Context", specs: list, candidates=CANDIDATES, sample: int | None = None,
                       objective: str = "wall") -> tuple[int, dict]:
    """Times a batched solve of `specs` (the first `sample` of them if given)
    for each candidate budget and sets the fastest on `ctx`.

    objective "wall": the launch's wall time (what a batch of these tables
    costs); "sum": the sum of the tables' own device times (throughput when
    far more tables than workgroups are queued).  Returns (best budget,
    {budget: seconds})."""
    sel = specs if sample is None else specs[:sample]
    timings = {}
    for b in candidates:
        ctx.set_schedule(b)
        t0 = time.perf_counter()
        ids = ctx.solve(sel, fdgpu.BATCHED)
        wall = time.perf_counter() - t0
        if objective == "sum":
            tick = 1e-8   # device wall-clock ticks of the per-table timers (100 MHz)
            cost = 0.0
            for i in ids:
                s = ctx.stats(i)
                cost += (s["t_gather"] + s["t_sort"] + s["t_level"] + s["t_change"] + s["t_init"] + s["t_wave"]) * tick
        else:
            cost = wall
        for i in ids:
            ctx.release(i)
        timings[b] = cost
    best = min(timings, key=timings.get)
    ctx.set_schedule(best)
    return best, timings


def tune_schedule(ctx: "fdgpu.Context", specs: list, candidates=CANDIDATES, sample: int | None = None,
                  objective: str = "wall") -> tuple[tuple[int, int], dict]:
    """Times the passes face by face under each candidate budget (flags 0), the
    first pass as a wavefront with reverse passes face by face (flags 1), and
    every pass as a wavefront (flags 3); with wavefronts, also the team
    choices of TEAM_CANDIDATES against the context's own.  Keeps the fastest
    on `ctx` and returns ((wave flags, budget), {"w<flags>/b<budget>", "w3" or
    "w<flags>/team<size>/<share>": seconds})."""
    ctx.set_wave(1)
    best_b, t1 = tune_change_budget(ctx, specs, candidates, sample, objective)
    timings = {f"w1/b{b}": v for b, v in t1.items()}
    ctx.set_wave(0)
    best_b0, t0 = tune_change_budget(ctx, specs, candidates, sample, objective)
    timings.update({f"w0/b{b}": v for b, v in t0.items()})
    ctx.set_wave(3)
    _, t3 = tune_change_budget(ctx, specs, (best_b,), sample, objective)
    timings["w3"] = t3[best_b]
    options = {(1, best_b): t1[best_b], (0, best_b0): t0[best_b0], (3, best_b): t3[best_b]}
    wave, budget = min(options, key=options.get)
    ctx.set_wave(wave)
    ctx.set_schedule(budget)
    if wave != 0:
        # teams for the largest tables' wavefront passes: the context's choice
        # (timed above) against none and against a plan at a higher load
        base_team = ctx.get_team()
        best_t, best_s = base_team, options[(wave, budget)]
        for cand in TEAM_CANDIDATES:
            if cand == base_team:
                continue
            ctx.set_team(*cand)
            _, tt = tune_change_budget(ctx, specs, (budget,), sample, objective)
            timings[f"w{wave}/team{cand[0]}/{cand[1]}"] = tt[budget]
            if tt[budget] < best_s:
                best_t, best_s = cand, tt[budget]
        ctx.set_team(*best_t)
        ctx.set_schedule(budget)
    return (wave, budget), timings
