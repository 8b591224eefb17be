"""Benchmark of the MI355X sphfd path.

Default workload (the configuration the north-star target is stated on, C3):
256 station solves whose boxes are clipped to per-station radii uniform in
40-200 km -- windows from 41 x 41 x 41 to 201 x 201 x 101 nodes of one
resident 401 x 401 x 101 model at 1 km centred at 23.5N 121E (FP32 gradient
5.0 -> 8.0 km/s over 0-64 km times three seeded +-8 % Gaussian anomalies of
sigma 10-30 km), reverse=50.  A step is one batched solve of all 256 windows
(one hipGraph-captured persistent launch); the model is resident in HBM
before timing starts, and the per-face schedule is picked by the autotuner
during the untimed warm-up.

The line reports station-solves/s and grid nodes/s of the batched path, the
plain path (one table per launch) on a sample, p50/p99 step latency, the HBM
roofline of the sweep kernel (kernel time from HIP events on its stream), the
analytic FP64 error of the same grid in the unperturbed gradient medium, and
the reference sphfd (compiled from its own sources, oracle/_ref) timed on the
host's cores, one single-threaded process per table, on a size-stratified
sample of the same windows (the largest included), whose outputs are also
compared with the GPU tables bit for bit.

Other workloads (one JSON line each, same keys where they apply):
  --workload c1   one 65x65x33 gradient table: plain GPU latency, the
                  reference on one core, error against the analytic
                  linear-gradient time
  --workload c2   48 stations x (P, S) = 96 tables of 161x161x81 (gradient x
                  +-5 % checkerboard) in one launch, plain sample, reference
                  sample
  --workload c4   serving: a Poisson stream of requests (81^3 .. 257x257x129,
                  each with its own updated velocities), 200 GB HBM arena,
                  p50/p99 latency and stations/s
  --workload c5   the 1000-station P+S set on 321x321x161, all 2000 tables
                  resident through 5 inversion iterations (models updated
                  and tables re-solved in place), sample tables written
  --workload c2f  c2f's coarse-to-fine interpolation (the row beside the
                  path): a 31x31x7 coarse model to the 301x301x61 fine grid
                  into two resident models per step, the reference c2f timed
                  beside it
  --workload c2w  C2 with every table written to a file each step by the
                  background writer, overlapping the next solve; the same
                  steps with synchronous writes beside it
  --workload loc  sphfdloc's grid-search location (the row after c2f): 256
                  events of 24-40 picks over the 96 resident tables of the
                  default workload, all events per step, the oracle's location
                  timed on one core on a sample of the events
  --workload rays sphrayderv's ray tracing (the row after loc): 1000 events x
                  48 stations x (P, S) = 96,000 rays through the 96 resident
                  tables of 161x161x81, derivatives on a 33x33x17 coarse grid,
                  all rays per step; the ray oracle timed on the host's cores
                  on a sample of the rays

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload c3]
    python bench.py --calibrate-copy MIB     # one known-size copy, for HBM counter scaling
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import json
import math
import os
import shutil
import subprocess
import sys
import tempfile
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)

from fdtomoc_amd import analytic, autotune, fdgpu, synth, tables, workloads  # noqa: E402

METRIC = "stations/s & nodes/s batched vs plain at <=1e-4 s err; p50/p99 solve latency"
HBM_PEAK_GBS = 8000.0
BYTES_PER_VISIT = 12          # slowness read + time read + time write per node visit
# FP64 lane-operations of one launch of the C2 workload at change budget 0,
# from VALU instruction counters (64 x (ADD + MUL + 2 FMA + TRANS), an upper
# bound) (profiles/sphfd_sweep_kernel_c2_96tables_fp64.md)
FP64_OPS_C2 = {"change_budget": 0, "ops": 8.557e12, "source": "profiles/sphfd_sweep_kernel_c2_96tables_fp64.md"}
FP64_PEAK_TFLOPS = 78.6
REF_EXE = os.path.join(REPO, "oracle", "_ref", "sphfd_ref")
# HBM bytes of one c2f model update (its three kernels, FETCH_SIZE x 1.99992 +
# WRITE_SIZE) (profiles/c2f_fine_kernel_301x301x61.md)
TRAFFIC_BYTES_C2F = 51.39e6
# HBM bytes per node-search launch of the location workload (average over
# its 5 passes, FETCH_SIZE x 1.99992 + WRITE_SIZE, copy-calibrated)
# (profiles/loc_nodes_kernel_256events_96tables.md)
TRAFFIC_BYTES_LOC = 2.6258e9


# ------------------------------------------------------------------ helpers

def host_info() -> dict:
    """The host's CPUs as this process sees them: nproc, the CPUs it may run
    on, the cgroup CPU quota and the CPU model."""
    info = {"nproc": os.cpu_count(), "affinity": len(os.sched_getaffinity(0)), "cgroup_cpus": None,
            "model": None}
    try:
        q, per = open("/sys/fs/cgroup/cpu.max").read().split()[:2]
        if q != "max":
            info["cgroup_cpus"] = int(q) / int(per)
    except (OSError, ValueError):
        pass
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                info["model"] = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return info


def host_cores(requested: int) -> int:
    """All the cores this process may use (CPU affinity, capped by the cgroup
    quota) unless --cpu-threads says otherwise."""
    if requested:
        return requested
    info = host_info()
    n = info["affinity"]
    if info["cgroup_cpus"]:
        n = min(n, max(1, int(info["cgroup_cpus"])))
    return n


def cpu_reference(items, cores: int):
    """Runs the compiled reference sphfd once per table, `cores` processes at
    a time (one process per station table).  `items` = [(par dict, velocity
    (nz, ny, nx) float32, header or None)].  Returns ([seconds per table],
    [times grid per table])."""
    base = tempfile.mkdtemp(prefix="sphfd_cpu_")

    def one(k):
        par, vel, head = items[k]
        wd = os.path.join(base, f"t{k}")
        os.makedirs(os.path.join(wd, "out"))
        nz, ny, nx = vel.shape
        if head is None:
            head = tables.make_header(nx, ny, nz, par["x0"], par["y0"], par["z0"], 0.0, 0.0, par["h"])
        tables.write_grid(os.path.join(wd, "vel.mod"), head, vel)
        txt = "".join(f"{k2}={v!r}\n" for k2, v in par.items()) + "timefile=sta.times\nvelfile=vel.mod\n"
        open(os.path.join(wd, "sta.par"), "w").write(txt)
        open(os.path.join(wd, "parlist"), "w").write("par=sta.par\n\n")
        open(os.path.join(wd, "spec"), "w").write("parlist parlist\ntimedir out\n")
        env = dict(os.environ, OMP_NUM_THREADS="1")
        t0 = time.perf_counter()
        r = subprocess.run([REF_EXE], input="spec\n", cwd=wd, env=env, capture_output=True, text=True)
        dt = time.perf_counter() - t0
        if r.returncode != 0:
            raise RuntimeError("reference sphfd failed: " + r.stderr[-1000:])
        return dt, np.array(tables.read_grid(os.path.join(wd, "out", "sta.times")).data)

    try:
        with cf.ThreadPoolExecutor(max_workers=cores) as ex:
            res = list(ex.map(one, range(len(items))))
    finally:
        shutil.rmtree(base, ignore_errors=True)
    return [r[0] for r in res], [r[1] for r in res]


def compare(gpu: list, ref: list):
    maxerr, ndiff = 0.0, 0
    for a, b in zip(gpu, ref):
        maxerr = max(maxerr, float(np.max(np.abs(a.astype(np.float64) - b))))
        ndiff += int((a.view(np.uint32) != b.view(np.uint32)).sum())
    return maxerr, ndiff


def cpu_block(secs, sizes, cores, what):
    """cpu_baseline object: throughput of `cores` concurrent single-threaded
    reference processes = tables / (sum of per-table seconds / cores)."""
    busy = float(np.sum(secs)) / cores
    return {"value": round(len(secs) / busy, 5), "unit": "station-solves/s", "cores": cores, "kind": "reference",
            "nodes_per_s": round(float(np.sum(sizes)) / busy, 1), "host": host_info(),
            "sample": what + f"; compiled reference sphfd, one single-threaded process per table, {cores} at a "
                             f"time; wall time per table incl. file I/O, rate = tables / (sum of times / {cores})"}


def base_line(args, value, unit, ms_per_step, config, dtype="f64"):
    return {"metric": METRIC, "value": value, "unit": unit, "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": ms_per_step, "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": dtype,
            "data": "synthetic (seeded models and stations, reference binary layout)", "config": config}


# HBM bytes per launch (FETCH_SIZE x calibration + WRITE_SIZE x calibration)
# counted by rocprofv3 for one launch of a workload at the schedule (wavefront
# flags, CHANGE budget) it was profiled with; attached to a bench line only
# when the run uses that schedule.
PROFILED_TRAFFIC = {
    # FETCH_SIZE x 1.99992 + WRITE_SIZE x 1.0
    "c2": [{"wave": 0, "change_budget": 0, "bytes": 323.92e9,
            "source": "profiles/sphfd_sweep_kernel_c2_96tables_hbm_traffic.md"}],
    "c3": [{"wave": 0, "change_budget": 2, "bytes": 398.23e9,
            "source": "profiles/sphfd_sweep_kernel_c3_256windows.md"},
           {"wave": 3, "change_budget": 2, "bytes": 1484.32e9,
            "source": "profiles/sphfd_sweep_kernel_c3_256windows_wavefront.md"},
           {"wave": 3, "change_budget": 2, "team": [0, 150], "bytes": 1717.57e9,
            "source": "profiles/sphfd_sweep_kernel_c3_256windows_teams.md"}],
}


def profiled_traffic(workload: str, sched: dict):
    """The profiled HBM traffic of the schedule a run picked: wavefront flags,
    CHANGE budget and, with wavefronts, the team choice (none: [1, x])."""
    team = list(sched.get("team") or [1, 0])
    for p in PROFILED_TRAFFIC.get(workload, []):
        if p["wave"] != sched.get("wave") or p["change_budget"] != sched.get("change_budget"):
            continue
        if p["wave"] == 0 or list(p.get("team", [1, 0])) == team or (team[0] == 1 and p.get("team", [1])[0] == 1):
            return p["bytes"], p["source"]
    return None, None


def roofline(visits, kern_s, traffic=None, source=None):
    achieved = BYTES_PER_VISIT * visits / kern_s / 1e9
    out = {"bound": "hbm", "achieved": round(achieved, 3), "peak": HBM_PEAK_GBS, "unit": "GB/s",
           "frac": round(achieved / HBM_PEAK_GBS, 6), "traffic": traffic,
           "kernel": "sphfd_sweep_kernel", "kernel_ms": round(kern_s * 1e3, 2),
           "algorithmic_bytes": BYTES_PER_VISIT * visits}
    if source:
        out["traffic_source"] = source
    return out


def tune(ctx, specs, args, sample=None):
    """Per-face schedule picked by the autotuner before timing (untimed, like
    the warmup); the tables are identical for every choice."""
    if args.team >= 0 or args.team_share >= 0:
        size, share = ctx.get_team()
        ctx.set_team(args.team if args.team >= 0 else size, args.team_share if args.team_share >= 0 else share)
    if args.wave >= 0:
        ctx.set_wave(args.wave)
    if args.schedule >= 0:
        ctx.set_schedule(args.schedule)
        return {"wave": ctx.get_wave(), "change_budget": args.schedule, "team": list(ctx.get_team()), "tuned": False}
    if args.no_autotune:
        return {"wave": ctx.get_wave(), "change_budget": ctx.get_schedule(), "team": list(ctx.get_team()),
                "tuned": False}
    if args.wave >= 0:
        best, timings = autotune.tune_change_budget(ctx, specs, sample=sample)
        cand = {str(k): round(v, 3) for k, v in timings.items()}
    else:
        (_, best), timings = autotune.tune_schedule(ctx, specs, sample=sample)
        cand = {k: round(v, 3) for k, v in timings.items()}
    print(f"# autotune wave {ctx.get_wave()} change_budget {best}: {cand}", file=sys.stderr, flush=True)
    return {"wave": ctx.get_wave(), "change_budget": best, "team": list(ctx.get_team()), "tuned": True,
            "candidates_s": cand}


def timed_steps(ctx, specs, args, mode=fdgpu.BATCHED):
    for _ in range(args.warmup):
        for i in ctx.solve(specs, mode):
            ctx.release(i)
    wall, kern, last = [], [], None
    for s in range(args.steps):
        t0 = time.perf_counter()
        ids = ctx.solve(specs, mode)
        dt = time.perf_counter() - t0
        wall.append(dt)
        kern.append(ctx.last_kernel_ms() / 1e3)
        print(f"# step {s}: {dt * 1e3:.1f} ms wall, {kern[-1] * 1e3:.1f} ms kernel", file=sys.stderr, flush=True)
        if s < args.steps - 1:
            for i in ids:
                ctx.release(i)
        else:
            last = ids
    return np.array(wall), np.array(kern), last


def plain_rate(ctx, specs, sample):
    if not sample:
        return None
    t0 = time.perf_counter()
    ids = ctx.solve([specs[i] for i in sample], fdgpu.PLAIN)
    dt = time.perf_counter() - t0
    for i in ids:
        ctx.release(i)
    return len(sample) / dt


def analytic_error(ctx, par: dict, what: str, sample: int = 20000) -> dict:
    """FD error of the workload's grid: one table of the same grid and source
    solved on the device (untimed) in the unperturbed gradient medium
    Vp = 5.0 + 3.0 z/64 km/s (z the depth k h below the grid top), against
    the analytic FP64 spherical ray-integral time (fdtomoc_amd/analytic.py) on
    a seeded sample of nodes.  The GPU tables are bit-identical to the
    reference's, so this is the reference's own discretisation error."""
    nz, ny, nx = par["nz"], par["ny"], par["nx"]
    depth = np.arange(nz, dtype=np.float64) * par["h"]
    v = np.ascontiguousarray(np.broadcast_to((5.0 + 3.0 * depth / 64.0)[:, None, None], (nz, ny, nx)),
                             dtype=np.float32)
    m = ctx.model(v)
    (tid,) = ctx.solve([fdgpu.TableSpec(par, m)], fdgpu.BATCHED)
    t = ctx.read(tid, v.shape)
    ctx.release(tid)
    ctx.release_model(m)
    geom = analytic.geometry(par)
    e = analytic.table_error(t, geom, lambda r, q, f: analytic.gradient_time(geom, 5.0, 3.0 / 64.0, geom["z0r"],
                                                                             r, q, f), sample=sample)
    e.update({"max": round(e["max"], 5), "mean": round(e["mean"], 5), "unit": "s",
              "what": what + ", Vp 5.0 + 3.0 z/64 km/s, vs the FP64 spherical linear-gradient ray-integral time"})
    return e


# ------------------------------------------------------------------ C2 (default)

def run_c2(args):
    g, vp, vs, stations = workloads.c2()
    ctx = fdgpu.Context(0)
    mp, ms = ctx.model(vp), ctx.model(vs)
    specs, items = [], []
    for st in stations:
        for m, v in ((mp, vp), (ms, vs)):
            p = workloads.table_par(g, st)
            specs.append(fdgpu.TableSpec(p, m))
            items.append((p, v, None))
    ntab = len(specs)
    sched = tune(ctx, specs, args)
    wall, kern, last = timed_steps(ctx, specs, args)
    restarts = np.array([ctx.stats(i)["restarts"] for i in last])
    visits = int(np.sum(g.n * (1 + restarts)))
    stations_s = ntab / float(wall.mean())
    plain_s = plain_rate(ctx, specs, list(range(min(args.plain_sample, ntab))))
    out = base_line(args, round(stations_s, 3), "station-solves/s", round(float(wall.mean()) * 1e3, 2),
                    {"workload": "C2: 48 stations P+S = 96 tables, 161x161x81 at 1 km, gradient 5-8 km/s x "
                                 "+-5% 20 km checkerboard, reverse=50, one batched launch per step",
                     "tables": ntab, "nodes_per_table": g.n, "reverse_passes_mean": float(restarts.mean())})
    out["nodes_per_s"] = round(ntab * g.n / float(wall.mean()), 1)
    out["schedule"] = sched
    out["plain_station_solves_per_s"] = None if plain_s is None else round(plain_s, 4)
    out["batched_over_plain"] = None if plain_s is None else round(stations_s / plain_s, 2)
    out["latency_ms"] = {"p50": round(float(np.percentile(wall, 50)) * 1e3, 2),
                         "p99": round(float(np.percentile(wall, 99)) * 1e3, 2),
                         "what": "batched solve of all 96 tables (submit to results resident)"}
    out["roofline"] = roofline(visits, float(kern.mean()), *profiled_traffic("c2", sched))
    if sched.get("change_budget") == FP64_OPS_C2["change_budget"]:
        fp64 = FP64_OPS_C2["ops"] / float(kern.mean()) / 1e12
        out["fp64_rate"] = {"achieved_max": round(fp64, 3), "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                            "frac": round(fp64 / FP64_PEAK_TFLOPS, 5), "ops_per_launch": FP64_OPS_C2["ops"],
                            "source": FP64_OPS_C2["source"]}
    if not args.no_analytic:
        out["analytic_error"] = analytic_error(ctx, specs[0].par, "the first station's 161x161x81 grid")
    out["cpu_baseline"] = None
    if not args.no_cpu and os.path.exists(REF_EXE):
        cores = host_cores(args.cpu_threads)
        n = max(1, min(cores if args.cpu_sample < 0 else args.cpu_sample, ntab))
        secs, refs = cpu_reference(items[:n], cores)
        out["cpu_baseline"] = cpu_block(secs, [g.n] * n, cores, f"the first {n} of the 96 tables")
        out["gpu_over_cpu"] = round(stations_s / out["cpu_baseline"]["value"], 2)
        err, nd = compare([ctx.read(last[i], refs[i].shape) for i in range(n)], refs)
        out["max_abs_err_vs_reference_s"] = err
        out["nodes_differing_vs_reference"] = nd
    for i in last:
        ctx.release(i)
    ctx.close()
    return out


# ------------------------------------------------------------------ C2 with files

def run_c2w(args):
    """C2 with every table written to a file in the reference layout each step:
    the step solves the 96 tables and hands them to the background writer
    (fdgpu_tables_write_async), whose device-to-host copies and file writes
    overlap the next step's solve; a step ends once the previous step's files
    are written.  Beside it: the same steps with the files written after each
    solve (fdgpu_table_write, one table after another)."""
    g, vp, vs, stations = workloads.c2()
    ctx = fdgpu.Context(0)
    mp, ms = ctx.model(vp), ctx.model(vs)
    specs = [fdgpu.TableSpec(workloads.table_par(g, st), m) for st in stations for m in (mp, ms)]
    ntab = len(specs)
    sched = tune(ctx, specs, args)
    outdir = tempfile.mkdtemp(prefix="sphfd_tables_")
    paths = [os.path.join(outdir, f"t{k:03d}.times") for k in range(ntab)]
    try:
        def step_overlapped(prev):
            ids = ctx.solve(specs, fdgpu.BATCHED)
            if prev:
                ctx.write_wait()
                for i in prev:
                    ctx.release(i)
            ctx.write_async(ids, paths)
            return ids

        prev = None
        for _ in range(args.warmup):
            prev = step_overlapped(prev)
        walls = []
        for s in range(args.steps):
            t0 = time.perf_counter()
            prev = step_overlapped(prev)
            walls.append(time.perf_counter() - t0)
            print(f"# overlapped step {s}: {walls[-1]:.3f} s", file=sys.stderr, flush=True)
        t0 = time.perf_counter()
        ctx.write_wait()
        drain = time.perf_counter() - t0
        for i in prev:
            ctx.release(i)
        seq = []
        for s in range(max(1, min(args.steps, 2))):
            t0 = time.perf_counter()
            ids = ctx.solve(specs, fdgpu.BATCHED)
            for i, p in zip(ids, paths):
                ctx.write(i, p, None)
            seq.append(time.perf_counter() - t0)
            for i in ids:
                ctx.release(i)
            print(f"# sequential step {s}: {seq[-1]:.3f} s", file=sys.stderr, flush=True)
        file_bytes = sum(os.path.getsize(p) for p in paths)
    finally:
        shutil.rmtree(outdir, ignore_errors=True)
    wall = float(np.mean(walls))
    out = base_line(args, round(ntab / wall, 3), "station-solves/s (tables written to files)", round(wall * 1e3, 2),
                    {"workload": "C2 with files: 96 tables 161x161x81 solved per step and written in the "
                                 "reference layout by the background writer while the next step solves",
                     "tables": ntab, "nodes_per_table": g.n})
    out["metric"] = METRIC + " | with table files"
    out["schedule"] = sched
    out["sequential_station_solves_per_s"] = round(ntab / float(np.mean(seq)), 3)
    out["bytes_written_per_step"] = file_bytes
    out["final_drain_s"] = round(drain, 3)
    out["cpu_baseline"] = None
    out["roofline"] = None
    ctx.release_model(mp)
    ctx.release_model(ms)
    ctx.close()
    return out


# ------------------------------------------------------------------ C1

def run_c1(args):
    g, v, st = workloads.c1()
    par = workloads.table_par(g, st)
    ctx = fdgpu.Context(0)
    m = ctx.model(v)
    specs = [fdgpu.TableSpec(par, m)]
    wall, kern, last = timed_steps(ctx, specs, args, fdgpu.PLAIN)
    t = ctx.read(last[0], v.shape)
    geom = analytic.geometry(par)
    err = analytic.table_error(t, geom, lambda r, q, f: analytic.gradient_time(geom, 5.0, 3.0 / 64.0, geom["z0r"],
                                                                               r, q, f))
    out = base_line(args, round(1.0 / float(wall.mean()), 3), "station-solves/s",
                    round(float(wall.mean()) * 1e3, 2),
                    {"workload": "C1: one station, 65x65x33 at 1 km, Vp 5.0 + 3.0 z/64 km/s, plain launch",
                     "tables": 1, "nodes_per_table": g.n})
    out["latency_ms"] = {"p50": round(float(np.percentile(wall, 50)) * 1e3, 2),
                         "p99": round(float(np.percentile(wall, 99)) * 1e3, 2), "what": "one plain solve"}
    st_ = ctx.stats(last[0])
    out["roofline"] = roofline(g.n * (1 + st_["restarts"]), float(kern.mean()))
    err.update({"unit": "s", "what": "every node of the C1 table vs the FP64 spherical linear-gradient "
                                     "ray-integral time (the workload's own medium)"})
    out["analytic_error"] = err
    out["cpu_baseline"] = None
    if not args.no_cpu and os.path.exists(REF_EXE):
        secs, refs = cpu_reference([(par, v, None)], 1)
        out["cpu_baseline"] = cpu_block(secs, [g.n], 1, "the C1 table")
        e, nd = compare([t], refs)
        out["max_abs_err_vs_reference_s"] = e
        out["nodes_differing_vs_reference"] = nd
    ctx.release(last[0])
    ctx.close()
    return out


# ------------------------------------------------------------------ C3

def run_c3(args):
    G, v, wins = workloads.c3()
    ctx = fdgpu.Context(0)
    m = ctx.model(v)
    specs = [fdgpu.TableSpec(w.par, m, w.origin) for w in wins]
    sizes = np.array([w.par["nx"] * w.par["ny"] * w.par["nz"] for w in wins])
    sched = tune(ctx, specs, args)
    wall, kern, last = timed_steps(ctx, specs, args)
    restarts = np.array([ctx.stats(i)["restarts"] for i in last])
    team_sizes = np.array([ctx.stats(i)["team"] for i in last])
    visits = int(np.sum(sizes * (1 + restarts)))
    stations_s = len(specs) / float(wall.mean())
    rng = np.random.default_rng(33)
    psel = sorted(rng.choice(len(specs), size=min(args.plain_sample, len(specs)), replace=False).tolist())
    plain_s = plain_rate(ctx, specs, psel)
    # plain rate scaled to the workload's mean table size
    plain_adj = None if plain_s is None else plain_s * float(np.mean(sizes[psel])) / float(sizes.mean())
    out = base_line(args, round(stations_s, 3), "station-solves/s", round(float(wall.mean()) * 1e3, 2),
                    {"workload": "C3: 256 station windows of a 401x401x101 model, radii U(40,200) km, extents "
                                 "41^3..201x201x101, gradient x 3 Gaussian +-8% anomalies, one batched launch",
                     "tables": len(specs), "nodes_total": int(sizes.sum()),
                     "reverse_passes_mean": float(restarts.mean())})
    out["nodes_per_s"] = round(float(sizes.sum()) / float(wall.mean()), 1)
    out["schedule"] = sched
    out["team_tables"] = int((team_sizes > 1).sum())
    out["plain_station_solves_per_s"] = None if plain_adj is None else round(plain_adj, 4)
    out["batched_over_plain"] = None if plain_adj is None else round(stations_s / plain_adj, 2)
    out["plain_sample"] = f"{len(psel)} random tables, one launch each, rate scaled by mean size"
    out["latency_ms"] = {"p50": round(float(np.percentile(wall, 50)) * 1e3, 2),
                         "p99": round(float(np.percentile(wall, 99)) * 1e3, 2), "what": "batched solve of 256 windows"}
    out["roofline"] = roofline(visits, float(kern.mean()), *profiled_traffic("c3", sched))
    big = wins[int(np.argmax(sizes))]
    if not args.no_analytic:
        out["analytic_error"] = analytic_error(ctx, big.par, "the largest window's grid and station "
                                                             f"({big.par['nx']}x{big.par['ny']}x{big.par['nz']})")
    out["cpu_baseline"] = None
    if not args.no_cpu and os.path.exists(REF_EXE):
        cores = host_cores(args.cpu_threads)
        n = max(2, cores if args.cpu_sample < 0 else args.cpu_sample)
        # a sample stratified over the whole size range, smallest and largest
        # window included: its mean size is the workload's, so its rate in
        # stations/s stands for the whole batch
        order = np.argsort(sizes, kind="stable")
        pick = [int(order[int(round(q))]) for q in np.linspace(0, len(order) - 1, n)]
        items = []
        for k in pick:
            w = wins[k]
            ox, oy, oz = w.origin
            p = w.par
            items.append((p, np.ascontiguousarray(v[oz:oz + p["nz"], oy:oy + p["ny"], ox:ox + p["nx"]]), None))
        secs, refs = cpu_reference(items, cores)
        cb = cpu_block(secs, sizes[pick], cores, f"{n} of the 256 windows at evenly spaced size ranks, the "
                                                 f"smallest and the largest included (sample mean "
                                                 f"{float(np.mean(sizes[pick])):.0f} nodes, workload mean "
                                                 f"{float(sizes.mean()):.0f})")
        out["cpu_baseline"] = cb
        out["gpu_over_cpu"] = round(stations_s / cb["value"], 2)
        out["gpu_over_cpu_nodes"] = round(out["nodes_per_s"] / cb["nodes_per_s"], 2)
        err, nd = compare([ctx.read(last[k], refs[q].shape) for q, k in enumerate(pick)], refs)
        out["max_abs_err_vs_reference_s"] = err
        out["nodes_differing_vs_reference"] = nd
    for i in last:
        ctx.release(i)
    ctx.close()
    return out


# ------------------------------------------------------------------ C4

def run_c4(args):
    reqs = workloads.c4_requests(args.requests, args.rate)
    vels = [workloads.c4_velocity(r) for r in reqs]
    ctx = fdgpu.Context(0)
    ctx.serve_start((257, 257, 129), arena_bytes=int(200e9), staging_bytes=8 << 30)
    keep = {0, 1, 2}            # requests re-solved alone afterwards
    lat, done_at, sizes, restarts, kept = [], [], [], [], {}
    outstanding = {}
    busy_retries = 0
    i = 0
    t0 = time.perf_counter()
    last_print = 0.0
    while i < len(reqs) or outstanding:
        now = time.perf_counter() - t0
        while i < len(reqs) and reqs[i].t_arrival <= now:
            try:
                rid = ctx.serve_submit(fdgpu.TableSpec(reqs[i].par, -1), vels[i])
            except fdgpu.FdgpuError as e:
                if e.code != fdgpu.EBUSY:
                    raise
                busy_retries += 1
                break
            outstanding[rid] = i
            i += 1
        for rid in list(outstanding):
            if ctx.serve_poll(rid):
                k = outstanding.pop(rid)
                t = time.perf_counter() - t0
                lat.append(t - reqs[k].t_arrival)
                done_at.append(t)
                st = ctx.serve_stats(rid)
                sizes.append(int(np.prod(reqs[k].shape)))
                restarts.append(st["restarts"])
                if k in keep:
                    kept[k] = ctx.serve_read(rid, reqs[k].shape)
                ctx.serve_release(rid)
        if now - last_print > 10.0:
            print(f"# t={now:.0f}s submitted {i} done {len(done_at)}", file=sys.stderr, flush=True)
            last_print = now
        time.sleep(0.0005)
    span = max(done_at) - reqs[0].t_arrival
    info = ctx.serve_info()
    ctx.serve_stop()
    # the same requests alone: one batched launch each, compared bit for bit
    nd = 0
    for k, t in kept.items():
        mm = ctx.model(vels[k])
        (tid,) = ctx.solve([fdgpu.TableSpec(reqs[k].par, mm)])
        nd += int((ctx.read(tid, t.shape).view(np.uint32) != t.view(np.uint32)).sum())
        ctx.release(tid)
        ctx.release_model(mm)
    ctx.close()
    lat = np.array(lat)
    out = base_line(args, round(len(reqs) / span, 3), "station-solves/s", None,
                    {"workload": f"C4: Poisson stream of {len(reqs)} requests at {args.rate}/s, sizes 81^3.."
                                 f"257x257x129, each with its own +-3% Gaussian-blob velocity update, 200 GB HBM "
                                 f"arena, continuous batching", "requests": len(reqs), "rate_per_s": args.rate})
    out["steps"] = None
    out["warmup"] = None
    out["nodes_per_s"] = round(float(np.sum(sizes)) / span, 1)
    out["latency_ms"] = {"p50": round(float(np.percentile(lat, 50)) * 1e3, 1),
                         "p99": round(float(np.percentile(lat, 99)) * 1e3, 1),
                         "what": "arrival to result resident (queueing + admission + solve)"}
    out["serve"] = {"server_launches": info["launches"], "busy_retries": busy_retries,
                    "reverse_passes_mean": float(np.mean(restarts)),
                    "mix_vs_alone_nodes_differing": nd, "mix_vs_alone_checked": len(kept)}
    out["cpu_baseline"] = None
    return out


# ------------------------------------------------------------------ C5

def run_c5(args):
    """C5: 1000 stations x (P, S) = 2000 tables of 321 x 321 x 161 (66.4 MB
    each, 132.7 GB), all kept resident in HBM through `--iterations`
    simulated inversion iterations.  Iteration 0 solves every table
    (fdgpu_solve); each later iteration replaces both velocity models in place
    (a fresh random-layer perturbation of the gradient, as an inversion step
    would) and re-solves every table in place (fdgpu_resolve: no allocation).
    A step is one re-solve iteration (model update + all tables), in launches
    of 500 tables.  Each iteration writes 8 tables in the reference layout
    through the background writer (overlapping the next iteration) and checks
    one table against the reference compiled from its sources, run
    concurrently on the host (the CPU baseline)."""
    g, vp, vs, sts = workloads.c5()
    ntab = min(args.tables, 2 * len(sts))
    ctx = fdgpu.Context(0)
    mp, ms = ctx.model(vp), ctx.model(vs)
    specs = []
    for st in sts:
        for m in (mp, ms):
            specs.append(fdgpu.TableSpec(workloads.table_par(g, st), m))
    specs = specs[:ntab]
    chunk = 500
    outdir = tempfile.mkdtemp(prefix="c5_tables_")
    cores = host_cores(args.cpu_threads)
    ref_pool = cf.ThreadPoolExecutor(max_workers=max(1, min(args.iterations, cores)))
    ref_jobs, iters, ids = [], [], []
    pending = set()   # tables whose background writes may still run
    t_start = time.perf_counter()
    try:
        for it in range(args.iterations):
            t0 = time.perf_counter()
            if it > 0:
                rng = np.random.default_rng(500 + it)
                base = synth.gradient_velocity(g, 5.0, 8.0, 64.0)
                vp = np.ascontiguousarray(base * synth.layered_perturbation(g, rng, 0.04, 2.0, 12.0), dtype=np.float32)
                vs = (vp / np.float32(1.73)).astype(np.float32)
                t0 = time.perf_counter()
                ctx.update_model(mp, vp)
                ctx.update_model(ms, vs)
            kern = 0.0
            for c0 in range(0, ntab, chunk):
                if it == 0:
                    try:
                        ids += ctx.solve(specs[c0:c0 + chunk], fdgpu.BATCHED)
                    except fdgpu.FdgpuError as e:
                        if e.ids:
                            bad = [(q, ctx.stats(q)) for q in e.ids if ctx.stats(q)["error"] & ~1]
                            print(f"# {len(bad)} tables with error bits: " + "; ".join(
                                f"table {q}: error {st['error']} passes {st['passes']} wave {st['wave_passes']} "
                                f"level {st['faces_level']} change {st['faces_change']}" for q, st in bad[:8]),
                                file=sys.stderr, flush=True)
                        raise
                else:
                    if any(c0 <= k < c0 + chunk for k in pending):
                        ctx.write_wait()   # the files of these tables are written first
                        pending.clear()
                    ctx.resolve(ids[c0:c0 + chunk], fdgpu.BATCHED)
                kern += ctx.last_kernel_ms() / 1e3
                print(f"# iteration {it}: tables {c0}..{min(c0 + chunk, ntab) - 1} done at "
                      f"{time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)
            dt = time.perf_counter() - t0
            restarts = np.array([ctx.stats(i)["restarts"] for i in ids])
            iters.append({"iteration": it, "s": round(dt, 2), "kernel_s": round(kern, 2),
                          "reverse_passes_mean": float(restarts.mean()), "call": "solve" if it == 0 else "resolve"})
            print(f"# iteration {it}: {dt:.1f} s wall, {kern:.1f} s kernels, reverse passes mean "
                  f"{restarts.mean():.2f}", file=sys.stderr, flush=True)
            # one table against the reference on the host, concurrently
            q = (it * 401) % ntab
            if not args.no_cpu and os.path.exists(REF_EXE):
                gpu_t = ctx.read(ids[q], (g.nz, g.ny, g.nx))
                vel = vp if q % 2 == 0 else vs
                ref_jobs.append((it, q, gpu_t, ref_pool.submit(cpu_reference, [(specs[q].par, vel.copy(), None)], 1)))
            # a sample of the last launch's tables to files in the reference
            # layout, in the background while the next iteration's first
            # launches run
            ctx.write_wait()
            for f in os.listdir(outdir):
                os.remove(os.path.join(outdir, f))
            lo = (ntab - 1) // chunk * chunk
            sample = sorted({lo + (it * 8 + k) * 37 % (ntab - lo) for k in range(8)})
            ctx.write_async([ids[k] for k in sample], [os.path.join(outdir, f"t{k:04d}.times") for k in sample])
            pending = set(sample)
        ctx.write_wait()
        files_ok = all(os.path.getsize(os.path.join(outdir, f)) == 232 + 4 * g.n for f in os.listdir(outdir))
        total = time.perf_counter() - t_start
        resolve = [x["s"] for x in iters if x["call"] == "resolve"]
        step = float(np.mean(resolve)) if resolve else iters[0]["s"]
        out = base_line(args, round(ntab / step, 3), "station-solves/s", round(step * 1e3, 1),
                        {"workload": f"C5: {ntab} tables (1000 stations x P, S) of 321x321x161, gradient x random "
                                     f"+-4% layers, all resident through {args.iterations} inversion iterations "
                                     f"(model update in place + re-solve of every table in place per step)",
                         "tables": ntab, "nodes_per_table": g.n, "iterations": args.iterations})
        out["steps"] = len(resolve)
        out["warmup"] = 1
        out["nodes_per_s"] = round(ntab * g.n / step, 1)
        out["iterations"] = iters
        out["run_s"] = round(total, 1)
        out["resident_table_bytes"] = 4 * g.n * ntab
        out["tables_solved_again"] = ctx.resolved_again()
        out["sample_files_written_ok"] = bool(files_ok)
        last = iters[-1]
        out["roofline"] = roofline(int(ntab * g.n * (1 + last["reverse_passes_mean"])), last["kernel_s"])
        out["cpu_baseline"] = None
        if ref_jobs:
            secs, nd, err = [], 0, 0.0
            for it, q, gpu_t, fut in ref_jobs:
                s1, refs = fut.result()
                secs += s1
                e, d = compare([gpu_t], refs)
                nd += d
                err = max(err, e)
            busy = float(np.mean(secs))
            out["cpu_baseline"] = {"value": round(cores / busy, 5), "unit": "station-solves/s", "cores": cores,
                                   "kind": "reference", "nodes_per_s": round(cores * g.n / busy, 1),
                                   "host": host_info(),
                                   "sample": f"one table per iteration ({len(secs)} tables), compiled reference sphfd, "
                                             f"one single-threaded process each, run concurrently with the GPU "
                                             f"iterations; rate = {cores} cores / mean wall time per table "
                                             f"({busy:.1f} s incl. file I/O)"}
            out["gpu_over_cpu"] = round(out["value"] / out["cpu_baseline"]["value"], 2)
            out["max_abs_err_vs_reference_s"] = err
            out["nodes_differing_vs_reference"] = nd
    finally:
        ref_pool.shutdown(wait=True)
        shutil.rmtree(outdir, ignore_errors=True)
    for i in ids:
        ctx.release(i)
    ctx.close()
    return out


# ------------------------------------------------------------------ c2f

C2F_EXE = os.path.join(REPO, "oracle", "_ref", "c2f_ref")


def c2f_reference(co, fine_shape):
    """The reference c2f compiled from its own sources, one process on one core,
    on the same coarse model: (wall seconds incl. file I/O, fine velocities)."""
    wd = tempfile.mkdtemp(prefix="c2f_cpu_")
    try:
        with open(os.path.join(wd, "spec"), "w") as f:
            f.write(workloads.c2f_spec_text(co.nxc, co.nyc, co.nzc, co.igridx, co.igridy, co.igridz, co.x0,
                                            co.y0, co.z0, co.h))
        head = tables.make_header(co.nxc, co.nyc, co.nzc, co.x0, co.y0, co.z0, co.h, co.h, co.h,
                                  text=tables.header_text("CORS", "CART", "BMOD"))
        with open(os.path.join(wd, "coarse.mod"), "wb") as f:
            f.write(np.asarray(head, dtype=tables.VHEAD_DTYPE).tobytes())
            for a in (co.igridx, co.igridy, co.igridz):
                f.write(np.asarray(a, dtype="<i4").tobytes())
            f.write(np.asarray(co.velocity, dtype="<f4").tobytes())
        t0 = time.perf_counter()
        r = subprocess.run([C2F_EXE], input="spec\n", cwd=wd, capture_output=True, text=True, timeout=600)
        wall = time.perf_counter() - t0
        if r.returncode != 0:
            raise RuntimeError(f"reference c2f failed: {r.stdout[-500:]}")
        n = int(np.prod(fine_shape))
        raw = np.fromfile(os.path.join(wd, "fine.mod"), dtype=np.uint8)
        return wall, raw[232:232 + 8 * n].view("<f4").reshape((2,) + tuple(fine_shape)).copy()
    finally:
        shutil.rmtree(wd, ignore_errors=True)


def run_c2f(args):
    """c2f's coarse-to-fine interpolation (src/c2f.c:549-575, 673-727) on the
    device: a 31 x 31 x 7 coarse model (P and S) to the 301 x 301 x 61 fine grid,
    straight into two resident models; a step is one inversion iteration's
    model update (coarse model up, interpolation, range, models replaced)."""
    co = fdgpu.Coarse(*workloads.c2f_coarse())
    nx, ny, nz = co.dims()
    n = nx * ny * nz
    ctx = fdgpu.Context(0)
    mp, ms, fine = ctx.models_from_coarse(co, want_fine=True)
    # the coarse model as its C record, built once (an inversion loop hands the
    # library its own record); the timed region is the update call
    pc = fdgpu.PreparedCoarse(co)
    for _ in range(args.warmup):
        ctx.update_models_from_prepared(mp, ms, pc)
    walls, kerns = [], []
    for s in range(args.steps):
        t0 = time.perf_counter()
        ctx.update_models_from_prepared(mp, ms, pc)
        walls.append(time.perf_counter() - t0)
        kerns.append(ctx.last_kernel_ms())
        print(f"# step {s}: {walls[-1] * 1e3:.3f} ms wall, {kerns[-1]:.3f} ms kernel", file=sys.stderr, flush=True)
    wall = float(np.mean(walls))
    kern_s = float(np.mean(kerns)) * 1e-3
    alg = 8 * n   # the P and S slowness of every fine node written once
    out = {
        "metric": METRIC + " | c2f fine-model nodes/s",
        "value": round(n / wall, 1),
        "unit": "fine nodes/s (P and S), device-resident",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(wall * 1e3, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded coarse model, gradient x +-4 % per node, Vs = Vp/1.73 x +-2 %)",
        "config": {"workload": "c2f: 31x31x7 coarse model (P and S, 10 fine intervals per coarse interval) "
                               "to the 301x301x61 fine grid, two resident models updated per step",
                   "fine_nodes": n, "coarse_nodes": co.nxc * co.nyc * co.nzc},
        "kernel_nodes_per_s": round(n / kern_s, 1),
        "roofline": {"bound": "hbm", "achieved": round(alg / kern_s / 1e9, 3), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(alg / kern_s / 1e9 / HBM_PEAK_GBS, 6), "traffic": TRAFFIC_BYTES_C2F,
                     "traffic_source": "profiles/c2f_fine_kernel_301x301x61.md",
                     "kernel": "c2f_prep_kernel + c2f_fine_kernel + c2f_range_kernel", "kernel_ms": round(kern_s * 1e3, 4),
                     "algorithmic_bytes": alg},
    }
    if not args.no_cpu and os.path.exists(C2F_EXE):
        cw, cfine = c2f_reference(co, (nz, ny, nx))
        out["cpu_baseline"] = {"value": round(n / cw, 1), "unit": "fine nodes/s (P and S)", "cores": 1,
                               "kind": "reference",
                               "sample": "the whole workload: the reference c2f (oracle/_ref/c2f_ref) on the same "
                                         "coarse model, one process, wall time incl. reading the coarse file and "
                                         "writing its three fine-model files"}
        out["gpu_over_cpu"] = round(cw / wall, 1)
        out["nodes_differing_vs_reference"] = int((cfine.view(np.uint32) != fine.view(np.uint32)).sum())
    else:
        out["cpu_baseline"] = None
    ctx.release_model(mp)
    ctx.release_model(ms)
    ctx.close()
    return out


# ------------------------------------------------------------------ location

def loc_oracle_rate(params, events, host, sample):
    """cpu_baseline leg: the oracle's location (oracle/loc_oracle.c, a scalar
    restatement of sphfdloc's per-event loop) on one core over the first
    `sample` events; also returns its results for the comparison."""
    sys.path.insert(0, os.path.join(REPO, "tests"))
    import oracle_bind as ob
    grid = ob.loc_grid(**params)
    outs, t0 = [], time.perf_counter()
    for ev in events[:sample]:
        outs.append(ob.locate(grid, [host[int(q)] for q in ev["qs"]], ev["obstime"], ev["pwt"], ev["phase"],
                              ev["isgood"]))
    return sample / (time.perf_counter() - t0), outs


def run_loc(args):
    """sphfdloc's grid search (src/sphfdloc.c:752-1130, find_time 1305-1345) over
    resident tables: the 96 tables of the default workload (161 x 161 x 81,
    48 stations P and S) solved once, untimed; a step locates 256 synthetic
    events (24-40 picks each, outliers and flagged picks included): every
    search pass on the device, the outlier decisions on the host between
    passes."""
    g, vp, vs, stations = workloads.c2()
    ctx = fdgpu.Context(0)
    mp, ms = ctx.model(vp), ctx.model(vs)
    specs = []
    for st in stations:
        for m in (mp, ms):
            specs.append(fdgpu.TableSpec(workloads.table_par(g, st), m))
    ids = ctx.solve(specs, fdgpu.BATCHED)
    host = [ctx.read(i, (g.nz, g.ny, g.nx)) for i in ids]
    events = workloads.loc_events(g, lambda q, k, j, i: float(host[q][k, j, i]), len(ids), count=args.events)
    for ev in events:
        ev["table"] = [ids[int(q)] for q in ev["qs"]]
    params = workloads.loc_params(g)
    cp = fdgpu.loc_params(**params)
    # the events as C records, built once (a C caller hands its own records
    # to fdgpu_locate); the timed region is the fdgpu_locate call
    batch = fdgpu.LocBatch(events)
    for _ in range(args.warmup):
        batch.reset()
        ctx.locate_batch(cp, batch)
    walls, node_ms, node_bytes, pass_ms, passes = [], [], [], [], []
    for s in range(args.steps):
        batch.reset()
        t0 = time.perf_counter()
        ctx.locate_batch(cp, batch)
        walls.append(time.perf_counter() - t0)
        t = ctx.locate_timing()
        node_ms.append(t["node_ms"])
        node_bytes.append(t["node_bytes"])
        pass_ms.append(t["pass_ms"])
        passes.append(t["passes"])
        print(f"# step {s}: {walls[-1] * 1e3:.2f} ms wall, {t['passes']} passes, {t['pass_ms']:.2f} ms device, "
              f"{t['node_ms']:.2f} ms node search", file=sys.stderr, flush=True)
    wall = float(np.mean(walls))
    got = batch.results()
    nev = len(events)
    located = sum(r["status"] == 1 for r in got)
    err = [max(abs(r["node"][0] - ev["src"][0]), abs(r["node"][1] - ev["src"][1]), abs(r["node"][2] - ev["src"][2]))
           for r, ev in zip(got, events) if r["status"] == 1]
    ach = float(np.sum(node_bytes)) / (float(np.sum(node_ms)) * 1e-3) / 1e9
    out = {
        "metric": METRIC + " | sphfdloc events/s",
        "value": round(nev / wall, 2),
        "unit": "events located/s (tables resident)",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(wall * 1e3, 2), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32/f64",
        "data": "synthetic (events at random nodes of the default workload's tables, 0.05 s noise, 5 % outliers)",
        "config": {"workload": f"loc: {nev} events x 24-40 picks over 96 resident tables (48 stations P+S) of "
                               "161x161x81, sphfdloc defaults (kmin 2, ndiv 20, ndiv2 20, nthres 8); timed: the "
                               "fdgpu_locate call on the events' prepared C records",
                   "events": nev, "picks": int(sum(len(e["qs"]) for e in events)),
                   "nodes_searched": g.nx * g.ny * (g.nz - 2)},
        "passes_per_step": float(np.mean(passes)),
        "device_ms_per_step": round(float(np.mean(pass_ms)), 3),
        "located": located, "max_node_offset_from_source": max(err) if err else None,
        "roofline": {"bound": "hbm", "achieved": round(ach, 3), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(ach / HBM_PEAK_GBS, 6), "traffic": TRAFFIC_BYTES_LOC,
                     "traffic_source": "profiles/loc_nodes_kernel_256events_96tables.md", "kernel": "loc_nodes_kernel",
                     "kernel_ms": round(float(np.sum(node_ms)) / float(np.sum(passes)), 4),
                     "algorithmic_bytes": float(np.sum(node_bytes)) / float(np.sum(passes)),
                     "what": "per launch: 4 bytes per table a usable pick of the pass names, per searched node"},
    }
    out["cpu_baseline"] = None
    if not args.no_cpu:
        n = max(1, min(args.loc_cpu_events, nev))
        rate, outs = loc_oracle_rate(params, events, host, n)
        out["cpu_baseline"] = {"value": round(rate, 4), "unit": "events located/s", "cores": 1, "kind": "port",
                               "sample": f"the first {n} of the {nev} events: the oracle's location "
                                         "(oracle/loc_oracle.c, scalar C) over the same tables in host memory"}
        out["gpu_over_cpu"] = round(out["value"] / rate, 1)
        same = all(r["node"] == o["node"] and r["exm"] == o["exm"] and r["eym"] == o["eym"] and r["ezm"] == o["ezm"]
                   and np.float32(r["stdmin"]).view(np.uint32) == np.float32(o["stdmin"]).view(np.uint32)
                   and r["passes"] == o["passes"] for r, o in zip(got, outs))
        out["sample_identical_to_oracle"] = bool(same)
    for i in ids:
        ctx.release(i)
    ctx.release_model(mp)
    ctx.release_model(ms)
    ctx.close()
    return out


# ------------------------------------------------------------------ rays

# Per segment, the eight corner times of the station's table a step reads
# (the algorithmic bytes of the ray kernel; derivatives add 8 bytes each).
RAY_BYTES_PER_SEGMENT = 32
# HBM bytes of one ray launch of the rays workload (FETCH_SIZE x 1.99992 +
# WRITE_SIZE, copy-calibrated) (profiles/ray_kernel_rays_96000.md)
TRAFFIC_BYTES_RAYS = 18.06e9


def ray_workload(events: int, seed: int = 9):
    """The 96 tables of 161 x 161 x 81 (48 stations, P and S, the c2 models),
    sphrayderv's grid over them with a 33 x 33 x 17 coarse grid (every fifth
    node), and `events` events at random positions 2 cells inside the sides
    and below 2 km, each with P and S picks at every station."""
    from fdtomoc_amd import rayderv
    g, vp, vs, stations = workloads.c2()
    ig = lambda n: np.full(n - 1, 5, dtype=np.int64)  # noqa: E731
    grid = rayderv.grid_from_spec(33, 33, 17, ig(33), ig(33), ig(17), g.x0, g.y0, g.z0, g.h)
    assert (grid.nx, grid.ny, grid.nz) == (g.nx, g.ny, g.nz)
    sta = [rayderv.station(st.name, st.lat, st.lon, int(round(-st.depth * 1000))) for st in stations]
    rng = np.random.default_rng(seed)
    evs = []
    for _ in range(events):
        fi, fj, fk = rng.uniform(2, g.nx - 3), rng.uniform(2, g.ny - 3), rng.uniform(2, g.nz - 3)
        evs.append((grid.x0 + fi * grid.df, grid.y0 + fj * grid.dq, grid.z0 + fk * grid.h))
    return g, vp, vs, stations, grid, sta, evs


def run_rays(args):
    """sphrayderv's rays (src/sphrayderv.c:1306-2250) through resident tables:
    the 96 tables solved once, untimed; a step traces every pick of every
    event (fdgpu_rays_trace: host set-up, the launches, the derivatives back
    in host memory)."""
    sys.path.insert(0, os.path.join(REPO, "tests"))
    g, vp, vs, stations, grid, sta, evs = ray_workload(args.ray_events)
    ctx = fdgpu.Context(0)
    mp, ms = ctx.model(vp), ctx.model(vs)
    specs = []
    for st in stations:
        for m in (mp, ms):
            specs.append(fdgpu.TableSpec(workloads.table_par(g, st), m))
    ids = ctx.solve(specs, fdgpu.BATCHED)
    keys, coords = [], []
    for e in evs:
        for q, s in enumerate(sta):
            for ph in range(2):
                coords.append(e + (s.lon, s.colat, s.z))
                keys.append(2 * q + ph)
    keys = np.array(keys)
    # the rays as fdgpu_ray records, built once (a C caller hands its own)
    rays = fdgpu.ray_array(np.asarray(ids)[keys], coords)
    # derivative arrays sized by the first call and reused by the others, as
    # an inversion loop that traces every iteration keeps its buffers
    res, idx, val = ctx.trace_rays(grid, rays)
    bufs = (np.empty(len(idx), np.int32), np.empty(len(idx), np.float32))
    bufs[0][:] = 0
    bufs[1][:] = 0
    for _ in range(args.warmup):
        res, idx, val = ctx.trace_rays(grid, rays, out=bufs)
    walls, kms = [], []
    for s in range(args.steps):
        t0 = time.perf_counter()
        res, idx, val = ctx.trace_rays(grid, rays, out=bufs)
        walls.append(time.perf_counter() - t0)
        kms.append(ctx.last_kernel_ms())
        print(f"# step {s}: {walls[-1] * 1e3:.1f} ms wall, {kms[-1]:.1f} ms kernels", file=sys.stderr, flush=True)
    wall, kern = float(np.mean(walls)), float(np.mean(kms)) * 1e-3
    nok = int((res["status"] == 1).sum())
    nseg = int(res["nseg"].sum())
    alg = RAY_BYTES_PER_SEGMENT * nseg + 8 * len(idx)
    ach = alg / kern / 1e9
    out = {
        "metric": METRIC + " | sphrayderv rays/s",
        "value": round(len(rays) / wall, 1),
        "unit": "rays traced/s (tables resident, derivatives in host memory)",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(wall * 1e3, 2), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64/f32",
        "data": "synthetic (events at random positions in the c2 models' 96 tables, P and S at 48 stations)",
        "config": {"workload": f"rays: {len(evs)} events x 48 stations x (P, S) = {len(rays)} rays through 96 "
                               f"resident tables of 161x161x81, derivatives on a 33x33x17 coarse grid (18,513 "
                               f"variables); timed: the fdgpu_rays_trace call",
                   "rays": len(rays), "segments": nseg, "derivatives": int(len(idx))},
        "kernel_ms_per_step": round(kern * 1e3, 3),
        "rays_reaching_station": nok,
        "status_counts": {int(k): int(v) for k, v in zip(*np.unique(res["status"], return_counts=True))},
        "segments_per_ray": {"mean": round(float(res["nseg"].mean()), 1), "p99": int(np.percentile(res["nseg"], 99)),
                             "max": int(res["nseg"].max())},
        "segments_per_s": round(nseg / kern, 1),
        "roofline": {"bound": "hbm", "achieved": round(ach, 3), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(ach / HBM_PEAK_GBS, 6),
                     "traffic": TRAFFIC_BYTES_RAYS if len(evs) == 1000 else None,
                     "traffic_source": "profiles/ray_kernel_rays_96000.md", "kernel": "ray_kernel",
                     "kernel_ms": round(kern * 1e3, 3), "algorithmic_bytes": alg,
                     "what": "per step: 32 bytes of table corners per segment + 8 bytes per derivative; the walk "
                             "is a chain of dependent FP64 steps per lane, so latency, not bandwidth, bounds it"},
    }
    out["cpu_baseline"] = None
    if not args.no_cpu:
        import oracle_bind as ob
        cores = host_cores(args.cpu_threads)
        host = [ctx.read(i, (g.nz, g.ny, g.nx)) for i in ids]
        n = max(1, min(args.ray_cpu_rays, len(rays)))
        pick = np.linspace(0, len(rays) - 1, n).astype(int)
        t0 = time.perf_counter()
        ref = ob.ray_trace(grid, [host[keys[r]] for r in pick], [coords[r] for r in pick], nthreads=cores)
        rate = n / (time.perf_counter() - t0)
        out["cpu_baseline"] = {"value": round(rate, 2), "unit": "rays traced/s", "cores": cores, "kind": "port",
                               "host": host_info(),
                               "sample": f"{n} of the {len(rays)} rays, evenly spaced: the ray oracle "
                                         f"(oracle/ray_oracle.c, the reference's loop in C, equal to the "
                                         f"reference program's outputs) on {cores} host threads"}
        out["gpu_over_cpu"] = round(out["value"] / rate, 1)
        same = exact = 0
        worst = 0.0
        for r, o in zip(pick, ref):
            d = res[r]
            if d["status"] != o["status"] or d["nseg"] != o["nseg"]:
                continue
            same += 1
            if d["status"] == 1:
                gv = val[d["offset"]:d["offset"] + d["ndu"]]
                gi = idx[d["offset"]:d["offset"] + d["ndu"]]
                if np.array_equal(gi, o["idx"]):
                    exact += bool(np.array_equal(gv.view(np.uint32), o["val"].view(np.uint32)))
                    worst = max(worst, float(np.max(np.abs(gv - o["val"]) / np.abs(o["val"]))) if len(gv) else 0.0)
                else:
                    worst = float("inf")
            else:
                exact += 1
        out["sample_vs_oracle"] = {"rays": n, "same_status_and_segments": same, "bit_identical": exact,
                                   "max_rel_derivative_diff": worst,
                                   "what": "same status and segment count, and derivatives bit for bit; the "
                                           "device's sines, cosines and arctangents are correctly rounded, the "
                                           "host C library's not in about one call in a thousand"}
    for i in ids:
        ctx.release(i)
    ctx.release_model(mp)
    ctx.release_model(ms)
    ctx.close()
    return out


# ------------------------------------------------------------------ calibration

def calibrate_copy(mib: int) -> None:
    """One device-to-device copy of known size, for scaling the FETCH_SIZE /
    WRITE_SIZE counters of a profiled run (each should read MIB MiB)."""
    import torch
    n = mib * (1 << 20) // 4
    a = torch.arange(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    b.copy_(a)
    torch.cuda.synchronize()
    assert bool(torch.equal(a[-4:], b[-4:]))
    print(json.dumps({"copy_bytes": 4 * n, "read_bytes": 4 * n, "write_bytes": 4 * n}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="c3", choices=["c1", "c2", "c2w", "c3", "c4", "c5", "c2f", "loc", "rays"])
    ap.add_argument("--plain-sample", type=int, default=2)
    ap.add_argument("--cpu-sample", type=int, default=-1, help="tables for the CPU baseline (-1: one per core)")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-analytic", action="store_true", help="skip the untimed analytic-error solve (profiling runs)")
    ap.add_argument("--no-autotune", action="store_true", help="keep the default per-face schedule")
    ap.add_argument("--schedule", type=int, default=-1, help="set the per-face schedule (CHANGE budget) instead of tuning")
    ap.add_argument("--wave", type=int, default=-1, help="wavefront flags (1 first pass, 3 all passes) instead of tuning")
    ap.add_argument("--team", type=int, default=-1, help="workgroups per large table's wavefront passes (0 auto, 1 never)")
    ap.add_argument("--team-share", type=int, default=-1, help="teams for tables of at least this %% of the mean share")
    ap.add_argument("--requests", type=int, default=96, help="c4: requests in the stream")
    ap.add_argument("--rate", type=float, default=3.0, help="c4: mean arrivals per second")
    ap.add_argument("--tables", type=int, default=2000, help="c5: resident tables (1000 stations x P, S)")
    ap.add_argument("--iterations", type=int, default=5, help="c5: inversion iterations (the first solves)")
    ap.add_argument("--events", type=int, default=256, help="loc: events per step")
    ap.add_argument("--loc-cpu-events", type=int, default=8, help="loc: events in the CPU baseline sample")
    ap.add_argument("--ray-events", type=int, default=1000, help="rays: events per step (96 rays each)")
    ap.add_argument("--ray-cpu-rays", type=int, default=4000, help="rays: rays in the CPU baseline sample")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run a device copy of MIB MiB (HBM counter calibration) and exit")
    args = ap.parse_args()
    if args.calibrate_copy:
        return calibrate_copy(args.calibrate_copy)
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: measures one GPU; the station tables shard across GPUs with no "
              "communication on the hot path, so multi-GPU runs are not part of it.", file=sys.stderr)
        sys.exit(2)
    run = {"c1": run_c1, "c2": run_c2, "c3": run_c3, "c4": run_c4, "c5": run_c5, "c2f": run_c2f, "loc": run_loc, "c2w": run_c2w,
           "rays": run_rays}[args.workload]
    print(json.dumps(run(args)), flush=True)


if __name__ == "__main__":
    main()
