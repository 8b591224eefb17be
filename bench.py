"""Benchmark of the batched MI355X sphfd path.

Workload (default, the configuration the headline metric is quoted on): 48
stations uniform-random on the top surface of a 161 x 161 x 81 grid at 1 km
centred at 23.5N 121E, P and S tables (Vs = Vp/1.73) -> 96 station solves per
step, FP32 1D gradient 5.0 -> 8.0 km/s over 0-64 km (continued below) times a
+-5% checkerboard of 20 km cells, reverse=50 as make_par_file writes.

A step is one batched solve of all 96 tables (one persistent launch); the
velocity models are resident in HBM before timing starts.  The line reports
station-solves/s and grid nodes/s of the batched path, the plain path (one
table per launch) on a sample, the HBM roofline of the sweep kernel and the
reference sphfd (compiled from its own sources, oracle/_ref) timed on the host
on a bounded sample of the same tables, whose outputs are also compared with
the GPU tables bit for bit.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--cpu-sample N] [--plain-sample M]
    python bench.py --calibrate-copy MIB     # one known-size copy, for HBM counter scaling
"""
from __future__ import annotations

import argparse
import json
import os
import shutil
import subprocess
import sys
import tempfile
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)

from fdtomoc_amd import fdgpu, synth, tables  # noqa: E402

METRIC = "stations/s & nodes/s batched vs plain at <=1e-4 s err; p50/p99 solve latency"
HBM_PEAK_GBS = 8000.0
BYTES_PER_VISIT = 12          # slowness read + time read + time write per node visit
# HBM bytes (FETCH_SIZE x 1.99992 + WRITE_SIZE x 1.0, copy-calibrated) that
# rocprofv3 counted for one launch of the default 96-table workload
# (profiles/sphfd_sweep_kernel_c2_96tables_hbm_traffic.md)
TRAFFIC_BYTES_C2 = 367.03e9
REF_EXE = os.path.join(REPO, "oracle", "_ref", "sphfd_ref")


def c2_workload(seed: int = 2):
    g = synth.centred_grid(161, 161, 81, 1.0)
    vp = synth.gradient_velocity(g, 5.0, 8.0, 64.0) * synth.checkerboard(g, 20.0, 0.05)
    vs = (vp / np.float32(1.73)).astype(np.float32)
    stations = synth.random_surface_stations(g, np.random.default_rng(seed), 48)
    return g, vp.astype(np.float32), vs, stations


def table_par(g, st, reverse=50):
    return dict(nx=g.nx, ny=g.ny, nz=g.nz, h=g.h, x0=g.x0, y0=g.y0, z0=g.z0,
                fxs=st.lon, fys=st.lat, fzs=st.depth, reverse=reverse)


def cpu_reference(g, vp, vs, stations, sample: int, threads: int):
    """Runs the compiled reference sphfd on `sample` tables with OpenMP threads
    (its own parallel-for over tables, src/sphfd.c:247).  Returns (seconds,
    {table index: times})."""
    wd = tempfile.mkdtemp(prefix="sphfd_cpu_")
    try:
        head = tables.make_header(g.nx, g.ny, g.nz, g.x0, g.y0, g.z0, g.df_deg(), g.dq_deg(), g.h)
        tables.write_grid(os.path.join(wd, "VP.mod"), head, vp)
        tables.write_grid(os.path.join(wd, "VS.mod"), head, vs)
        os.makedirs(os.path.join(wd, "out"))
        lines = []
        picked = []
        for t in range(sample):
            st = stations[t // 2]
            phase = t % 2
            p = table_par(g, st)
            name = f"{st.name}.{'ptimes' if phase == 0 else 'stimes'}"
            txt = "".join(f"{k}={v!r}\n" for k, v in p.items())
            txt += f"timefile={name}\nvelfile={'VP.mod' if phase == 0 else 'VS.mod'}\n"
            pf = os.path.join(wd, f"t{t}.par")
            open(pf, "w").write(txt)
            lines.append(f"par=t{t}.par")
            picked.append((2 * (t // 2) + phase, name))
        open(os.path.join(wd, "parlist"), "w").write("\n".join(lines) + "\n\n")
        open(os.path.join(wd, "spec"), "w").write("parlist parlist\ntimedir out\n")
        env = dict(os.environ, OMP_NUM_THREADS=str(threads))
        t0 = time.perf_counter()
        r = subprocess.run([REF_EXE], input="spec\n", cwd=wd, env=env, capture_output=True, text=True)
        dt = time.perf_counter() - t0
        if r.returncode != 0:
            raise RuntimeError("reference sphfd failed: " + r.stderr[-1000:])
        outs = {}
        for idx, name in picked:
            outs[idx] = np.array(tables.read_grid(os.path.join(wd, "out", name)).data)
        return dt, outs
    finally:
        shutil.rmtree(wd, ignore_errors=True)


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
    ap.add_argument("--plain-sample", type=int, default=4)
    ap.add_argument("--cpu-sample", type=int, default=-1, help="tables for the CPU baseline (-1: one per core)")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run a device copy of MIB MiB (HBM counter calibration) and exit")
    args = ap.parse_args()
    if args.calibrate_copy:
        return calibrate_copy(args.calibrate_copy)
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: this build measures one GPU; the station tables shard across GPUs with no "
              "communication on the hot path, so multi-GPU runs are not part of it.", file=sys.stderr)
        sys.exit(2)

    g, vp, vs, stations = c2_workload()
    ctx = fdgpu.Context(0)
    mp, ms = ctx.model(vp), ctx.model(vs)
    specs = []
    for st in stations:
        specs.append(fdgpu.TableSpec(table_par(g, st), mp))
        specs.append(fdgpu.TableSpec(table_par(g, st), ms))
    ntab = len(specs)

    def step(mode=fdgpu.BATCHED, which=None):
        sel = specs if which is None else [specs[i] for i in which]
        t0 = time.perf_counter()
        ids = ctx.solve(sel, mode)
        dt = time.perf_counter() - t0
        return ids, dt, ctx.last_kernel_ms()

    for _ in range(args.warmup):
        ids, _, _ = step()
        for i in ids:
            ctx.release(i)

    wall, kern = [], []
    last_ids = None
    for s in range(args.steps):
        ids, dt, kms = step()
        wall.append(dt)
        kern.append(kms)
        print(f"# step {s}: {dt * 1e3:.1f} ms wall, {kms:.1f} ms kernel", file=sys.stderr, flush=True)
        if s < args.steps - 1:
            for i in ids:
                ctx.release(i)
        else:
            last_ids = ids
    stats = [ctx.stats(i) for i in last_ids]
    restarts = np.array([s["restarts"] for s in stats])
    n = g.n
    visits = int(np.sum(n * (1 + restarts)))
    wall = np.array(wall)
    kern_s = np.array(kern) / 1e3
    ms_per_step = float(np.mean(wall) * 1e3)
    stations_s = ntab / float(np.mean(wall))
    nodes_s = ntab * n / float(np.mean(wall))
    achieved = BYTES_PER_VISIT * visits / float(np.mean(kern_s)) / 1e9

    # plain path: one table per launch, on a sample
    plain_s = None
    psel = list(range(min(args.plain_sample, ntab)))
    if psel:
        pids, pdt, _ = step(fdgpu.PLAIN, psel)
        for i in pids:
            ctx.release(i)
        plain_s = len(psel) / pdt

    out = {
        "metric": METRIC, "value": round(stations_s, 3), "unit": "station-solves/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_per_step, 2),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded gradient x checkerboard model, random surface stations)",
        "config": {"workload": "48 stations P+S = 96 tables, 161x161x81 at 1 km, gradient 5-8 km/s x "
                               "+-5% 20 km checkerboard, reverse=50, one batched launch per step",
                   "tables": ntab, "nodes_per_table": n, "reverse_passes_mean": float(restarts.mean())},
        "nodes_per_s": round(nodes_s, 1),
        "plain_station_solves_per_s": None if plain_s is None else round(plain_s, 4),
        "batched_over_plain": None if plain_s is None else round(stations_s / plain_s, 2),
        "latency_ms": {"p50": round(float(np.percentile(wall, 50)) * 1e3, 2),
                       "p99": round(float(np.percentile(wall, 99)) * 1e3, 2),
                       "what": "batched solve of all 96 tables (submit to results resident)"},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 3), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 6), "traffic": TRAFFIC_BYTES_C2,
                     "traffic_source": "profiles/sphfd_sweep_kernel_c2_96tables_hbm_traffic.md",
                     "kernel": "sphfd_sweep_kernel", "kernel_ms": round(float(np.mean(kern_s)) * 1e3, 2),
                     "algorithmic_bytes": BYTES_PER_VISIT * visits},
    }

    if not args.no_cpu and os.path.exists(REF_EXE):
        cores = args.cpu_threads or min(os.cpu_count() or 1, 16)
        sample = cores if args.cpu_sample < 0 else args.cpu_sample
        sample = max(1, min(sample, ntab))
        dt, outs = cpu_reference(g, vp, vs, stations, sample, cores)
        maxerr, ndiff = 0.0, 0
        for idx, tref in outs.items():
            tg = ctx.read(last_ids[idx], tref.shape)
            maxerr = max(maxerr, float(np.max(np.abs(tg.astype(np.float64) - tref))))
            ndiff += int((tg.view(np.uint32) != tref.view(np.uint32)).sum())
        out["cpu_baseline"] = {"value": round(sample / dt, 5), "unit": "station-solves/s", "cores": cores,
                               "kind": "reference",
                               "sample": f"{sample} of the 96 tables, compiled reference sphfd, OpenMP "
                                         f"parallel-for over tables with {cores} threads, wall time incl. file I/O"}
        out["gpu_over_cpu"] = round(stations_s / (sample / dt), 2)
        out["max_abs_err_vs_reference_s"] = maxerr
        out["nodes_differing_vs_reference"] = ndiff
    else:
        out["cpu_baseline"] = None
    for i in last_ids:
        ctx.release(i)
    ctx.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
