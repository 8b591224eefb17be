"""GPU: sphrayderv's rays through resident tables (fdgpu_rays_trace) against
the ray oracle (oracle/ray_oracle.c, itself equal to the reference program's
outputs bit for bit on tests/golden/ray_cases.npz).

The tables are solved on the device (bit-identical to the sphfd oracle) and
stay resident; the rays of every pick of every event are traced in one call.
Bar: every ray's status and segment count equal; the event time and its
derivatives bit for bit (no transcendental functions are involved); the
derivatives' indices equal and values within 1e-5 relative for every ray,
bit for bit for at least 95% of them.  The device computes the sines,
cosines and arctangents of a ray's positions correctly rounded
(fdtomoc_amd/csrc/cr_math.h); the host C library the reference and the
oracle use is not correctly rounded in about one call in a thousand, and a
float derived from such a position can then differ by a unit
(tests/test_cr_math.py); incidence and azimuth within 1e-4 degrees (float
arctangents of the device library)."""
from __future__ import annotations

import os

import numpy as np
import pytest

import oracle_bind as ob
import ray_cases as rc
from cases import par_dict

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    from fdtomoc_amd import fdgpu
    c = fdgpu.Context(0)
    yield c
    c.close()


def _device_tables(ctx, case):
    """The case's P and S tables solved on the device; returns
    {(station, phase): table id} and the device tables as arrays."""
    from fdtomoc_amd import fdgpu
    mp, ms = ctx.model(case.vp), ctx.model(case.vs)
    keys, specs = [], []
    for st in case.stations:
        par = par_dict(ob.Par(**case.table_par(st)))
        for ph, m in (("P", mp), ("S", ms)):
            keys.append((st[0], ph))
            specs.append(fdgpu.TableSpec(par, m))
    ids = ctx.solve(specs, fdgpu.BATCHED)
    ctx.release_model(mp)
    ctx.release_model(ms)
    shape = (case.grid.nz, case.grid.ny, case.grid.nx)
    return dict(zip(keys, ids)), {k: ctx.read(i, shape) for k, i in zip(keys, ids)}


def _compare(res, idx, val, ref, label):
    exact = 0
    for q, (d, o) in enumerate(zip(res, ref)):
        assert d["status"] == o["status"] and d["nseg"] == o["nseg"], (label, q, d, o)
        for f in ("tev", "dtdx", "dtdy", "dtdz"):
            assert np.float64(d[f]).tobytes() == np.float64(o[f]).tobytes() or d[f] == o[f], (label, q, f)
        if d["status"] != 1:
            assert d["ndu"] == 0
            continue
        gi = idx[d["offset"]:d["offset"] + d["ndu"]]
        gv = val[d["offset"]:d["offset"] + d["ndu"]]
        np.testing.assert_array_equal(gi, o["idx"], err_msg=f"{label} ray {q}")
        np.testing.assert_allclose(gv, o["val"], rtol=1e-5, atol=0, err_msg=f"{label} ray {q}")
        same = bool(np.array_equal(gv.view(np.uint32), o["val"].view(np.uint32)))
        exact += same
        if not same:
            dif = gv.view(np.uint32) != o["val"].view(np.uint32)
            rel = np.max(np.abs(gv[dif] - o["val"][dif]) / np.abs(o["val"][dif]))
            print(f"{label} ray {q}: {int(dif.sum())} of {len(gv)} derivatives differ, max rel {rel:.2e}, "
                  f"ulps {np.abs(gv[dif].view(np.int32) - o['val'][dif].view(np.int32)).max()}", flush=True)
        assert d["first_in_cube"] == o["first_in_cube"]
        if not o["first_in_cube"]:
            assert d["ai"] == pytest.approx(o["ai"], abs=1e-4) and d["az"] == pytest.approx(o["az"], abs=1e-4)
        assert d["rdev"] == pytest.approx(o["rdev"], rel=1e-4, abs=1e-6, nan_ok=True)
    return exact


@pytest.mark.parametrize("seed", [11, 12, 13])
def test_rays_match_oracle_on_reference_cases(ctx, seed):
    """The three cases whose oracle results equal the reference program's
    files (tests/test_ray_oracle.py), traced on the device."""
    case = rc.ray_case(seed)
    ids, dev = _device_tables(ctx, case)
    host = rc.case_tables(case)
    for k in host:   # the device tables are the oracle's, bit for bit
        assert np.array_equal(dev[k].view(np.uint32), host[k].view(np.uint32)), k
    rays = rc.case_rays(case)
    res, idx, val = ctx.trace_rays(case.grid, [(ids[r["key"]],) + tuple(r["ray"]) for r in rays])
    ref = ob.ray_trace(case.grid, [host[r["key"]] for r in rays], [r["ray"] for r in rays])
    exact = _compare(res, idx, val, ref, f"seed {seed}")
    ok = sum(o["status"] == 1 for o in ref)
    print(f"seed {seed}: {len(rays)} rays, {ok} reach their station, {exact} bit-identical derivatives", flush=True)
    assert exact >= 0.95 * ok
    for i in ids.values():
        ctx.release(i)


@pytest.mark.parametrize("ido1d", [0, 1])
def test_many_events_one_call(ctx, ido1d):
    """A larger batch: 6 stations, 40 events, P and S at every station (480
    rays) in one call, three- and one-dimensional derivatives."""
    case = rc.ray_case(50 + ido1d, nsta=6, nev=40, ido1d=ido1d)
    ids, dev = _device_tables(ctx, case)
    rays = rc.case_rays(case)
    res, idx, val = ctx.trace_rays(case.grid, [(ids[r["key"]],) + tuple(r["ray"]) for r in rays])
    ref = ob.ray_trace(case.grid, [dev[r["key"]] for r in rays], [r["ray"] for r in rays])
    exact = _compare(res, idx, val, ref, f"ido1d {ido1d}")
    ok = sum(o["status"] == 1 for o in ref)
    print(f"ido1d {ido1d}: {len(rays)} rays, {ok} reach their station, {exact} bit-identical", flush=True)
    assert ok >= 0.8 * len(rays) and exact >= 0.95 * ok
    for i in ids.values():
        ctx.release(i)


def test_small_output_arrays_report_sizes(ctx):
    """Arrays too small: FDGPU_ESPACE with every ray's count, then the same
    results into arrays of the stated size."""
    from fdtomoc_amd import fdgpu
    case = rc.ray_case(12)
    ids, _ = _device_tables(ctx, case)
    rays = [(ids[r["key"]],) + tuple(r["ray"]) for r in rc.case_rays(case)]
    with pytest.raises(fdgpu.FdgpuError) as e:
        ctx.trace_rays(case.grid, rays, du_cap=10)
    assert e.value.code == fdgpu.ESPACE
    res, idx, val = ctx.trace_rays(case.grid, rays)
    total = sum(d["ndu"] for d in res)
    res2, idx2, val2 = ctx.trace_rays(case.grid, rays, du_cap=total)
    assert [d["offset"] for d in res] == [d["offset"] for d in res2]
    assert np.array_equal(idx, idx2) and np.array_equal(val.view(np.uint32), val2.view(np.uint32))
    for i in ids.values():
        ctx.release(i)


@pytest.mark.parametrize("seed", [11, 12, 13])
def test_program_files_with_device_rays(ctx, tmp_path, seed):
    """sphrayderv's program (fdtomoc_amd.rayderv.run) with the rays traced on
    the device through tables loaded from the run's files: its eight output
    files against the reference program's for the same inputs
    (tests/golden/ray_cases.npz), byte for byte."""
    from fdtomoc_amd import rayderv
    golden = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "ray_cases.npz"))
    case = rc.ray_case(seed)
    spec = rc.write_inputs(case, str(tmp_path), rc.case_tables(case))
    trace, ids = rayderv.device_trace(ctx, str(tmp_path), "ttimes")
    rayderv.run(spec, trace, workdir=str(tmp_path))
    for i in ids.values():
        ctx.release(i)
    for f in ("dtds.bin", "res.bin", "dtdh.txt", "book.txt", "hit.txt", "scle.txt", "dot.txt", "head.txt"):
        assert open(tmp_path / f, "rb").read() == golden[f"c{seed}_file_{f}"].tobytes(), f


def test_program_entry_point(tmp_path):
    """`python -m fdtomoc_amd.rayderv <<< spec` in a run directory, as run.sh
    calls the reference program: the same eight files as the reference's."""
    import subprocess
    import sys
    golden = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "ray_cases.npz"))
    case = rc.ray_case(12)
    spec = rc.write_inputs(case, str(tmp_path), rc.case_tables(case))
    repo = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    p = subprocess.run([sys.executable, "-m", "fdtomoc_amd.rayderv"], input=spec + "\n", cwd=str(tmp_path),
                       capture_output=True, text=True, timeout=300, env=dict(os.environ, PYTHONPATH=repo))
    assert p.returncode == 0, p.stderr[-2000:]
    for f in ("dtds.bin", "res.bin", "dtdh.txt", "book.txt", "hit.txt", "scle.txt", "dot.txt", "head.txt"):
        assert open(tmp_path / f, "rb").read() == golden[f"c12_file_{f}"].tobytes(), f
