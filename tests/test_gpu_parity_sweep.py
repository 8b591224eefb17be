"""A wider parity sweep of the sweep kernel against the oracle, bit for bit:
96 seeded small cases, each with random combinations of the par options the
reference exposes (growth rates of all six faces, integer source nodes,
explicit dq/df, maxoff, NCUBE, reverse, headpref), solved in ONE batched
launch of 96 different models and sizes (src/sphfd.c:383-531 keys, the sweep
1008-5678)."""
import numpy as np
import pytest

import oracle_bind as ob
from cases import make_case, par_dict

pytestmark = pytest.mark.gpu


def _case(seed):
    par, v = make_case(seed, nmin=5, nmax=24)
    rng = np.random.default_rng(seed + 7777)
    for key in ("iminus", "iplus", "jminus", "jplus", "kminus", "kplus"):
        if rng.random() < 0.2:
            setattr(par, key, 2)
    explicit = rng.random() < 0.2
    if explicit:
        # explicit angular spacings (degrees), close to the derived ones; the
        # grid's extent changes, so the source becomes an integer node
        par.dq = float(np.float64(par.h / 6371.0 / 0.017453292519943295) * rng.uniform(0.9, 1.1))
        par.df = float(par.dq * rng.uniform(1.0, 1.3))
    if explicit or rng.random() < 0.25:
        par.floatsrc = 0
        par.xs = int(rng.integers(0, par.nx))
        par.ys = int(rng.integers(0, par.ny))
        par.zs = int(rng.integers(0, par.nz))
    if rng.random() < 0.1:
        par.maxoff = float(rng.uniform(1.5, 6.0)) * par.h
    return par, v


def test_ninety_six_option_mixes_in_one_launch():
    from fdtomoc_amd import fdgpu
    ctx = fdgpu.Context(0)
    try:
        cases = [_case(1000 + s) for s in range(96)]
        models = [ctx.model(v) for _, v in cases]
        specs = [fdgpu.TableSpec(par_dict(p), m) for (p, _), m in zip(cases, models)]
        ids = ctx.solve(specs, fdgpu.BATCHED)
        bad = []
        for k, ((par, v), tid) in enumerate(zip(cases, ids)):
            t_ref, st = ob.solve(par, ob.slowness(v))
            t = ctx.read(tid, v.shape)
            gst = ctx.stats(tid)
            nd = int((t.view(np.uint32) != t_ref.view(np.uint32)).sum())
            if nd or gst["passes"] != st["passes"] or gst["error"] & ~1:
                bad.append((1000 + k, v.shape, nd, gst["passes"], st["passes"], gst["error"]))
        print(f"{len(cases)} cases, {len(bad)} differing: {bad[:5]}", flush=True)
        assert not bad
        for tid in ids:
            ctx.release(tid)
        for m in models:
            ctx.release_model(m)
    finally:
        ctx.close()


def test_larger_option_mixes():
    """Eight cases of 40..72 nodes per axis with the same random option mixes
    (head waves and several reverse passes are more common at these sizes):
    bit for bit and pass for pass equal to the oracle."""
    from fdtomoc_amd import fdgpu
    ctx = fdgpu.Context(0)
    try:
        cases = []
        for s in range(8):
            par, v = make_case(3000 + s, nmin=40, nmax=72)
            rng = np.random.default_rng(3000 + s + 7777)
            for key in ("iminus", "iplus", "jminus", "jplus", "kminus", "kplus"):
                if rng.random() < 0.2:
                    setattr(par, key, 2)
            cases.append((par, v))
        models = [ctx.model(v) for _, v in cases]
        ids = ctx.solve([fdgpu.TableSpec(par_dict(p), m) for (p, _), m in zip(cases, models)], fdgpu.BATCHED)
        for k, ((par, v), tid) in enumerate(zip(cases, ids)):
            t_ref, st = ob.solve(par, ob.slowness(v))
            t = ctx.read(tid, v.shape)
            gst = ctx.stats(tid)
            print(f"case {3000 + k} {v.shape} passes {st['passes']}", flush=True)
            assert gst["passes"] == st["passes"] and gst["error"] & ~1 == 0
            assert int((t.view(np.uint32) != t_ref.view(np.uint32)).sum()) == 0, 3000 + k
        for tid in ids:
            ctx.release(tid)
        for m in models:
            ctx.release_model(m)
    finally:
        ctx.close()
