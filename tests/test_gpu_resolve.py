"""Re-solving resident tables in place (fdgpu_resolve) after model updates, as
an inversion's iterations do, and the repeated-call paths of the c2f and
location bindings:

  * tables solved once, then for three model updates re-solved in place in
    one batched launch: each equal bit for bit to the oracle on the updated
    velocities and to a fresh fdgpu_solve, the tables' device buffers
    unchanged, restart counts following the new models;
  * plain-mode re-solves, refused tables (srctype 3, unknown ids);
  * a prepared coarse record whose velocities are changed in place between
    updates equals updates from a freshly built record;
  * a location batch located again after reset() equals a fresh locate.
"""
import numpy as np
import pytest

import oracle_bind as ob
from cases import make_case, par_dict

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    from fdtomoc_amd import fdgpu
    c = fdgpu.Context(0)
    yield c
    c.close()


def _bits(a, b):
    return int((np.asarray(a, np.float32).view(np.uint32) != np.asarray(b, np.float32).view(np.uint32)).sum())


def _world(seed=41, n=(44, 40, 24)):
    from fdtomoc_amd import synth
    rng = np.random.default_rng(seed)
    g = synth.centred_grid(*n, 1.0)
    base = synth.gradient_velocity(g, 4.5, 7.5, 20.0).astype(np.float32)
    pars = []
    for st in synth.random_surface_stations(g, rng, 6):
        pars.append(ob.Par(nx=g.nx, ny=g.ny, nz=g.nz, h=g.h, x0=g.x0, y0=g.y0, z0=g.z0, fxs=st.lon, fys=st.lat,
                           fzs=st.depth, reverse=50))
    return rng, g, base, pars


def _update(rng, g, base, it):
    """An inversion iteration's model: the base times fresh layered and
    Gaussian perturbations (a strong low-velocity layer on odd iterations,
    so reverse passes change between iterations)."""
    from fdtomoc_amd import synth
    v = base * synth.layered_perturbation(g, rng, 0.06, 1.0, 4.0) * synth.gaussian_anomalies(g, rng, 2, 0.05, (3, 8))
    if it % 2:
        v[g.nz // 2:g.nz // 2 + 2] *= np.float32(0.8)
    return np.ascontiguousarray(v, dtype=np.float32)


def test_resolve_in_place_after_model_updates(ctx):
    from fdtomoc_amd import fdgpu
    rng, g, base, pars = _world()
    v0 = _update(rng, g, base, 0)
    vs0 = (v0 / np.float32(1.73)).astype(np.float32)
    mp, ms = ctx.model(v0), ctx.model(vs0)
    specs = []
    for p in pars:
        specs += [fdgpu.TableSpec(par_dict(p), mp), fdgpu.TableSpec(par_dict(p), ms)]
    ids = ctx.solve(specs, fdgpu.BATCHED)
    ptrs = [ctx.device_ptr(i) for i in ids]
    restarts = set()
    for it in range(1, 4):
        vp = _update(rng, g, base, it)
        vs = (vp / np.float32(1.73)).astype(np.float32)
        ctx.update_model(mp, vp)
        ctx.update_model(ms, vs)
        ctx.resolve(ids, fdgpu.BATCHED)
        fresh = ctx.solve(specs, fdgpu.BATCHED)
        for q, (tid, fid) in enumerate(zip(ids, fresh)):
            vel = vp if q % 2 == 0 else vs
            t = ctx.read(tid, vel.shape)
            st = ctx.stats(tid)
            t_ref, ost = ob.solve(pars[q // 2], ob.slowness(vel))
            restarts.add(st["restarts"])
            assert st["error"] & ~1 == 0 and st["passes"] == ost["passes"], (it, q)
            assert _bits(t, t_ref) == 0, (it, q)
            assert _bits(t, ctx.read(fid, vel.shape)) == 0, (it, q)
            ctx.release(fid)
        print(f"iteration {it}: {len(ids)} tables re-solved in place, equal to the oracle and to fresh solves",
              flush=True)
        assert [ctx.device_ptr(i) for i in ids] == ptrs
    assert len(restarts) > 1   # the updates change the reverse passes
    for i in ids:
        ctx.release(i)
    ctx.release_model(mp)
    ctx.release_model(ms)


def test_resolve_plain_mode_and_refusals(ctx):
    from fdtomoc_amd import fdgpu
    rng, g, base, pars = _world(43, (30, 26, 18))
    v = _update(rng, g, base, 1)
    m = ctx.model(v)
    ids = ctx.solve([fdgpu.TableSpec(par_dict(p), m) for p in pars[:3]], fdgpu.BATCHED)
    v2 = _update(rng, g, base, 2)
    ctx.update_model(m, v2)
    ctx.resolve(ids, fdgpu.PLAIN)
    for tid, p in zip(ids, pars):
        t_ref, _ = ob.solve(p, ob.slowness(v2))
        assert _bits(ctx.read(tid, v2.shape), t_ref) == 0
    # a restart from an old table (srctype 3): its initial grid is gone
    par, vv = make_case(5)
    par.srctype = 3
    mm = ctx.model(vv)
    old = np.full(vv.shape, 1.0e10, dtype=np.float32)
    old[0, 0, 0] = 0.0
    (t3,) = ctx.solve([fdgpu.TableSpec(par_dict(par), mm, preset=old)])
    for bad in ([t3], [ids[0], 987654]):
        with pytest.raises(fdgpu.FdgpuError) as e:
            ctx.resolve(bad)
        assert e.value.code == -5
    ctx.release(t3)
    ctx.release_model(mm)
    for i in ids:
        ctx.release(i)
    ctx.release_model(m)


def test_prepared_coarse_changed_in_place(ctx):
    """Context.update_models_from_prepared on one record whose velocity array
    is overwritten between updates equals update_models_from_coarse of fresh
    records with the same velocities."""
    from fdtomoc_amd import fdgpu
    rng = np.random.default_rng(5)
    nxc, nyc, nzc = 6, 5, 4
    ig = [np.full(n - 1, s, dtype=np.int32) for n, s in ((nxc, 7), (nyc, 6), (nzc, 5))]

    def coarse():
        vp = rng.uniform(4.5, 7.5, size=nxc * nyc * nzc).astype(np.float32)
        return fdgpu.Coarse(nxc, nyc, nzc, ig[0], ig[1], ig[2], 0.0, 0.0, 0.0, 1.0,
                            np.concatenate([vp, (vp / np.float32(1.7)).astype(np.float32)]))
    c0 = coarse()
    mp, ms = ctx.models_from_coarse(c0)
    ap, as_ = ctx.models_from_coarse(c0)
    prep = fdgpu.PreparedCoarse(c0)
    nx, ny, nz = c0.dims()
    for _ in range(3):
        c = coarse()
        prep.velocity[:] = c.velocity
        ctx.update_models_from_prepared(mp, ms, prep)
        fine = ctx.update_models_from_coarse(ap, as_, c, want_fine=True)
        for a, b in ((mp, ap), (ms, as_)):
            # solve one table on each pair of models and compare (the models'
            # slowness is not readable directly)
            (ta, tb) = ctx.solve([fdgpu.TableSpec(dict(nx=nx, ny=ny, nz=nz, h=1.0, x0=0.0, y0=40.0, z0=0.0,
                                                       xs=nx // 2, ys=ny // 2, zs=0, floatsrc=0, reverse=3), m)
                                  for m in (a, b)])
            assert _bits(ctx.read(ta, (nz, ny, nx)), ctx.read(tb, (nz, ny, nx))) == 0
            ctx.release(ta)
            ctx.release(tb)
        assert fine.shape == (2, nz, ny, nx)
    for m in (mp, ms, ap, as_):
        ctx.release_model(m)


def test_location_batch_reused_after_reset(ctx):
    """locate_batch twice on one LocBatch with reset() in between: both runs
    equal Context.locate on freshly built events."""
    from fdtomoc_amd import fdgpu, synth, workloads
    g = synth.centred_grid(36, 34, 20, 1.0)
    rng = np.random.default_rng(9)
    vp = (synth.gradient_velocity(g, 4.8, 7.6, 16.0) * synth.layered_perturbation(g, rng, amp=0.03)).astype(np.float32)
    m = ctx.model(vp)
    stations = synth.random_surface_stations(g, rng, 12)
    ids = ctx.solve([fdgpu.TableSpec(workloads.table_par(g, st), m) for st in stations])
    host = [ctx.read(i, (g.nz, g.ny, g.nx)) for i in ids]
    events = workloads.loc_events(g, lambda q, k, j, i: float(host[q][k, j, i]), len(ids), count=24, seed=3,
                                  picks=(8, 12))
    for ev in events:
        ev["table"] = [ids[int(q)] for q in ev["qs"]]
        ev["phase"] = "P" * len(ev["qs"])
    p = fdgpu.loc_params(**workloads.loc_params(g))
    want = ctx.locate(p, events)
    batch = fdgpu.LocBatch(events)
    for rep in range(2):
        batch.reset()
        ctx.locate_batch(p, batch)
        got = batch.results()
        for r, w in zip(got, want):
            for f in ("status", "ngood", "passes", "node", "exm", "eym", "ezm"):
                assert r[f] == w[f], (rep, f)
            assert _bits([r["stdmin"], r["avrmin"]], [w["stdmin"], w["avrmin"]]) == 0
            assert np.array_equal(r["isgood"], w["isgood"])
            assert _bits(r["res"], w["res"]) == 0 and _bits(r["tps"], w["tps"]) == 0
    for i in ids:
        ctx.release(i)
    ctx.release_model(m)
