"""The device chain of one location step (c2f -> sphfd -> sphfdloc, as
run.sh chains the programs) against the oracle chain on the same inputs,
bit for bit: the coarse model interpolated (oracle_c2f), each station's P
and S tables solved on the fine velocities (oracle_solve), and the events
located over those tables (oracle_locate)."""
import numpy as np
import pytest

import oracle_bind as ob
from cases import par_dict

pytestmark = pytest.mark.gpu


def test_location_step_equals_the_oracle_chain():
    from fdtomoc_amd import fdgpu, pipeline, synth
    rng = np.random.default_rng(31)
    nxc, nyc, nzc = 7, 6, 4
    vp = rng.uniform(4.8, 7.4, size=nxc * nyc * nzc).astype(np.float32)
    vs = (vp / np.float32(1.74)).astype(np.float32)
    co = fdgpu.Coarse(nxc, nyc, nzc, np.full(nxc - 1, 5), np.full(nyc - 1, 5), np.array([4, 5, 5]), 0.0, 0.0, 0.0,
                      1.0, np.concatenate([vp, vs]))
    nx, ny, nz = co.dims()
    g = synth.centred_grid(nx, ny, nz, 1.0)
    stations = synth.random_surface_stations(g, rng, 9)
    pars = [ob.Par(nx=nx, ny=ny, nz=nz, h=1.0, x0=g.x0, y0=g.y0, z0=g.z0, fxs=st.lon, fys=st.lat, fzs=st.depth,
                   reverse=50) for st in stations]
    # the oracle chain: fine velocities, then P and S tables
    _, fine = ob.c2f(nxc, nyc, nzc, co.igridx, co.igridy, co.igridz, 0.0, 0.0, 0.0, 1.0, co.velocity)
    fine = fine.reshape(2, nz, ny, nx)
    tabs = [(ob.solve(p, ob.slowness(fine[0]))[0], ob.solve(p, ob.slowness(fine[1]))[0]) for p in pars]
    loc, _ = fdgpu.loc_params_from_grid(nx, ny, nz, g.x0, g.y0, g.z0, 1.0)
    loc.ndiv, loc.ndiv2, loc.nthres = 8, 6, 6
    events = []
    for e in range(10):
        i, j, k = int(rng.integers(2, nx - 2)), int(rng.integers(2, ny - 2)), int(rng.integers(3, nz - 2))
        picks = []
        for s in rng.choice(9, size=int(rng.integers(7, 10)), replace=False):
            for ph, q in (("P", 0), ("S", 1)):
                if rng.random() < 0.8:
                    t = float(tabs[int(s)][q][k, j, i]) + 2.0 + rng.normal(0, 0.03)
                    picks.append(pipeline.Pick(int(s), ph, t, float(rng.choice([1.0, 4.0, 25.0]))))
        events.append(picks)
    ctx = fdgpu.Context(0)
    try:
        got, _ = pipeline.locate_with_coarse_model(ctx, co, [par_dict(p) for p in pars], events, loc)
    finally:
        ctx.close()
    grid = ob.loc_grid(**{f: getattr(loc, f) for f, _ in ob.LocGrid._fields_})
    for picks, r in zip(events, got):
        o = ob.locate(grid, [tabs[k.station][0 if k.phase == "P" else 1] for k in picks],
                      np.array([k.obstime for k in picks], np.float32), np.array([k.pwt for k in picks], np.float32),
                      "".join(k.phase for k in picks), np.array([k.good for k in picks], np.int32))
        for f in ("status", "ngood", "passes", "node", "exm", "eym", "ezm"):
            assert r[f] == o[f], f
        assert np.float32(r["stdmin"]).view(np.uint32) == np.float32(o["stdmin"]).view(np.uint32)
        assert np.array_equal(r["res"].view(np.uint32), o["res"].view(np.uint32))
