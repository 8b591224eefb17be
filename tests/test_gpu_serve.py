"""Serving: the device-resident request loop (fdgpu_serve_*).

Requests of different sizes arrive while others are running; each result is
bit-identical to the same table solved alone (oracle and a batched launch),
admission by the HBM arena refuses work that does not fit, and the loop
drains, idles out and relaunches correctly."""
import time

import numpy as np
import pytest

import oracle_bind as ob
from cases import par_dict

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    from fdtomoc_amd import fdgpu
    c = fdgpu.Context(0)
    yield c
    c.close()


def _bits(a, b):
    return int((a.view(np.uint32) != b.view(np.uint32)).sum())


def _world(seed=31):
    from fdtomoc_amd import synth
    rng = np.random.default_rng(seed)
    G = synth.centred_grid(56, 48, 28, 1.0)
    V = (synth.gradient_velocity(G, 4.5, 7.5, 24.0) * synth.gaussian_anomalies(G, rng, 4, 0.1, (3, 8))).astype(np.float32)
    return rng, G, V


def _window_request(rng, V, nmin=10, nmax=34):
    from fdtomoc_amd import synth
    NZ, NY, NX = V.shape
    nx, ny, nz = int(rng.integers(nmin, min(nmax, NX))), int(rng.integers(nmin, min(nmax, NY))), \
        int(rng.integers(6, min(22, NZ)))
    ox, oy, oz = int(rng.integers(0, NX - nx + 1)), int(rng.integers(0, NY - ny + 1)), int(rng.integers(0, NZ - nz + 1))
    g = synth.centred_grid(nx, ny, nz, 1.0, lat=float(rng.uniform(-40, 40)), lon=float(rng.uniform(-100, 100)))
    lon, lat = g.node_lonlat(rng.uniform(0, nx - 1), rng.uniform(0, ny - 1))
    par = ob.Par(nx=nx, ny=ny, nz=nz, h=1.0, x0=g.x0, y0=g.y0, z0=0.0, fxs=round(lon, 5), fys=round(lat, 5),
                 fzs=round(float(rng.uniform(0, nz - 1)), 3), reverse=int(rng.choice([1, 3, 50])))
    sub = np.ascontiguousarray(V[oz:oz + nz, oy:oy + ny, ox:ox + nx])
    return par, (ox, oy, oz), sub


def test_requests_in_the_mix_equal_alone(ctx):
    from fdtomoc_amd import fdgpu
    rng, G, V = _world()
    m = ctx.model(V)
    reqs = []
    for k in range(12):
        par, origin, sub = _window_request(rng, V)
        own = None
        if k % 3 == 2:
            # a request carrying its own (updated) velocities
            own = (sub * (1.0 + 0.03 * rng.standard_normal(sub.shape))).astype(np.float32)
            origin = (0, 0, 0)
        reqs.append((par, origin, sub if own is None else own, own))
    ctx.serve_start((56, 48, 28))
    try:
        ids = []
        for k, (par, origin, vel, own) in enumerate(reqs):
            ids.append(ctx.serve_submit(fdgpu.TableSpec(par_dict(par), m if own is None else -1, origin), own))
            if k in (1, 5):
                time.sleep(0.05)     # later requests arrive while earlier ones run
        info = ctx.serve_info()
        print("serve info after submits", info, flush=True)
        assert info["outstanding"] >= 1
        got = []
        for rid, (par, _, vel, _) in zip(ids, reqs):
            ctx.serve_wait(rid, 120.0)
            got.append((ctx.serve_read(rid, vel.shape), ctx.serve_stats(rid)))
            ctx.serve_release(rid)
    finally:
        ctx.serve_stop()
    # alone: the oracle, and each table in its own batched launch
    for (par, origin, vel, own), (t, st) in zip(reqs, got):
        t_ref, ost = ob.solve(par, ob.slowness(vel))
        assert st["error"] & ~1 == 0
        assert st["passes"] == ost["passes"]
        assert _bits(t, t_ref) == 0
    for (par, origin, vel, own), (t, _) in list(zip(reqs, got))[:4]:
        mm = ctx.model(vel) if own is not None else m
        (tid,) = ctx.solve([fdgpu.TableSpec(par_dict(par), mm, origin if own is None else (0, 0, 0))])
        alone = ctx.read(tid, vel.shape)
        ctx.release(tid)
        if own is not None:
            ctx.release_model(mm)
        assert _bits(t, alone) == 0
    ctx.release_model(m)


def test_admission_by_arena_and_busy_calls(ctx):
    from fdtomoc_amd import fdgpu
    rng, G, V = _world(32)
    m = ctx.model(V)
    par, origin, sub = _window_request(rng, V, 24, 25)
    n = par.nx * par.ny * par.nz
    # one request's arena block (time grid, sin table, counters, stamps, last
    # full-plane faces, geometry table): at least the time grid and stamps;
    # room for two and a half
    a = lambda b: (b + 255) // 256 * 256
    per = fdgpu.serve_request_bytes(nx=par.nx, ny=par.ny, nz=par.nz)
    assert per >= a(4 * n) + a(4 * n)
    assert fdgpu.serve_request_bytes(True, nx=par.nx, ny=par.ny, nz=par.nz) >= per + 4 * n
    ctx.serve_start((56, 48, 28), arena_bytes=a(int(2.5 * per)))
    try:
        spec = fdgpu.TableSpec(par_dict(par), m, origin)
        a = ctx.serve_submit(spec)
        b = ctx.serve_submit(spec)
        with pytest.raises(fdgpu.FdgpuError) as e:
            ctx.serve_submit(spec)
        assert e.value.code == fdgpu.EBUSY
        with pytest.raises(fdgpu.FdgpuError) as e:   # no device allocation while serving
            ctx.model(V)
        assert e.value.code == fdgpu.EBUSY
        ctx.serve_wait(a, 120.0)
        with pytest.raises(fdgpu.FdgpuError):          # too big for the configured maximum
            big = dict(par_dict(par), nx=60, ny=60, nz=30)
            ctx.serve_submit(fdgpu.TableSpec(big, m, (0, 0, 0)))
        ctx.serve_release(a)
        c = ctx.serve_submit(spec)                      # admitted once memory is back
        for r in (b, c):
            ctx.serve_wait(r, 120.0)
        t_ref = ob.solve(par, ob.slowness(sub))[0]
        for r in (b, c):
            assert _bits(ctx.serve_read(r, sub.shape), t_ref) == 0
    finally:
        ctx.serve_stop()
    ctx.release_model(m)


def test_idle_exit_relaunch_and_stop_drains(ctx):
    from fdtomoc_amd import fdgpu
    rng, G, V = _world(33)
    m = ctx.model(V)
    ctx.serve_start((56, 48, 28), workgroups=8)
    try:
        p1, o1, s1 = _window_request(rng, V, 10, 16)
        r1 = ctx.serve_submit(fdgpu.TableSpec(par_dict(p1), m, o1))
        ctx.serve_wait(r1, 120.0)
        time.sleep(0.8)                                 # the server goes idle and leaves
        p2, o2, s2 = _window_request(rng, V, 10, 16)
        r2 = ctx.serve_submit(fdgpu.TableSpec(par_dict(p2), m, o2))
        ctx.serve_wait(r2, 120.0)
        info = ctx.serve_info()
        print("serve info", info, flush=True)
        assert info["launches"] >= 2
        assert _bits(ctx.serve_read(r2, s2.shape), ob.solve(p2, ob.slowness(s2))[0]) == 0
        # queue more than the 8 workgroups take at once, then stop at once: stop
        # finishes every admitted request before it returns
        for _ in range(12):
            p, o, _s = _window_request(rng, V, 10, 16)
            ctx.serve_submit(fdgpu.TableSpec(par_dict(p), m, o))
    finally:
        ctx.serve_stop()
    assert ctx.serve_info()["on"] == 0
    # the context is usable for batched work afterwards
    (tid,) = ctx.solve([fdgpu.TableSpec(par_dict(p1), m, o1)])
    assert _bits(ctx.read(tid, s1.shape), ob.solve(p1, ob.slowness(s1))[0]) == 0
    ctx.release(tid)
    ctx.release_model(m)


def test_coarse_model_requests(ctx):
    """Requests that carry one phase of a coarse model (an inversion update)
    arrive while a fine-velocity request runs: the workgroup that takes each
    interpolates the slowness on the device; every result equals the oracle
    chain (c2f, then sphfd) and the same velocities submitted as fine
    velocities, bit for bit.  A coarse model whose fine grid is not the
    table's, or a phase other than 0/1, is refused."""
    from fdtomoc_amd import fdgpu, synth
    rng = np.random.default_rng(61)
    nxc, nyc, nzc = 6, 5, 4
    vp = rng.uniform(4.5, 7.5, size=nxc * nyc * nzc).astype(np.float32)
    vs = (vp / np.float32(1.73)).astype(np.float32)
    co = fdgpu.Coarse(nxc, nyc, nzc, np.full(nxc - 1, 6), np.array([7, 6, 6, 5]), np.array([5, 5, 4]),
                      0.0, 0.0, 0.0, 1.0, np.concatenate([vp, vs]))
    nx, ny, nz = co.dims()
    (_, _, _), fine = ob.c2f(nxc, nyc, nzc, co.igridx, co.igridy, co.igridz, 0.0, 0.0, 0.0, 1.0, co.velocity)
    fine = fine.reshape(2, nz, ny, nx)
    g = synth.centred_grid(nx, ny, nz, 1.0)
    st = synth.surface_node_station(g, rng, margin=2)
    par = ob.Par(nx=nx, ny=ny, nz=nz, h=1.0, x0=g.x0, y0=g.y0, z0=g.z0, fxs=st.lon, fys=st.lat, fzs=st.depth,
                 reverse=50)
    spec = fdgpu.TableSpec(par_dict(par), -1)
    ctx.serve_start((nx, ny, nz))
    try:
        rf = ctx.serve_submit(spec, fine[1])            # S as fine velocities
        rp = ctx.serve_submit_coarse(spec, co, 0)       # P and S as the coarse model
        rs = ctx.serve_submit_coarse(spec, co, 1)
        small = fdgpu.Coarse(nxc, nyc, nzc, np.full(nxc - 1, 5), co.igridy, co.igridz, 0.0, 0.0, 0.0, 1.0,
                             co.velocity)
        for bad, ph in ((small, 0), (co, 2)):
            with pytest.raises(fdgpu.FdgpuError) as e:
                ctx.serve_submit_coarse(spec, bad, ph)
            assert e.value.code == -5
        for r in (rf, rp, rs):
            ctx.serve_wait(r, 120.0)
        t = {r: ctx.serve_read(r, (nz, ny, nx)) for r in (rf, rp, rs)}
        for r in (rf, rp, rs):
            ctx.serve_release(r)
    finally:
        ctx.serve_stop()
    t_p = ob.solve(par, ob.slowness(fine[0]))[0]
    t_s = ob.solve(par, ob.slowness(fine[1]))[0]
    assert _bits(t[rp], t_p) == 0
    assert _bits(t[rs], t_s) == 0
    assert _bits(t[rf], t[rs]) == 0
