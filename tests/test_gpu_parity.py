"""GPU parity: the gfx950 sweep kernel against the C oracle, bit for bit.

Every case compares the whole travel-time grid as float32 bit patterns
(integer-exact: the path is a deterministic FP64 computation stored in FP32,
and the kernel reproduces the reference's operation order).
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


def _bits_equal(a, b):
    return int((a.view(np.uint32) != b.view(np.uint32)).sum())


@pytest.mark.parametrize("seed", list(range(24)))
def test_kernel_matches_oracle_batched(ctx, seed):
    from fdtomoc_amd import fdgpu
    par, v = make_case(seed)
    t_ref, st = ob.solve(par, ob.slowness(v))
    m = ctx.model(v)
    (tid,) = ctx.solve([fdgpu.TableSpec(par_dict(par), m)], fdgpu.BATCHED)
    t = ctx.read(tid, v.shape)
    gst = ctx.stats(tid)
    ctx.release(tid)
    ctx.release_model(m)
    print(f"seed {seed} shape {v.shape} oracle {st} gpu {gst}", flush=True)
    assert gst["error"] & ~1 == 0
    assert gst["passes"] == st["passes"]
    assert gst["wave_passes"] == 0   # default: every pass face by face
    assert _bits_equal(t, t_ref) == 0


@pytest.mark.parametrize("wave", [1, 3])
def test_kernel_matches_oracle_wave_modes(ctx, wave):
    """The 24 seeded option mixes again with the first pass as one dataflow
    wavefront (wave 1) and with every pass as one (wave 3), all in one launch;
    the default (every pass face by face) runs in the test above."""
    from fdtomoc_amd import fdgpu
    cases = [make_case(seed) for seed in range(24)]
    models = [ctx.model(v) for _, v in cases]
    ctx.set_wave(wave)
    try:
        ids = ctx.solve([fdgpu.TableSpec(par_dict(par), m) for (par, _), m in zip(cases, models)], fdgpu.BATCHED)
    finally:
        ctx.set_wave(0)
    for (par, v), tid in zip(cases, ids):
        t_ref, st = ob.solve(par, ob.slowness(v))
        gst = ctx.stats(tid)
        t = ctx.read(tid, v.shape)
        ctx.release(tid)
        assert gst["error"] & ~1 == 0
        assert gst["passes"] == st["passes"]
        assert gst["wave_passes"] == (gst["passes"] if wave == 3 else 1)
        assert _bits_equal(t, t_ref) == 0
    for m in models:
        ctx.release_model(m)


def _golden_par(keys, vals):
    ints = {"nx", "ny", "nz", "iminus", "iplus", "jminus", "jplus", "kminus", "kplus", "floatsrc", "srctype",
            "srcwall", "xs", "ys", "zs", "ncube", "reverse", "headpref", "parse", "fill", "invh", "invq", "invf",
            "invx0", "invy0", "invz0"}
    return ob.Par(**{str(k): (int(v) if str(k) in ints else float(v)) for k, v in zip(keys, vals)})


def test_golden_fixtures_one_launch(ctx):
    """M001 (the reference's shipped verification table) and the 48 seeded
    cases the compiled reference solved, all in ONE batched launch, against
    the fixture times bit for bit."""
    import os
    from fdtomoc_amd import fdgpu
    gdir = os.path.join(os.path.dirname(__file__), "golden")
    z = np.load(os.path.join(gdir, "ref_cases.npz"))
    cases = []
    for s in sorted({k.split("_")[0] for k in z.files if k.startswith("s")}):
        cases.append((s, _golden_par(z["par_keys"], z[f"{s}_par"]), z[f"{s}_vel"], z[f"{s}_times"]))
    m1 = np.load(os.path.join(gdir, "m001.npz"))
    cases.append(("m001", _golden_par(m1["par_keys"], m1["par_vals"]), m1["vel"], m1["times"]))
    models = [ctx.model(v) for _, _, v, _ in cases]
    specs = [fdgpu.TableSpec(par_dict(p), m) for (_, p, _, _), m in zip(cases, models)]
    ids = ctx.solve(specs, fdgpu.BATCHED)
    bad = []
    for (name, _, v, t_ref), tid in zip(cases, ids):
        t = ctx.read(tid, v.shape)
        st = ctx.stats(tid)
        if st["error"] & ~1 or _bits_equal(t, t_ref):
            bad.append((name, st["error"], _bits_equal(t, t_ref)))
        ctx.release(tid)
    for m in models:
        ctx.release_model(m)
    print(f"{len(cases)} golden tables in one launch, mismatches {bad}", flush=True)
    assert not bad


@pytest.mark.parametrize("seed", [100, 101, 102, 103])
def test_plain_mode_matches_oracle(ctx, seed):
    """PLAIN launches one table at a time with the whole device (the drop-in
    per-station call) -- same bits as BATCHED and the oracle."""
    from fdtomoc_amd import fdgpu
    par, v = make_case(seed, 12, 40)
    t_ref, st = ob.solve(par, ob.slowness(v))
    m = ctx.model(v)
    (tid,) = ctx.solve([fdgpu.TableSpec(par_dict(par), m)], fdgpu.PLAIN)
    t = ctx.read(tid, v.shape)
    ctx.release(tid)
    ctx.release_model(m)
    assert _bits_equal(t, t_ref) == 0


def test_windows_of_one_resident_model(ctx):
    """Tables that cover different sub-blocks (origins) of one resident model
    equal the oracle run on the cut-out block."""
    from fdtomoc_amd import fdgpu, synth
    rng = np.random.default_rng(77)
    G = synth.centred_grid(48, 40, 24, 1.0)
    V = (synth.gradient_velocity(G, 4.5, 7.5, 24.0) * synth.gaussian_anomalies(G, rng, 4, 0.1, (3, 8))).astype(np.float32)
    m = ctx.model(V)
    specs, refs = [], []
    for w in range(6):
        nx, ny, nz = int(rng.integers(10, 30)), int(rng.integers(10, 28)), int(rng.integers(6, 20))
        ox, oy, oz = int(rng.integers(0, 48 - nx + 1)), int(rng.integers(0, 40 - ny + 1)), int(rng.integers(0, 24 - nz + 1))
        sub = np.ascontiguousarray(V[oz:oz + nz, oy:oy + ny, ox:ox + nx])
        g = synth.centred_grid(nx, ny, nz, 1.0, lat=float(rng.uniform(-40, 40)), lon=float(rng.uniform(-100, 100)))
        lon, lat = g.node_lonlat(rng.uniform(0, nx - 1), rng.uniform(0, ny - 1))
        par = ob.Par(nx=nx, ny=ny, nz=nz, h=1.0, x0=g.x0, y0=g.y0, z0=0.0, fxs=round(lon, 5), fys=round(lat, 5),
                     fzs=round(float(rng.uniform(0, nz - 1)), 3), reverse=50)
        specs.append(fdgpu.TableSpec(par_dict(par), m, (ox, oy, oz)))
        refs.append(ob.solve(par, ob.slowness(sub))[0])
    ids = ctx.solve(specs, fdgpu.BATCHED)
    for tid, t_ref in zip(ids, refs):
        t = ctx.read(tid, t_ref.shape)
        ctx.release(tid)
        assert _bits_equal(t, t_ref) == 0
    ctx.release_model(m)


def test_model_update_then_resolve(ctx):
    """A resident model updated in place (a tomography iteration) gives the
    oracle's tables for the new velocities."""
    from fdtomoc_amd import fdgpu
    par, v = make_case(200, 14, 30)
    m = ctx.model(v)
    for it in range(3):
        v2 = (v * np.float32(1.0 + 0.02 * it)).astype(np.float32)
        v2[: v.shape[0] // 2] *= np.float32(0.97)
        ctx.update_model(m, v2)
        (tid,) = ctx.solve([fdgpu.TableSpec(par_dict(par), m)], fdgpu.BATCHED)
        t = ctx.read(tid, v.shape)
        ctx.release(tid)
        assert _bits_equal(t, ob.solve(par, ob.slowness(v2))[0]) == 0, it
    ctx.release_model(m)


def test_error_paths(ctx):
    """Bad requests are refused with the documented codes, not solved."""
    from fdtomoc_amd import fdgpu
    par, v = make_case(3)
    m = ctx.model(v)
    d = par_dict(par)
    with pytest.raises(fdgpu.FdgpuError):     # window outside the model
        ctx.solve([fdgpu.TableSpec(d, m, (1, 0, 0))])
    far = dict(d, fxs=d["fxs"] + 50.0)
    with pytest.raises(fdgpu.FdgpuError, match="source"):
        ctx.solve([fdgpu.TableSpec(far, m)])
    with pytest.raises(fdgpu.FdgpuError):     # unknown model
        ctx.solve([fdgpu.TableSpec(d, m + 1000)])
    ctx.release_model(m)


@pytest.mark.parametrize("budget", [0, 1, 32])
def test_schedule_knob_does_not_change_results(ctx, budget):
    """Every per-face schedule choice gives the oracle's bits (budget 1 hands
    nearly every reverse-pass face to LEVEL, 0 never does)."""
    from fdtomoc_amd import fdgpu
    cases = [make_case(s) for s in range(300, 312)]
    models = [ctx.model(v) for _, v in cases]
    ctx.set_schedule(budget)
    try:
        ids = ctx.solve([fdgpu.TableSpec(par_dict(p), m) for (p, _), m in zip(cases, models)])
    finally:
        ctx.set_schedule(8)
    for (p, v), tid in zip(cases, ids):
        assert _bits_equal(ctx.read(tid, v.shape), ob.solve(p, ob.slowness(v))[0]) == 0
        ctx.release(tid)
    for m in models:
        ctx.release_model(m)


@pytest.mark.parametrize("lanes", [256, 512])
def test_workgroup_shapes_give_the_same_bits(ctx, lanes):
    """Both sweep builds (256 lanes, two tables per CU; 512 lanes, one per CU)
    reproduce the oracle."""
    from fdtomoc_amd import fdgpu
    cases = [make_case(s) for s in range(400, 416)]
    models = [ctx.model(v) for _, v in cases]
    ctx.set_workgroup(lanes)
    try:
        ids = ctx.solve([fdgpu.TableSpec(par_dict(p), m) for (p, _), m in zip(cases, models)])
    finally:
        ctx.set_workgroup(0)
    for (p, v), tid in zip(cases, ids):
        st = ctx.stats(tid)
        assert st["error"] & ~1 == 0
        assert _bits_equal(ctx.read(tid, v.shape), ob.solve(p, ob.slowness(v))[0]) == 0
        ctx.release(tid)
    for m in models:
        ctx.release_model(m)


def test_captured_launch_replays(ctx):
    """Batched solves replay one captured graph per launch shape: the same
    batch solved three times, with other work in between, gives the oracle's
    bits every time."""
    from fdtomoc_amd import fdgpu
    cases = [make_case(s) for s in range(500, 506)]
    refs = [ob.solve(p, ob.slowness(v))[0] for p, v in cases]
    models = [ctx.model(v) for _, v in cases]
    specs = [fdgpu.TableSpec(par_dict(p), m) for (p, _), m in zip(cases, models)]
    for rep in range(3):
        ids = ctx.solve(specs)
        for tid, t_ref in zip(ids, refs):
            assert _bits_equal(ctx.read(tid, t_ref.shape), t_ref) == 0, rep
            ctx.release(tid)
        # eager work of another shape between the replays
        (other,) = ctx.solve(specs[rep:rep + 1])
        assert _bits_equal(ctx.read(other, refs[rep].shape), refs[rep]) == 0
        ctx.release(other)
    for m in models:
        ctx.release_model(m)


def test_grid_beyond_geometry_lds_is_refused(ctx):
    """A grid whose geometry row/layer tables do not fit the wide build's LDS
    (7 (ny + 2) + 6 (nz + 2) > 6800 doubles, e.g. 1000 rows) is refused
    before any device work, with FDGPU_EINPUT."""
    from fdtomoc_amd import fdgpu
    nx, ny, nz = 3, 1000, 6
    v = np.full((nz, ny, nx), 6.0, dtype=np.float32)
    m = ctx.model(v)
    par = dict(nx=nx, ny=ny, nz=nz, h=1.0, x0=120.0, y0=24.0, z0=0.0, fxs=120.0, fys=23.99, fzs=0.0)
    with pytest.raises(fdgpu.FdgpuError) as e:
        ctx.solve([fdgpu.TableSpec(par, m)], fdgpu.BATCHED)
    assert e.value.code == -5
    ctx.release_model(m)
