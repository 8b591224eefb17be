"""GPU parity of the coarse-to-fine model path (c2f's interpolation,
src/c2f.c:549-575 and 673-727, on the device): the fine velocities against
the reference c2f's own output (golden fixtures written by the compiled
reference) and against the oracle, bit for bit, and travel-time tables
solved on the resulting resident models against tables solved on models
made from the same velocities the usual way."""
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
    return int((np.asarray(a).view(np.uint32) != np.asarray(b).view(np.uint32)).sum())


def _coarse(c):
    from fdtomoc_amd import fdgpu
    return fdgpu.Coarse(c["nxc"], c["nyc"], c["nzc"], c["igridx"], c["igridy"], c["igridz"], c["x0"], c["y0"],
                        c["z0"], c["h"], c["vel"])


def test_fine_models_match_reference_c2f(ctx):
    """All 16 golden cases (headered and headerless coarse files): the fine
    velocities equal what the compiled reference c2f wrote."""
    for c in ob.c2f_golden_cases():
        print(f"c2f golden seed {c['seed']} coarse {c['nxc']}x{c['nyc']}x{c['nzc']}", flush=True)
        mp, ms, fine = ctx.models_from_coarse(_coarse(c), want_fine=True)
        assert fine.shape[1:] == tuple(int(v) for v in c["dims"][::-1])
        assert _bits(fine.ravel(), c["fine"]) == 0, c["seed"]
        ctx.release_model(mp)
        ctx.release_model(ms)


def _sphfd_coarse(seed, nxc=6, nyc=6, nzc=4, step=(6, 6, 5)):
    """A coarse model whose fine grid is a 31 x 31 x 16 sphfd grid."""
    from fdtomoc_amd import fdgpu
    rng = np.random.default_rng(seed)
    vp = rng.uniform(4.5, 7.5, size=nxc * nyc * nzc).astype(np.float32)
    vs = (vp / np.float32(1.73)).astype(np.float32)
    return fdgpu.Coarse(nxc, nyc, nzc, np.full(nxc - 1, step[0]), np.full(nyc - 1, step[1]),
                        np.full(nzc - 1, step[2]), 0.0, 0.0, 0.0, 1.0, np.concatenate([vp, vs]))


def _table_par(nx, ny, nz, seed):
    from fdtomoc_amd import synth
    g = synth.centred_grid(nx, ny, nz, 1.0)
    st = synth.surface_node_station(g, np.random.default_rng(seed), margin=2)
    return ob.Par(nx=g.nx, ny=g.ny, nz=g.nz, h=g.h, x0=g.x0, y0=g.y0, z0=g.z0, fxs=st.lon, fys=st.lat,
                  fzs=st.depth, reverse=50)


def test_tables_on_coarse_models(ctx):
    """Tables solved on the models made on the device equal the tables solved
    on models created from the fine velocities, and the oracle chain
    (oracle c2f, then oracle sphfd) on the same inputs."""
    from fdtomoc_amd import fdgpu
    co = _sphfd_coarse(41)
    (nx, ny, nz), fine_ref = ob.c2f(co.nxc, co.nyc, co.nzc, co.igridx, co.igridy, co.igridz, co.x0, co.y0, co.z0,
                                    co.h, co.velocity)
    mp, ms, fine = ctx.models_from_coarse(co, want_fine=True)
    assert _bits(fine.ravel(), fine_ref) == 0
    vp, vs = fine[0], fine[1]
    fp, fs = ctx.model(vp), ctx.model(vs)
    par = _table_par(nx, ny, nz, 5)
    specs = [fdgpu.TableSpec(par_dict(par), m) for m in (mp, ms, fp, fs)]
    ids = ctx.solve(specs, fdgpu.BATCHED)
    t = [ctx.read(i, vp.shape) for i in ids]
    assert _bits(t[0], t[2]) == 0 and _bits(t[1], t[3]) == 0
    assert _bits(t[0], ob.solve(par, ob.slowness(vp))[0]) == 0
    assert _bits(t[1], ob.solve(par, ob.slowness(vs))[0]) == 0
    for i in ids:
        ctx.release(i)
    for m in (mp, ms, fp, fs):
        ctx.release_model(m)


def test_update_from_coarse_and_errors(ctx):
    """An inversion step: the models are replaced in place from an updated
    coarse model (same fine grid); results equal fresh models.  Mismatched
    fine dimensions and one model given twice are refused."""
    from fdtomoc_amd import fdgpu
    a, b = _sphfd_coarse(42), _sphfd_coarse(43)
    mp, ms = ctx.models_from_coarse(a)
    fine_b = ctx.update_models_from_coarse(mp, ms, b, want_fine=True)
    _, ref_b = ob.c2f(b.nxc, b.nyc, b.nzc, b.igridx, b.igridy, b.igridz, b.x0, b.y0, b.z0, b.h, b.velocity)
    assert _bits(fine_b.ravel(), ref_b) == 0
    np2, ns2 = ctx.models_from_coarse(b)
    nx, ny, nz = b.dims()
    par = _table_par(nx, ny, nz, 9)
    ids = ctx.solve([fdgpu.TableSpec(par_dict(par), m) for m in (mp, ms, np2, ns2)], fdgpu.BATCHED)
    shape = (nz, ny, nx)
    t = [ctx.read(i, shape) for i in ids]
    assert _bits(t[0], t[2]) == 0 and _bits(t[1], t[3]) == 0
    small = _sphfd_coarse(44, step=(5, 6, 5))
    with pytest.raises(fdgpu.FdgpuError) as e:
        ctx.update_models_from_coarse(mp, ms, small)
    assert e.value.code == -5
    with pytest.raises(fdgpu.FdgpuError) as e:
        ctx.update_models_from_coarse(mp, mp, b)
    assert e.value.code == -5
    for i in ids:
        ctx.release(i)
    for m in (mp, ms, np2, ns2):
        ctx.release_model(m)
