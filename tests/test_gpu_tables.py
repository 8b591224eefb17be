"""GPU tests of the table service beside the solver: background writes of
resident tables (fdgpu_tables_write_async) give files byte-identical to
fdgpu_table_write's (the reference layout, src/sphfd.c:5702-5803) while the
next solve runs; tables with writes in flight are not released; write
failures are reported by fdgpu_write_wait."""
import os

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


def _specs(m, g, seeds, **extra):
    from fdtomoc_amd import fdgpu, synth
    out = []
    for s in seeds:
        st = synth.surface_node_station(g, np.random.default_rng(s), margin=1)
        par = ob.Par(nx=g.nx, ny=g.ny, nz=g.nz, h=g.h, x0=g.x0, y0=g.y0, z0=g.z0, fxs=st.lon, fys=st.lat,
                     fzs=st.depth, reverse=50)
        d = par_dict(par)
        d.update(extra)
        out.append(fdgpu.TableSpec(d, m))
    return out


def test_background_writes_equal_sync_writes(ctx, tmp_path):
    """Eight tables (parse 1 and 2, swab 0 and 2, with and without a velocity
    header) written in the background while another batch solves: every file
    equals the synchronous write byte for byte."""
    from fdtomoc_amd import fdgpu, synth, tables
    g = synth.centred_grid(31, 29, 16, 1.0)
    v = synth.gradient_velocity(g, 5.0, 7.5, 16.0).astype(np.float32)
    m = ctx.model(v)
    head = np.asarray(tables.make_header(g.nx, g.ny, g.nz, g.x0, g.y0, g.z0, g.df_deg(), g.dq_deg(), g.h)).tobytes()
    specs = (_specs(m, g, range(300, 302)) + _specs(m, g, range(302, 304), parse=2)
             + _specs(m, g, range(304, 306), swab=2) + _specs(m, g, range(306, 308)))
    ids = ctx.solve(specs, fdgpu.BATCHED)
    heads = [head if k % 2 == 0 else None for k in range(len(ids))]
    (tmp_path / "a").mkdir()
    (tmp_path / "b").mkdir()
    for k, i in enumerate(ids):
        ctx.write(i, str(tmp_path / "a" / f"t{k}.times"), heads[k])
    ctx.write_async(ids, [str(tmp_path / "b" / f"t{k}.times") for k in range(len(ids))], heads)
    more = ctx.solve(_specs(m, g, range(400, 408)), fdgpu.BATCHED)   # overlaps the writes
    ctx.write_wait()
    for k in range(len(ids)):
        a = open(tmp_path / "a" / f"t{k}.times", "rb").read()
        b = open(tmp_path / "b" / f"t{k}.times", "rb").read()
        assert a == b, k
    for i in ids + more:
        ctx.release(i)
    ctx.release_model(m)


def test_release_waits_for_writes_and_errors_reported(ctx, tmp_path):
    """A table named by 40 queued writes is refused for release until the
    writes finish; a path in a missing directory makes fdgpu_write_wait
    return FDGPU_EIO, and the next wait is clean."""
    from fdtomoc_amd import fdgpu, synth
    g = synth.centred_grid(25, 23, 12, 1.0)
    v = synth.gradient_velocity(g, 5.0, 7.5, 12.0).astype(np.float32)
    m = ctx.model(v)
    ids = ctx.solve(_specs(m, g, [500, 501]), fdgpu.BATCHED)
    ctx.write_async([ids[0]] * 40, [str(tmp_path / f"r{k}.times") for k in range(40)])
    with pytest.raises(fdgpu.FdgpuError) as e:
        ctx.release(ids[0])
    assert e.value.code == -10
    ctx.write_wait()
    ctx.release(ids[0])
    assert all(os.path.getsize(tmp_path / f"r{k}.times") == 232 + 4 * g.n for k in range(40))
    ctx.write_async([ids[1]], [str(tmp_path / "missing" / "x.times")])
    with pytest.raises(fdgpu.FdgpuError) as e:
        ctx.write_wait()
    assert e.value.code == -6
    ctx.write_wait()
    ctx.release(ids[1])
    ctx.release_model(m)


def test_close_finishes_queued_writes(tmp_path):
    """Closing a context with background writes queued completes them first:
    every file exists in full afterwards."""
    from fdtomoc_amd import fdgpu, synth
    c = fdgpu.Context(0)
    g = synth.centred_grid(21, 19, 10, 1.0)
    m = c.model(synth.gradient_velocity(g, 5.0, 7.0, 10.0).astype(np.float32))
    ids = c.solve(_specs(m, g, [600, 601, 602]), fdgpu.BATCHED)
    c.write_async(ids * 10, [str(tmp_path / f"c{k}.times") for k in range(30)])
    c.close()
    assert all(os.path.getsize(tmp_path / f"c{k}.times") == 232 + 4 * g.n for k in range(30))
