"""GPU: resident tables against the launch slots' scratch under a tight HBM
budget.  The wavefront state of the launch slots (16 bytes per node per
slot) is only taken while it leaves half of the budget free, and is released
when a table or model would not fit otherwise, so the tables a caller keeps
resident are limited by the budget, not by scratch of earlier launches."""
import numpy as np
import pytest

import oracle_bind as ob

pytestmark = pytest.mark.gpu

BUDGET = 96 << 20


def test_resident_tables_reclaim_launch_scratch():
    from fdtomoc_amd import fdgpu, workloads
    g, v, st = workloads.c1(0)            # 65 x 65 x 33 (557,700 bytes per table)
    par = workloads.table_par(g, st)
    t_ref, ost = ob.solve(ob.Par(**par), ob.slowness(v))
    ctx = fdgpu.Context(0, BUDGET)
    try:
        m = ctx.model(v)
        ctx.set_wave(1)   # first passes as wavefronts: the slots hold wavefront state
        ids = ctx.solve([fdgpu.TableSpec(par, m)] * 8, fdgpu.BATCHED)
        first_wave = ctx.stats(ids[0])["wave_passes"]
        batch = 8
        while True:
            try:
                ids += ctx.solve([fdgpu.TableSpec(par, m)] * batch, fdgpu.BATCHED)
            except fdgpu.FdgpuError as e:
                assert "memory" in str(e), str(e)
                if batch == 1:
                    break
                batch = 1
        per_table = 4 * g.n
        print(f"{len(ids)} tables of {per_table} bytes resident in a {BUDGET}-byte budget; "
              f"first launch wave passes {first_wave}, last {ctx.stats(ids[-1])['wave_passes']}", flush=True)
        # the first launch ran with wavefront state; it was given up for tables
        assert first_wave == 1
        assert ctx.stats(ids[-1])["wave_passes"] == 0
        # what stays besides tables: the model and 8 slots' face scratch and
        # change stamps (about 0.8 MB each, room for about 167 tables); with
        # the wavefront state held (4.2 MB per slot) about 119 would fit
        assert len(ids) >= 150
        for tid in (ids[0], ids[-1]):
            st_ = ctx.stats(tid)
            assert st_["error"] & ~1 == 0 and st_["passes"] == ost["passes"]
            t = ctx.read(tid, v.shape)
            assert int((t.view(np.uint32) != t_ref.view(np.uint32)).sum()) == 0
        for tid in ids:
            ctx.release(tid)
        ctx.release_model(m)
    finally:
        ctx.close()
