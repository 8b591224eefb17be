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
    assert _bits_equal(t, t_ref) == 0
