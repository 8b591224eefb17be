"""The C-ABI library loads and exports every function include/*.h declares
(no GPU work: symbol resolution only)."""
import ctypes as C
import os
import re

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(REPO, "fdtomoc_amd", "libfdtomoc_amd.so")


def declared_functions():
    names = []
    inc = os.path.join(REPO, "include")
    for fn in sorted(os.listdir(inc)):
        if not fn.endswith(".h"):
            continue
        text = open(os.path.join(inc, fn)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?[A-Za-z_][\w\s\*]*?\b([A-Za-z_]\w*)\s*\([^;{]*\)\s*;", text, re.M):
            if m.group(1) not in ("if", "while", "return"):
                names.append(m.group(1))
    return names


def test_header_parse_finds_entry_points():
    names = declared_functions()
    for n in ("sphfd", "fdgpu_open", "fdgpu_solve", "fdgpu_table_write", "fdgpu_run_spec"):
        assert n in names


@pytest.mark.skipif(not os.path.exists(LIB), reason="library not built")
def test_library_exports_every_declared_symbol():
    lib = C.CDLL(LIB)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert missing == []


@pytest.mark.skipif(not os.path.exists(LIB), reason="library not built")
def test_defaults_and_error_strings_without_gpu():
    from fdtomoc_amd import fdgpu
    p = fdgpu.make_par(nx=3, ny=4, nz=5, h=1.0)
    # src/sphfd.c:266-367 defaults
    assert (p.y0, p.floatsrc, p.srctype, p.ncube, p.reverse, p.headpref, p.parse) == (90.0, 1, 1, 2, 1, 6, 1)
    assert p.maxoff == -1.0 and p.kminus == 1 and p.iplus == 1
    assert fdgpu.lib().fdgpu_strerror(-3).decode().startswith("source")


@pytest.mark.skipif(not os.path.exists(LIB), reason="library not built")
def test_coarse_dims_and_validation_without_gpu():
    """fdgpu_coarse_dims (host only): 1 + sum(igrid) per axis as c2f computes the
    fine grid (src/c2f.c:262-277); bad coarse models are refused."""
    import numpy as np
    from fdtomoc_amd import fdgpu
    vel = np.ones(2 * 4 * 3 * 2, dtype=np.float32)
    c = fdgpu.Coarse(4, 3, 2, np.array([2, 3, 1]), np.array([4, 1]), np.array([5]), 0.0, 0.0, 0.0, 1.0, vel)
    assert c.dims() == (7, 6, 6)
    for bad in (dict(nzc=1, igridz=np.array([], dtype=np.int32)), dict(igridx=np.array([2, 0, 1])), dict(h=0.0)):
        kw = dict(c.__dict__)
        kw.update(bad)
        with pytest.raises(fdgpu.FdgpuError) as e:
            fdgpu.Coarse(**kw).dims()
        assert e.value.code == -5


@pytest.mark.skipif(not os.path.exists(LIB), reason="library not built")
def test_serve_request_bytes_without_gpu():
    """fdgpu_serve_request_bytes (host only): a request that carries its own
    velocities takes at least one more grid of floats; one that carries a
    coarse model takes that, plus its coarse slownesses, coordinates and axis
    entries; a coarse model whose fine grid is not the table's is refused."""
    import numpy as np
    from fdtomoc_amd import fdgpu
    vel = np.ones(2 * 4 * 3 * 2, dtype=np.float32)
    c = fdgpu.Coarse(4, 3, 2, np.array([2, 3, 1]), np.array([4, 1]), np.array([5]), 0.0, 0.0, 0.0, 1.0, vel)
    nx, ny, nz = c.dims()
    base = fdgpu.serve_request_bytes(nx=nx, ny=ny, nz=nz)
    own = fdgpu.serve_request_bytes(True, nx=nx, ny=ny, nz=nz)
    co = fdgpu.serve_request_bytes(coarse=c, nx=nx, ny=ny, nz=nz)
    assert own >= base + 4 * nx * ny * nz
    assert co >= own + 4 * 4 * 3 * 2 + 4 * (4 + 3 + 2) + 8 * (nx + ny + nz)
    with pytest.raises(fdgpu.FdgpuError) as e:
        fdgpu.serve_request_bytes(coarse=c, nx=nx + 1, ny=ny, nz=nz)
    assert e.value.code == -5


@pytest.mark.skipif(not os.path.exists(LIB), reason="library not built")
def test_loc_params_defaults_without_gpu():
    """fdgpu_loc_params_defaults (host only): sphfdloc's parameter defaults
    (src/sphfdloc.c:111-128), grid fields zero."""
    import numpy as np
    from fdtomoc_amd import fdgpu
    p = fdgpu.loc_params()
    assert (p.kmin, p.ndiv, p.ndiv2, p.ivs, p.nthres) == (2, 20, 20, 1, 8)
    assert p.vpvs == np.float32(1.78) and p.resthres == np.float32(0.5)
    assert p.resthrep == np.float32(5.0) and p.stdmax == np.float32(15.0)
    assert (p.nx, p.ny, p.nz, p.x0, p.h) == (0, 0, 0, 0.0, 0.0)


@pytest.mark.skipif(not os.path.exists(LIB), reason="library not built")
def test_loc_geography_without_gpu():
    """fdgpu_loc_params_from_grid and fdgpu_loc_geographic (host only): the
    grid set-up equals the independent Python set-up of the same formulas
    (sphfdloc.c:382-393) bit for bit; the hypocentre conversion inverts the
    forward one (glathinv(glath(lat, depth)) = (lat, depth) to 1e-9 degrees
    and 1e-6 km over latitudes -85..85 and depths -3..200 km); at the equator
    the radius is the WGS84 semi-major axis less the depth."""
    import numpy as np
    from fdtomoc_amd import fdgpu, synth, workloads
    g = synth.centred_grid(41, 37, 21, 1.0)
    p, z0r = fdgpu.loc_params_from_grid(g.nx, g.ny, g.nz, g.x0, g.y0, g.z0, g.h)
    ref = workloads.loc_params(g)
    for f in ("x0", "y0", "z0", "h", "dq", "df"):
        assert getattr(p, f) == ref[f], f
    assert (p.nx, p.ny, p.nz, p.kmin, p.ndiv) == (41, 37, 21, 2, 20)
    for lat in np.linspace(-85.0, 85.0, 35):
        for dep in (-3.0, 0.0, 10.0, 55.5, 200.0):
            q, r = fdgpu.loc_params_from_grid(5, 5, 5, 121.25, float(lat), dep, 1.0)
            la, lo, d = fdgpu.loc_geographic(q.x0, q.y0, dep, dep, r)
            assert abs(la - lat) < 1e-9 and abs(d - dep) < 1e-6 and abs(lo - 121.25) < 1e-12, (lat, dep)
    _, r = fdgpu.loc_params_from_grid(5, 5, 5, 0.0, 0.0, 12.0, 1.0)
    assert abs(r - (6378.137 - 12.0)) < 1e-9


def test_ctypes_structs_match_the_header(tmp_path):
    """Every ctypes mirror in fdtomoc_amd/fdgpu.py has the size and the field
    offsets of its struct in include/fdtomoc_amd.h, as a C compiler lays it
    out (a drifted mirror would corrupt calls silently)."""
    import shutil
    import subprocess
    from fdtomoc_amd import fdgpu
    if not shutil.which("gcc"):
        pytest.skip("no C compiler")
    pairs = [("fdgpu_par", fdgpu.CPar), ("fdgpu_serve_config", fdgpu.CServeConfig),
             ("fdgpu_table_desc", fdgpu.CTable), ("fdgpu_coarse", fdgpu.CCoarse),
             ("fdgpu_loc_params", fdgpu.CLocParams), ("fdgpu_loc_event", fdgpu.CLocEvent)]
    lines = ["#include <stdio.h>", "#include <stddef.h>", '#include "fdtomoc_amd.h"', "int main(void) {"]
    for cname, cls in pairs:
        lines.append(f'  printf("{cname} size %zu\\n", sizeof({cname}));')
        for fname, _ in cls._fields_:
            lines.append(f'  printf("{cname} {fname} %zu\\n", offsetof({cname}, {fname}));')
    lines += ["  return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    inc = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include")
    subprocess.run(["gcc", "-std=c99", "-I", inc, "-o", str(exe), str(src)], check=True)
    got = {}
    for line in subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.splitlines():
        cname, field, val = line.split()
        got[(cname, field)] = int(val)
    for cname, cls in pairs:
        assert got[(cname, "size")] == _ctypes_sizeof(cls), cname
        for fname, _ in cls._fields_:
            assert got[(cname, fname)] == getattr(cls, fname).offset, (cname, fname)


def _ctypes_sizeof(cls):
    import ctypes
    return ctypes.sizeof(cls)
