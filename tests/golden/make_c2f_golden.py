"""Writes tests/golden/c2f_cases.npz: seeded coarse models interpolated to
fine models by the reference c2f compiled from its own sources
(oracle/_ref/c2f_ref, built by `make -C oracle ref` in a container that has
the reference at /root/reference).

Each case is a spec file (nxc nyc nzc h, igridx/igridy/igridz, x0 y0 z0) and a
coarse model file (P then S velocities, with the 232-byte header and the
igrid arrays, or headerless), run through c2f_ref; the fine model it writes
(header + P then S fine velocities) is stored.  Only data is stored: the
inputs and the output velocities.

    python tests/golden/make_c2f_golden.py
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, REPO)

from fdtomoc_amd import tables  # noqa: E402

C2F = os.path.join(REPO, "oracle", "_ref", "c2f_ref")
SEEDS = list(range(700, 716))


def make_case(seed: int) -> dict:
    """A random coarse grid and velocities; odd seeds are headerless files."""
    rng = np.random.default_rng(seed)
    nxc, nyc, nzc = (int(v) for v in rng.integers(2, 9, size=3))
    igx = rng.integers(1, 6, size=nxc - 1).astype(np.int32)
    igy = rng.integers(1, 6, size=nyc - 1).astype(np.int32)
    igz = rng.integers(1, 5, size=nzc - 1).astype(np.int32)
    # spacings and origins that are not exact binary fractions exercise the
    # float rounding of the node coordinates
    h = float(rng.choice([1.0, 0.5, 0.1, 2.0, 0.3, 1.7]))
    x0 = round(float(rng.uniform(-60, 60)), 3)
    y0 = round(float(rng.uniform(-60, 60)), 3)
    z0 = round(float(rng.uniform(-5, 5)), 2)
    nxyzc = nxc * nyc * nzc
    vp = rng.uniform(3.0, 8.5, size=nxyzc).astype(np.float32)
    vs = (vp / rng.uniform(1.6, 1.9, size=nxyzc)).astype(np.float32)
    vel = np.concatenate([vp, vs]).astype(np.float32)
    return dict(nxc=nxc, nyc=nyc, nzc=nzc, igridx=igx, igridy=igy, igridz=igz, h=h, x0=x0, y0=y0, z0=z0,
                vel=vel, headered=(seed % 2 == 0))


def spec_text(c: dict) -> str:
    def row(name, arr):
        # c2f reads nxc values from the igridx line (c2f.c:211-216); the last
        # is not used
        return name + " " + " ".join(str(int(v)) for v in list(arr) + [0]) + "\n"
    return (f"nxc {c['nxc']}\nnyc {c['nyc']}\nnzc {c['nzc']}\nh {c['h']!r}\n"
            f"oldvfil coarse.mod\ntgrdfil grid.txt\nfinevel fine.mod\n"
            f"x0 {c['x0']!r}\ny0 {c['y0']!r}\nz0 {c['z0']!r}\n"
            + row("igridx", c["igridx"]) + row("igridy", c["igridy"]) + row("igridz", c["igridz"]))


def run_c2f(c: dict) -> np.ndarray:
    """The fine velocities c2f_ref writes (P then S), as float32."""
    wd = tempfile.mkdtemp(prefix="c2fref_")
    try:
        with open(os.path.join(wd, "spec"), "w") as f:
            f.write(spec_text(c))
        with open(os.path.join(wd, "coarse.mod"), "wb") as f:
            if c["headered"]:
                head = tables.make_header(c["nxc"], c["nyc"], c["nzc"], c["x0"], c["y0"], c["z0"], c["h"], c["h"],
                                          c["h"], text=tables.header_text("CORS", "CART", "BMOD"))
                f.write(np.asarray(head, dtype=tables.VHEAD_DTYPE).tobytes())
                for k in ("igridx", "igridy", "igridz"):
                    f.write(np.asarray(c[k], dtype="<i4").tobytes())
            f.write(np.asarray(c["vel"], dtype="<f4").tobytes())
        r = subprocess.run([C2F], input="spec\n", cwd=wd, capture_output=True, text=True, timeout=600)
        if r.returncode != 0:
            raise RuntimeError(f"c2f_ref failed ({r.returncode}): {r.stdout[-1000:]} {r.stderr[-1000:]}")
        raw = np.fromfile(os.path.join(wd, "fine.mod"), dtype=np.uint8)
        head = raw[:232].view(tables.VHEAD_DTYPE)[0]
        n = int(head["nx"]) * int(head["ny"]) * int(head["nz"])
        fine = raw[232:232 + 8 * n].view("<f4").copy()
        assert fine.size == 2 * n
        return fine, (int(head["nx"]), int(head["ny"]), int(head["nz"]))
    finally:
        shutil.rmtree(wd, ignore_errors=True)


def main():
    if not os.path.exists(C2F):
        raise SystemExit("reference c2f not built: run `make -C oracle ref`")
    out = {}
    for s in SEEDS:
        c = make_case(s)
        fine, dims = run_c2f(c)
        for k, v in c.items():
            out[f"s{s}_{k}"] = np.asarray(v)
        out[f"s{s}_fine"] = fine
        out[f"s{s}_dims"] = np.asarray(dims, dtype=np.int32)
    out["seeds"] = np.asarray(SEEDS, dtype=np.int32)
    np.savez_compressed(os.path.join(HERE, "c2f_cases.npz"), **out)
    print(f"wrote {len(SEEDS)} cases")


if __name__ == "__main__":
    main()
