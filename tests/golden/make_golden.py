"""Writes the golden fixtures under tests/golden/ (run in a container that has
the reference sources at /root/reference; needs oracle/_ref/sphfd_ref, built
by `make -C oracle ref`).

  m001.npz      the reference's own verification vector
                (verification/cmod/tempvel P grid and verification/cmod/M001.ptimes),
                with the three CR bytes that the repository's `* text eol=lf`
                rule stripped from the binary table put back (the output of
                the compiled reference matches the repaired file byte for byte)
  ref_cases.npz seeded small cases (tests/cases.py) solved by the compiled
                reference: par values, velocities, output tables and headers

Only data is stored: velocity grids, tables and par values.
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import oracle_bind as ob  # noqa: E402
from cases import make_case, par_dict  # noqa: E402
from fdtomoc_amd import tables  # noqa: E402

REF = "/root/reference/verification/cmod"
# offsets (in the 232-byte-header-stripped payload of the reference's own
# output) where the repository copy lost a CR before an LF
CR_OFFSETS = (103421, 131353, 181824)
GOLDEN_SEEDS = list(range(500, 548))


def m001():
    raw = open(os.path.join(REF, "M001.ptimes"), "rb").read()
    head, pay = raw[:232], bytearray(raw[232:])
    for off in CR_OFFSETS:
        assert pay[off] == 0x0A, off
        pay[off:off] = b"\r"
    times = np.frombuffer(bytes(pay), dtype="<f4").reshape(31, 41, 41)
    vg = tables.read_grid(os.path.join(REF, "tempvel"), count=41 * 41 * 31)
    vel = np.array(vg.data, dtype="<f4").reshape(31, 41, 41)
    vhead = open(os.path.join(REF, "tempvel"), "rb").read()[:232]
    par = dict(nx=41, ny=41, nz=31, h=2.0, x0=120.9, y0=23.8, z0=-4.0, fxs=121.37915, fys=23.12651,
               fzs=-0.198, reverse=50)
    np.savez_compressed(os.path.join(HERE, "m001.npz"), vel=vel, times=times,
                        table_header=np.frombuffer(head, np.uint8), vel_header=np.frombuffer(vhead, np.uint8),
                        par_keys=np.array(list(par.keys())), par_vals=np.array(list(par.values()), dtype=np.float64))


def ref_cases():
    out = {}
    for s in GOLDEN_SEEDS:
        par, v = make_case(s, nmin=6, nmax=18)
        if not ob.geometry(par)["ok"]:
            continue
        head, t, _ = ob.run_reference(par, v)
        pd = par_dict(par)
        out[f"s{s}_vel"] = v
        out[f"s{s}_times"] = t
        out[f"s{s}_head"] = np.frombuffer(np.asarray(head).tobytes(), np.uint8)
        out[f"s{s}_par"] = np.array([float(pd[k]) for k in sorted(pd)], dtype=np.float64)
    out["par_keys"] = np.array(sorted(par_dict(make_case(0)[0])))
    np.savez_compressed(os.path.join(HERE, "ref_cases.npz"), **out)


if __name__ == "__main__":
    m001()
    ref_cases()
    for f in ("m001.npz", "ref_cases.npz"):
        print(f, os.path.getsize(os.path.join(HERE, f)))
