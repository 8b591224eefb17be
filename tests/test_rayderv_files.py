"""sphrayderv's program around the rays (fdtomoc_amd/rayderv.py `run`: its
input files, bookkeeping and output files) against the files the reference
program compiled from its sources wrote for the same inputs
(tests/golden/ray_cases.npz, tests/golden/make_ray_golden.py), byte for byte:
the binary derivative rows and residuals, the hypocentre partials, the
bookkeeping, hit and index files, and the data and header files.  The rays
here come from the oracle (equal to the reference's rays on these cases);
tests/test_gpu_rays.py runs the same program with the device's rays."""
from __future__ import annotations

import os

import numpy as np
import pytest

import oracle_bind as ob
import ray_cases as rc
from fdtomoc_amd import rayderv, tables

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "ray_cases.npz")
FILES = ("dtds.bin", "res.bin", "dtdh.txt", "book.txt", "hit.txt", "scle.txt", "dot.txt", "head.txt")


def table_files_trace(workdir: str, timedir: str = "ttimes"):
    """A trace function for rayderv.run: the tables read from the run's
    table files as the reference reads them, the rays by the oracle."""
    def trace(grid, keys, rays):
        tabs = {}
        for k in set(keys):
            path = os.path.join(workdir, timedir, f"{k[0]}.{'p' if k[1] == 'P' else 's'}times")
            tabs[k] = tables.read_grid(path).data
        return ob.ray_trace(grid, [tabs[k] for k in keys], rays)
    return trace


@pytest.mark.parametrize("seed", [11, 12, 13])
def test_program_files_equal_the_reference(tmp_path, seed):
    golden = np.load(GOLDEN)
    case = rc.ray_case(seed)
    tabs = rc.case_tables(case)
    spec = rc.write_inputs(case, str(tmp_path), tabs)
    info = rayderv.run(spec, table_files_trace(str(tmp_path)), workdir=str(tmp_path))
    assert info["events"] == len(case.events)
    for f in FILES:
        got = open(tmp_path / f, "rb").read()
        want = golden[f"c{seed}_file_{f}"].tobytes()
        assert got == want, f"{f} differs (seed {seed}): {len(got)} vs {len(want)} bytes"
