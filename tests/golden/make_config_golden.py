"""Writes tests/golden/config_tables.npz: the oracle's results for the tables
that the full-size parity tests check (tests/config_cases.py), each also
solved by the reference compiled from its own sources (oracle/_ref/sphfd_ref,
`make -C oracle ref`), which must agree bit for bit before anything is
recorded.

Per table: the SHA-256 of the input velocities (float32, file order) and of
the output times, the pass count, and 512 sampled (node index, time) pairs.
Only data is stored.  Runs for about ten minutes on eight cores.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import config_cases as cc  # noqa: E402
import oracle_bind as ob  # noqa: E402


def solve_both(case):
    key, par, vel = case
    t0 = time.time()
    p = ob.Par(**par)
    with cf.ThreadPoolExecutor(max_workers=1) as ex:
        fut = ex.submit(ob.run_reference, p, vel)
        t_or, st = ob.solve(p, ob.slowness(vel))
        _, t_ref, _ = fut.result()
    t_ref = np.asarray(t_ref, dtype=np.float32).reshape(t_or.shape)
    nd = int((t_ref.view(np.uint32) != t_or.view(np.uint32)).sum())
    print(f"{key} {t_or.shape} passes {st['passes']} oracle vs reference: {nd} nodes differ "
          f"({time.time() - t0:.0f} s)", flush=True)
    if nd:
        raise RuntimeError(f"{key}: the oracle and the compiled reference differ at {nd} nodes")
    idx = cc.sample_index(t_or.size)
    return key, cc.sha(vel), cc.sha(t_or), st["passes"], idx, t_or.reshape(-1)[idx]


def main():
    cases = cc.all_cases()
    # the slowest first: the 321x321x161 tables take minutes each
    cases.sort(key=lambda c: -c[2].size)
    with cf.ThreadPoolExecutor(max_workers=int(os.environ.get("JOBS", "4"))) as ex:
        rows = list(ex.map(solve_both, cases))
    rows.sort(key=lambda r: r[0])
    np.savez_compressed(os.path.join(HERE, "config_tables.npz"),
                        keys=np.array([r[0] for r in rows]), vel_sha=np.array([r[1] for r in rows]),
                        out_sha=np.array([r[2] for r in rows]), passes=np.array([r[3] for r in rows], np.int32),
                        sample_idx=np.stack([r[4] for r in rows]), sample_val=np.stack([r[5] for r in rows]))


if __name__ == "__main__":
    main()
