"""Writes tests/golden/ray_cases.npz: sphrayderv's derivative outputs for
seeded ray cases (tests/ray_cases.py), produced by the reference program
compiled from its own sources (oracle/_ref/sphrayderv_ref, built by
`make -C oracle ref`) in a container that has them.

For each case the reference reads the spec, the station list, the event file
and the P and S tables of every station (solved by the sphfd oracle, which
matches the compiled reference sphfd bit for bit; their SHA-256 are stored
so that a test can regenerate and check them), and writes its files; stored
here per written ray: the weighted derivatives with their variables
(dtdsfil, float bits), the weighted residual (resfile, float bits), the four
hypocentre derivatives (dtdhfil, six decimals), and the residual, azimuth and
incidence the data file prints; and the program's output files themselves
(FILES, byte for byte), for the file-level comparison of
fdtomoc_amd.rayderv.run.  Only data is stored.
"""
from __future__ import annotations

import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import ray_cases as rc  # noqa: E402

REF_RAY = os.path.join(os.path.dirname(os.path.dirname(HERE)), "oracle", "_ref", "sphrayderv_ref")
SEEDS = (11, 12, 13)
FILES = ("dtds.bin", "res.bin", "dtdh.txt", "book.txt", "hit.txt", "scle.txt", "dot.txt", "head.txt")


def run_reference(case, tabs, wd):
    spec = rc.write_inputs(case, wd, tabs)
    p = subprocess.run([REF_RAY], input=spec + "\n", cwd=wd, capture_output=True, text=True, timeout=300)
    if p.returncode != 0:
        raise RuntimeError(f"reference failed ({p.returncode}): {p.stdout[-2000:]} {p.stderr[-2000:]}")
    out = rc.read_outputs(wd)
    out["files"] = {f: open(os.path.join(wd, f), "rb").read() for f in FILES}
    return out


def main():
    out = {}
    for seed in SEEDS:
        case = rc.ray_case(seed)
        tabs = rc.case_tables(case)
        with tempfile.TemporaryDirectory() as wd:
            r = run_reference(case, tabs, wd)
        pre = f"c{seed}_"
        keys = sorted(tabs)
        out[pre + "table_keys"] = np.array([f"{k[0]}.{k[1]}" for k in keys])
        out[pre + "table_sha"] = np.array([rc.table_hash(tabs[k]) for k in keys])
        good = r["good"]
        out[pre + "good"] = np.array([f"{g['event']}|{g['station']}|{g['phase']}" for g in good])
        out[pre + "resmin"] = np.array([g["resmin"] for g in good])
        out[pre + "az"] = np.array([g["az"] for g in good])
        out[pre + "ai"] = np.array([g["ai"] for g in good])
        out[pre + "res"] = r["res"]
        out[pre + "dtdh"] = r["dtdh"]
        ptr = np.cumsum([0] + [len(w[0]) for w in r["written"]]).astype(np.int64)
        out[pre + "du_ptr"] = ptr
        out[pre + "du_idx"] = np.concatenate([w[0] for w in r["written"]]).astype(np.int32)
        out[pre + "du_val"] = np.concatenate([w[1] for w in r["written"]]).astype(np.float32)
        out[pre + "book"] = r["book"]
        for f in FILES:
            out[pre + "file_" + f] = np.frombuffer(r["files"][f], dtype=np.uint8)
        print(f"seed {seed}: {len(case.events)} events, {sum(len(e[9]) for e in case.events)} picks, "
              f"{len(good)} good, {len(r['written'])} written, {ptr[-1]} derivatives", flush=True)
    np.savez_compressed(os.path.join(HERE, "ray_cases.npz"), **out)


if __name__ == "__main__":
    main()
