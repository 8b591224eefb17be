"""ctypes binding of the host-thread executor of the sweep solver
(tests/native/libemu_solver.so).  Test infrastructure only."""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

import oracle_bind as ob

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EMU_SO = os.path.join(REPO, "tests", "native", "libemu_solver.so")
STAT_NAMES = ["passes", "restarts", "faces_level", "faces_change", "rounds_level", "rounds_change",
              "fallbacks", "evals", "defers", "error", "last_wall"] + [f"headw{i}" for i in range(7)] + \
    ["skips", "t_gather", "t_sort", "t_level", "t_change", "t_init", "t_level_pre", "t_level_eval", "walks", "t_change1",
     "wave_passes", "rounds_wave", "wave_seeds", "wave_carry", "wave_scans", "t_wave",
     "t_wave_group", "t_wave_eval", "t_wave_carry", "team", "wave_fail"]

_lib = None


def lib():
    global _lib
    if _lib is None:
        L = C.CDLL(EMU_SO)
        L.emu_solve.argtypes = [C.POINTER(ob._Par), C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p,
                                C.c_int, C.c_int]
        L.emu_solve.restype = C.c_int
        _lib = L
    return _lib


def solve(par: "ob.Par", slow: np.ndarray, nthreads: int = 4, extra=None, change_budget: int = 8,
          global_state: bool = False, wave: int = 0, team: int = 1, missing_member: bool = False):
    """global_state: keep every face's schedule state (LEVEL pending counters
    and successor masks, CHANGE queue marks) in global scratch, the device's
    path for faces too large for LDS.  wave: bit0 the first pass as one
    dataflow wavefront, bit1 the reverse passes too.  team: workgroups (of
    nthreads threads each) sharing the wavefront passes; missing_member: the
    team's last member never starts."""
    slow = np.ascontiguousarray(slow, dtype=np.float32)
    time = np.empty_like(slow)
    stats = np.zeros(len(STAT_NAMES), dtype=np.int32)
    ex = None if extra is None else np.ascontiguousarray(extra, dtype=np.float32)
    rc = lib().emu_solve(C.byref(par.to_c()), slow.ctypes.data, time.ctypes.data,
                         None if ex is None else ex.ctypes.data, nthreads, stats.ctypes.data, change_budget,
                         (1 if global_state else 0) | (wave << 1) | (team << 3) | (256 if missing_member else 0))
    if rc != 0:
        raise RuntimeError(f"emu_solve failed: {rc}")
    return time, dict(zip(STAT_NAMES, stats.tolist()))
