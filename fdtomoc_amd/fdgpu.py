"""Python binding of the sphfd GPU library (include/fdtomoc_amd.h).

The compute path is the gfx950 HIP kernel in libfdtomoc_amd.so; this module
only marshals arguments.  It raises if the library is not built or no GPU is
present -- there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import dataclasses
import os

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("FDTOMOC_LIB") or os.path.join(PKG, "libfdtomoc_amd.so")   # env: another build
EXE_PATH = os.path.join(PKG, "bin", "sphfd")

STAT_NAMES = ["passes", "restarts", "faces_level", "faces_change", "rounds_level", "rounds_change",
              "fallbacks", "evals", "defers", "error", "last_wall"] + [f"headw{i}" for i in range(7)] + \
    ["skips", "t_gather", "t_sort", "t_level", "t_change", "t_init", "t_level_pre", "t_level_eval", "walks", "t_change1",
     "wave_passes", "rounds_wave", "wave_seeds", "wave_carry", "wave_scans", "t_wave",
     "t_wave_group", "t_wave_eval", "t_wave_carry", "team", "wave_fail"]

BATCHED, PLAIN = 0, 1


class CPar(C.Structure):
    """fdgpu_par: par-file values of one table (src/sphfd.c:383-530 keys)."""
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
                ("h", C.c_double), ("dq", C.c_double), ("df", C.c_double),
                ("x0", C.c_double), ("y0", C.c_double), ("z0", C.c_double),
                ("iminus", C.c_int), ("iplus", C.c_int), ("jminus", C.c_int),
                ("jplus", C.c_int), ("kminus", C.c_int), ("kplus", C.c_int),
                ("floatsrc", C.c_int), ("srctype", C.c_int), ("srcwall", C.c_int),
                ("fxs", C.c_double), ("fys", C.c_double), ("fzs", C.c_double),
                ("xs", C.c_int), ("ys", C.c_int), ("zs", C.c_int),
                ("ncube", C.c_int), ("reverse", C.c_int), ("headpref", C.c_int),
                ("parse", C.c_int), ("fill", C.c_int), ("maxoff", C.c_float),
                ("invh", C.c_int), ("invq", C.c_int), ("invf", C.c_int),
                ("invx0", C.c_int), ("invy0", C.c_int), ("invz0", C.c_int), ("swab", C.c_int)]


class CServeConfig(C.Structure):
    """fdgpu_serve_config."""
    _fields_ = [("max_nx", C.c_int), ("max_ny", C.c_int), ("max_nz", C.c_int),
                ("arena_bytes", C.c_ulonglong), ("staging_bytes", C.c_ulonglong),
                ("queue_capacity", C.c_int), ("workgroups", C.c_int)]


class CTable(C.Structure):
    _fields_ = [("par", CPar), ("model", C.c_int), ("ox", C.c_int), ("oy", C.c_int), ("oz", C.c_int),
                ("preset", C.c_void_p), ("boxvals", C.c_void_p)]


class CCoarse(C.Structure):
    """fdgpu_coarse: a coarse model as c2f reads it (src/c2f.c)."""
    _fields_ = [("nxc", C.c_int), ("nyc", C.c_int), ("nzc", C.c_int),
                ("igridx", C.c_void_p), ("igridy", C.c_void_p), ("igridz", C.c_void_p),
                ("x0", C.c_double), ("y0", C.c_double), ("z0", C.c_double), ("h", C.c_double),
                ("velocity", C.c_void_p)]


class CLocParams(C.Structure):
    """fdgpu_loc_params: sphfdloc's search grid and parameters."""
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int), ("x0", C.c_double), ("y0", C.c_double),
                ("z0", C.c_double), ("df", C.c_double), ("dq", C.c_double), ("h", C.c_double), ("kmin", C.c_int),
                ("ndiv", C.c_int), ("ndiv2", C.c_int), ("ivs", C.c_int), ("vpvs", C.c_float), ("nthres", C.c_int),
                ("resthres", C.c_float), ("resthrep", C.c_float), ("stdmax", C.c_float)]


class CLocEvent(C.Structure):
    """fdgpu_loc_event: one event's picks and, on return, its location."""
    _fields_ = [("npick", C.c_int), ("table", C.c_void_p), ("obstime", C.c_void_p), ("pwt", C.c_void_p),
                ("phase", C.c_void_p), ("isgood", C.c_void_p), ("res", C.c_void_p), ("tcalc", C.c_void_p),
                ("status", C.c_int), ("ngood", C.c_int), ("passes", C.c_int), ("node", C.c_int * 3),
                ("stdmin", C.c_float), ("avrmin", C.c_float), ("exm", C.c_double), ("eym", C.c_double),
                ("ezm", C.c_double)]


class CRayGrid(C.Structure):
    """fdgpu_ray_grid: the fine and coarse grids of sphrayderv's rays."""
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int), ("x0", C.c_double), ("y0", C.c_double),
                ("z0", C.c_double), ("df", C.c_double), ("dq", C.c_double), ("h", C.c_double), ("nxc", C.c_int),
                ("nyc", C.c_int), ("nzc", C.c_int), ("gx", C.c_void_p), ("gy", C.c_void_p), ("gz", C.c_void_p),
                ("ido1d", C.c_int), ("max_segments", C.c_int)]


class CRay(C.Structure):
    _fields_ = [("table", C.c_int), ("ex", C.c_double), ("ey", C.c_double), ("ez", C.c_double),
                ("xs", C.c_double), ("ys", C.c_double), ("zs", C.c_double)]


class CRayResult(C.Structure):
    _fields_ = [("status", C.c_int), ("nseg", C.c_int), ("ndu", C.c_int), ("first_in_cube", C.c_int),
                ("tev", C.c_double), ("dtdx", C.c_float), ("dtdy", C.c_float), ("dtdz", C.c_float),
                ("ai", C.c_float), ("az", C.c_float), ("rdev", C.c_float), ("offset", C.c_longlong)]


RAY_FIELDS = [name for name, _ in CRayResult._fields_ if name != "offset"]


def _np_dtype(ct) -> np.dtype:
    """numpy view of a ctypes structure's layout."""
    names = [n for n, _ in ct._fields_]
    fmt = {C.c_int: "<i4", C.c_double: "<f8", C.c_float: "<f4", C.c_longlong: "<i8"}
    return np.dtype({"names": names, "formats": [fmt[t] for _, t in ct._fields_],
                     "offsets": [getattr(ct, n).offset for n in names], "itemsize": C.sizeof(ct)})


RAY_DTYPE = _np_dtype(CRay)                 # one fdgpu_ray
RAY_RESULT_DTYPE = _np_dtype(CRayResult)    # one fdgpu_ray_result


def ray_array(tables, coords) -> np.ndarray:
    """fdgpu_ray records: tables (n ids) and coords (n x 6: ex, ey, ez, xs,
    ys, zs in the grid's units)."""
    coords = np.asarray(coords, dtype=np.float64).reshape(-1, 6)
    a = np.zeros(len(coords), dtype=RAY_DTYPE)
    a["table"] = np.asarray(tables, dtype=np.int32)
    for k, f in enumerate(("ex", "ey", "ez", "xs", "ys", "zs")):
        a[f] = coords[:, k]
    return a


def loc_params(**kw) -> CLocParams:
    """fdgpu_loc_params with sphfdloc's defaults (src/sphfdloc.c:111-128), then kw."""
    p = CLocParams()
    lib().fdgpu_loc_params_defaults(C.byref(p))
    for k, v in kw.items():
        setattr(p, k, v)
    return p


class PreparedCoarse:
    """A coarse model's fdgpu_coarse record and arrays, built once; change
    `velocity` in place between updates (same shape)."""

    def __init__(self, coarse: "Coarse"):
        self._keep = []
        self.c = coarse.to_c(self._keep)
        self.velocity = self._keep[-1]   # the float32 payload the record points at


class LocBatch:
    """Events as fdgpu_loc_event records with their arrays, built once, for
    repeated fdgpu_locate calls (Context.locate_batch); reset() restores the
    use flags the outlier passes change."""

    def __init__(self, events: list):
        self.n = len(events)
        self.arr = (CLocEvent * self.n)()
        self._keep, self._outs, self._good0 = [], [], []
        for e, ev in zip(self.arr, events):
            n = len(ev["table"])
            tab = np.ascontiguousarray(ev["table"], dtype=np.int32)
            obs = np.ascontiguousarray(ev["obstime"], dtype=np.float32)
            pwt = np.ascontiguousarray(ev["pwt"], dtype=np.float32)
            ph = np.frombuffer(ev["phase"].encode(), dtype=np.uint8).copy()
            good = np.ascontiguousarray(ev.get("isgood", np.ones(n)), dtype=np.int32).copy()
            res, tc = np.zeros(n, dtype=np.float32), np.zeros(n, dtype=np.float32)
            if not (len(obs) == len(pwt) == len(ph) == len(good) == n):
                raise ValueError("pick arrays of different lengths")
            self._keep += [tab, obs, pwt, ph]
            e.npick = n
            e.table, e.obstime, e.pwt = tab.ctypes.data, obs.ctypes.data, pwt.ctypes.data
            e.phase, e.isgood, e.res, e.tcalc = ph.ctypes.data, good.ctypes.data, res.ctypes.data, tc.ctypes.data
            self._outs.append((good, res, tc))
            self._good0.append(good.copy())

    def reset(self):
        for (good, _, _), g0 in zip(self._outs, self._good0):
            good[:] = g0

    def results(self) -> list:
        out = []
        for e, (good, res, tc) in zip(self.arr, self._outs):
            out.append(dict(status=e.status, ngood=e.ngood, passes=e.passes, node=tuple(e.node), stdmin=e.stdmin,
                            avrmin=e.avrmin, exm=e.exm, eym=e.eym, ezm=e.ezm, isgood=good.copy(), res=res.copy(),
                            tps=tc.copy()))
        return out


def loc_params_from_grid(nx: int, ny: int, nz: int, x0: float, y0: float, z0: float, h: float,
                         dq: float = 0.0, df: float = 0.0):
    """(fdgpu_loc_params, z0r): sphfdloc's search grid from the spec's grid
    (degrees, geodetic latitude, km), as sphfdloc.c:382-393 sets it up."""
    p, r = CLocParams(), C.c_double()
    _check(lib().fdgpu_loc_params_from_grid(nx, ny, nz, x0, y0, z0, h, dq, df, C.byref(p), C.byref(r)),
           "fdgpu_loc_params_from_grid")
    return p, r.value


def loc_geographic(exm: float, eym: float, ezm: float, z0: float, z0r: float):
    """(latitude deg, longitude deg, depth km) of a hypocentre as sphfdloc prints it."""
    la, lo, d = C.c_double(), C.c_double(), C.c_double()
    _check(lib().fdgpu_loc_geographic(exm, eym, ezm, z0, z0r, C.byref(la), C.byref(lo), C.byref(d)),
           "fdgpu_loc_geographic")
    return la.value, lo.value, d.value


@dataclasses.dataclass
class Coarse:
    """A coarse model: node counts, fine intervals per coarse interval, the
    spec's origin and fine spacing, and 2*nxc*nyc*nzc velocities (P then S)."""
    nxc: int
    nyc: int
    nzc: int
    igridx: np.ndarray
    igridy: np.ndarray
    igridz: np.ndarray
    x0: float
    y0: float
    z0: float
    h: float
    velocity: np.ndarray

    def to_c(self, keep: list) -> CCoarse:
        c = CCoarse()
        c.nxc, c.nyc, c.nzc = self.nxc, self.nyc, self.nzc
        for name in ("igridx", "igridy", "igridz"):
            a = np.ascontiguousarray(getattr(self, name), dtype=np.int32)
            keep.append(a)
            setattr(c, name, a.ctypes.data)
        v = np.ascontiguousarray(self.velocity, dtype=np.float32)
        keep.append(v)
        c.velocity = v.ctypes.data
        c.x0, c.y0, c.z0, c.h = self.x0, self.y0, self.z0, self.h
        return c

    def dims(self) -> tuple:
        keep = []
        c = self.to_c(keep)
        nx, ny, nz = C.c_int(), C.c_int(), C.c_int()
        _check(lib().fdgpu_coarse_dims(C.byref(c), C.byref(nx), C.byref(ny), C.byref(nz)), "fdgpu_coarse_dims")
        return nx.value, ny.value, nz.value


@dataclasses.dataclass
class TableSpec:
    """One station table: par values plus where it sits in a device model."""
    par: dict
    model: int
    origin: tuple = (0, 0, 0)
    preset: np.ndarray | None = None
    boxvals: np.ndarray | None = None


class FdgpuError(RuntimeError):
    def __init__(self, msg: str, code: int = 0):
        super().__init__(msg)
        self.code = code
        self.ids = None   # EKERNEL from a solve: the tables stay registered (their stats name the error)


EKERNEL, EBUSY, ETIMEOUT, ESCRATCH, ESPACE = -9, -10, -11, -12, -13
SERVE_INFO = ["on", "outstanding", "arena_used", "arena_bytes", "launches", "workgroups"]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise FdgpuError(f"{LIB_PATH} is not built; run `python -m fdtomoc_amd.build`")
        L = C.CDLL(LIB_PATH)
        L.fdgpu_par_defaults.argtypes = [C.POINTER(CPar)]
        L.fdgpu_par_defaults.restype = None
        L.fdgpu_open.argtypes = [C.c_int, C.c_ulonglong, C.POINTER(C.c_void_p)]
        L.fdgpu_close.argtypes = [C.c_void_p]
        L.fdgpu_strerror.argtypes = [C.c_int]
        L.fdgpu_strerror.restype = C.c_char_p
        L.fdgpu_model_create.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_void_p, C.POINTER(C.c_int)]
        L.fdgpu_model_update.argtypes = [C.c_void_p, C.c_int, C.c_void_p]
        L.fdgpu_model_release.argtypes = [C.c_void_p, C.c_int]
        L.fdgpu_solve.argtypes = [C.c_void_p, C.POINTER(CTable), C.c_int, C.c_int, C.POINTER(C.c_int)]
        L.fdgpu_resolve.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_int]
        L.fdgpu_solve_noalloc.argtypes = [C.c_void_p, C.POINTER(CTable), C.c_void_p, C.c_void_p, C.c_int]
        L.fdgpu_last_kernel_ms.argtypes = [C.c_void_p]
        L.fdgpu_last_kernel_ms.restype = C.c_double
        L.fdgpu_table_read.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_ulonglong]
        L.fdgpu_table_write.argtypes = [C.c_void_p, C.c_int, C.c_char_p, C.c_void_p]
        L.fdgpu_table_stats.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_int]
        L.fdgpu_table_release.argtypes = [C.c_void_p, C.c_int]
        L.fdgpu_table_device_ptr.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_void_p)]
        L.fdgpu_run_spec.argtypes = [C.c_char_p]
        L.fdgpu_serve_request_bytes.argtypes = [C.POINTER(CPar), C.c_int, C.c_void_p, C.POINTER(C.c_ulonglong)]
        L.fdgpu_coarse_dims.argtypes = [C.POINTER(CCoarse), C.POINTER(C.c_int), C.POINTER(C.c_int),
                                        C.POINTER(C.c_int)]
        L.fdgpu_models_from_coarse.argtypes = [C.c_void_p, C.POINTER(CCoarse), C.POINTER(C.c_int),
                                               C.POINTER(C.c_int), C.c_void_p]
        L.fdgpu_models_update_from_coarse.argtypes = [C.c_void_p, C.c_int, C.c_int, C.POINTER(CCoarse),
                                                      C.c_void_p]
        L.fdgpu_serve_submit_coarse.argtypes = [C.c_void_p, C.POINTER(CTable), C.POINTER(CCoarse), C.c_int,
                                                C.POINTER(C.c_int)]
        L.fdgpu_loc_params_defaults.argtypes = [C.POINTER(CLocParams)]
        L.fdgpu_loc_params_defaults.restype = None
        L.fdgpu_locate.argtypes = [C.c_void_p, C.POINTER(CLocParams), C.POINTER(CLocEvent), C.c_int]
        L.fdgpu_locate_timing.argtypes = [C.c_void_p, C.POINTER(C.c_double), C.c_int]
        L.fdgpu_loc_params_from_grid.argtypes = [C.c_int, C.c_int, C.c_int] + [C.c_double] * 6 + [
            C.POINTER(CLocParams), C.POINTER(C.c_double)]
        L.fdgpu_loc_geographic.argtypes = [C.c_double] * 5 + [C.POINTER(C.c_double)] * 3
        L.fdgpu_table_load.argtypes = [C.c_void_p, C.c_char_p, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_int)]
        L.fdgpu_tables_write_async.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int]
        L.fdgpu_write_wait.argtypes = [C.c_void_p]
        L.fdgpu_set_schedule.argtypes = [C.c_void_p, C.c_int]
        L.fdgpu_get_schedule.argtypes = [C.c_void_p, C.POINTER(C.c_int)]
        L.fdgpu_set_workgroup.argtypes = [C.c_void_p, C.c_int]
        L.fdgpu_set_wave.argtypes = [C.c_void_p, C.c_int]
        L.fdgpu_get_wave.argtypes = [C.c_void_p, C.POINTER(C.c_int)]
        L.fdgpu_set_team.argtypes = [C.c_void_p, C.c_int, C.c_int]
        L.fdgpu_get_team.argtypes = [C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_int)]
        L.fdgpu_serve_config_defaults.argtypes = [C.POINTER(CServeConfig)]
        L.fdgpu_serve_config_defaults.restype = None
        L.fdgpu_serve_start.argtypes = [C.c_void_p, C.POINTER(CServeConfig)]
        L.fdgpu_serve_submit.argtypes = [C.c_void_p, C.POINTER(CTable), C.c_void_p, C.POINTER(C.c_int)]
        L.fdgpu_serve_poll.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_int)]
        L.fdgpu_serve_wait.argtypes = [C.c_void_p, C.c_int, C.c_double]
        L.fdgpu_serve_read.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_ulonglong]
        L.fdgpu_serve_stats.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_int]
        L.fdgpu_serve_write.argtypes = [C.c_void_p, C.c_int, C.c_char_p, C.c_void_p]
        L.fdgpu_serve_release.argtypes = [C.c_void_p, C.c_int]
        L.fdgpu_serve_info.argtypes = [C.c_void_p, C.c_void_p, C.c_int]
        L.fdgpu_serve_stop.argtypes = [C.c_void_p]
        L.fdgpu_context_counters.argtypes = [C.c_void_p, C.c_void_p, C.c_int]
        L.fdgpu_rays_trace.argtypes = [C.c_void_p, C.POINTER(CRayGrid), C.POINTER(CRay), C.c_int,
                                       C.POINTER(CRayResult), C.c_void_p, C.c_void_p, C.c_longlong]
        L.sphfd.argtypes = [C.c_int, C.POINTER(C.c_char_p), C.c_char_p]
        L.fdgpu_sphfd_info.argtypes = [C.c_void_p, C.c_int]
        _lib = L
    return _lib


def _check(rc: int, what: str):
    if rc != 0:
        raise FdgpuError(f"{what}: {lib().fdgpu_strerror(rc).decode()} ({rc})", rc)


def make_par(**kw) -> CPar:
    p = CPar()
    lib().fdgpu_par_defaults(C.byref(p))
    for k, v in kw.items():
        if not hasattr(p, k):
            raise KeyError(k)
        setattr(p, k, v)
    return p


def serve_request_bytes(own_velocity: bool = False, coarse: "Coarse | None" = None, **par) -> int:
    """HBM arena bytes one serving request of these sizes takes (carrying its
    own velocities, or the coarse model `coarse`)."""
    b = C.c_ulonglong()
    keep = []
    cc = C.byref(coarse.to_c(keep)) if coarse is not None else None
    _check(lib().fdgpu_serve_request_bytes(C.byref(make_par(**par)), int(own_velocity), cc, C.byref(b)),
           "fdgpu_serve_request_bytes")
    return int(b.value)


class Context:
    """A device context: resident models and tables on one GPU."""

    def __init__(self, device: int = 0, hbm_budget: int = 0):
        self._h = C.c_void_p()
        _check(lib().fdgpu_open(device, hbm_budget, C.byref(self._h)), "fdgpu_open")

    def close(self):
        if self._h:
            lib().fdgpu_close(self._h)
            self._h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # models
    def model(self, velocity: np.ndarray) -> int:
        v = np.ascontiguousarray(velocity, dtype=np.float32)
        nz, ny, nx = v.shape
        mid = C.c_int()
        _check(lib().fdgpu_model_create(self._h, nx, ny, nz, v.ctypes.data, C.byref(mid)), "fdgpu_model_create")
        return mid.value

    def update_model(self, model: int, velocity: np.ndarray):
        v = np.ascontiguousarray(velocity, dtype=np.float32)
        _check(lib().fdgpu_model_update(self._h, model, v.ctypes.data), "fdgpu_model_update")

    def release_model(self, model: int):
        _check(lib().fdgpu_model_release(self._h, model), "fdgpu_model_release")

    def locate(self, params: CLocParams, events: list) -> list:
        """Locates events over resident tables (sphfdloc's grid search).  Each
        event is a dict with `table` (ids), `obstime`, `pwt`, `phase` (a str of
        'P'/'S') and optionally `isgood`; returns one dict per event with the
        status, counts, node, deviation, mean residual, hypocentre, and the
        final use flags, residuals and computed times per pick."""
        batch = LocBatch(events)
        self.locate_batch(params, batch)
        return batch.results()

    def locate_batch(self, params: CLocParams, batch: "LocBatch"):
        """fdgpu_locate on a prepared batch (its events' C records built once):
        the results are in the batch afterwards (LocBatch.results)."""
        _check(lib().fdgpu_locate(self._h, C.byref(params), batch.arr, batch.n), "fdgpu_locate")

    def trace_rays(self, grid, rays, du_cap: int | None = None, out=None):
        """sphrayderv's rays through resident tables (fdgpu_rays_trace).
        grid: a fdtomoc_amd.rayderv.RayGrid; rays: a RAY_DTYPE array
        (ray_array) or (table, ex, ey, ez, xs, ys, zs) tuples.  Returns
        (results, du_index, du_value): a RAY_RESULT_DTYPE array (status,
        segments, derivative count and offset, event time and derivatives,
        angles) and the packed derivatives, each ray's in index order, rays
        in input order.  du_cap: size the derivative arrays to it (an
        FDGPU_ESPACE error when they are too small) instead of growing them.
        out: (int32 array, float32 array) of equal length to receive the
        derivatives (reused by repeated calls; grown if too small), in place
        of new arrays; the returned ones are views of them."""
        if not isinstance(rays, np.ndarray):
            rays = ray_array([r[0] for r in rays], [r[1:7] for r in rays])
        rays = np.ascontiguousarray(rays, dtype=RAY_DTYPE)
        gx, gy, gz = (np.ascontiguousarray(a, dtype=np.float32) for a in (grid.gx, grid.gy, grid.gz))
        g = CRayGrid(grid.nx, grid.ny, grid.nz, grid.x0, grid.y0, grid.z0, grid.df, grid.dq, grid.h, len(gx), len(gy),
                     len(gz), gx.ctypes.data, gy.ctypes.data, gz.ctypes.data, grid.ido1d, grid.max_segments)
        n = len(rays)
        res = np.zeros(max(n, 1), dtype=RAY_RESULT_DTYPE)
        if out is not None:
            idx, val = out
            if not (isinstance(idx, np.ndarray) and idx.dtype == np.int32 and idx.flags.c_contiguous and
                    isinstance(val, np.ndarray) and val.dtype == np.float32 and val.flags.c_contiguous and
                    len(idx) == len(val)):
                raise ValueError("out: contiguous int32 and float32 arrays of one length")
            cap = len(idx) if du_cap is None else min(du_cap, len(idx))
        else:
            cap = du_cap if du_cap is not None else 128 * max(n, 1)
            idx = val = None
        while True:
            if idx is None or len(idx) < max(cap, 1):
                idx = np.empty(max(cap, 1), dtype=np.int32)
                val = np.empty(max(cap, 1), dtype=np.float32)
            rc = lib().fdgpu_rays_trace(self._h, C.byref(g), rays.ctypes.data_as(C.POINTER(CRay)), n,
                                        res.ctypes.data_as(C.POINTER(CRayResult)), idx.ctypes.data,
                                        val.ctypes.data, cap)
            if rc != ESPACE or du_cap is not None:
                break
            cap = int(res["ndu"][:n].sum())
        _check(rc, "fdgpu_rays_trace")
        total = int(res["ndu"][:n].sum())
        return res[:n], idx[:total], val[:total]

    def write_async(self, tables: list, paths: list, vel_headers: list | None = None):
        """Background writes of resident tables (fdgpu_tables_write_async);
        fdgpu_write_wait (write_wait) reports their outcome."""
        n = len(tables)
        ids = np.ascontiguousarray(tables, dtype=np.int32)
        enc = [C.create_string_buffer(os.fsencode(p)) for p in paths]
        pp = (C.c_char_p * n)(*[C.cast(e, C.c_char_p) for e in enc])
        hb, hp = [], None
        if vel_headers is not None:
            hb = [None if h is None else C.create_string_buffer(bytes(h), 232) for h in vel_headers]
            hp = (C.c_void_p * n)(*[None if h is None else C.cast(h, C.c_void_p) for h in hb])
        _check(lib().fdgpu_tables_write_async(self._h, ids.ctypes.data, pp, hp, n), "fdgpu_tables_write_async")

    def write_wait(self):
        _check(lib().fdgpu_write_wait(self._h), "fdgpu_write_wait")

    def load_table(self, path: str, shape: tuple) -> int:
        """A table file (headered or headerless, as FDtomoC's consumers read
        it) loaded into a new resident table; shape is (nz, ny, nx)."""
        nz, ny, nx = shape
        tid = C.c_int()
        _check(lib().fdgpu_table_load(self._h, os.fsencode(path), nx, ny, nz, C.byref(tid)), "fdgpu_table_load")
        return tid.value

    def locate_timing(self) -> dict:
        """Measurements of the last locate: passes, pass and node-search device
        milliseconds, and the table bytes the node searches read."""
        t = (C.c_double * 4)()
        _check(lib().fdgpu_locate_timing(self._h, t, 4), "fdgpu_locate_timing")
        return dict(passes=int(t[0]), pass_ms=t[1], node_ms=t[2], node_bytes=t[3])

    def models_from_coarse(self, coarse: Coarse, want_fine: bool = False):
        """(P model, S model[, fine velocities (2, nz, ny, nx)]) of a coarse
        model, interpolated on the device as c2f does."""
        keep = []
        c = coarse.to_c(keep)
        nx, ny, nz = coarse.dims()
        fine = np.empty((2, nz, ny, nx), dtype=np.float32) if want_fine else None
        mp, ms = C.c_int(), C.c_int()
        _check(lib().fdgpu_models_from_coarse(self._h, C.byref(c), C.byref(mp), C.byref(ms),
                                              None if fine is None else fine.ctypes.data),
               "fdgpu_models_from_coarse")
        return (mp.value, ms.value, fine) if want_fine else (mp.value, ms.value)

    def update_models_from_prepared(self, model_p: int, model_s: int, prepared: "PreparedCoarse"):
        """fdgpu_models_update_from_coarse on a coarse model whose C record was
        built once (repeated updates, as an inversion's iterations make)."""
        _check(lib().fdgpu_models_update_from_coarse(self._h, model_p, model_s, C.byref(prepared.c), None),
               "fdgpu_models_update_from_coarse")

    def update_models_from_coarse(self, model_p: int, model_s: int, coarse: Coarse, want_fine: bool = False):
        keep = []
        c = coarse.to_c(keep)
        nx, ny, nz = coarse.dims()
        fine = np.empty((2, nz, ny, nx), dtype=np.float32) if want_fine else None
        _check(lib().fdgpu_models_update_from_coarse(self._h, model_p, model_s, C.byref(c),
                                                     None if fine is None else fine.ctypes.data),
               "fdgpu_models_update_from_coarse")
        return fine

    # tables
    @staticmethod
    def _fill(dst: CTable, spec: TableSpec, keep: list):
        dst.par = make_par(**spec.par)
        dst.model = spec.model
        dst.ox, dst.oy, dst.oz = spec.origin
        for name in ("preset", "boxvals"):
            a = getattr(spec, name)
            if a is not None:
                a = np.ascontiguousarray(a, dtype=np.float32)
                keep.append(a)
                setattr(dst, name, a.ctypes.data)

    def solve(self, tables: list[TableSpec], mode: int = BATCHED) -> list[int]:
        n = len(tables)
        arr = (CTable * max(n, 1))()
        keep = []
        for t, spec in enumerate(tables):
            self._fill(arr[t], spec, keep)
        ids = (C.c_int * max(n, 1))()
        rc = lib().fdgpu_solve(self._h, arr, n, mode, ids)
        if rc == EKERNEL:
            e = FdgpuError(f"fdgpu_solve: {lib().fdgpu_strerror(rc).decode()} ({rc})", rc)
            e.ids = list(ids[:n])
            raise e
        _check(rc, "fdgpu_solve")
        return list(ids[:n])

    def resolve(self, tables: list[int], mode: int = BATCHED):
        """fdgpu_resolve: re-solve resident tables in place against their
        models' current velocities."""
        ids = np.ascontiguousarray(tables, dtype=np.int32)
        _check(lib().fdgpu_resolve(self._h, ids.ctypes.data, len(ids), mode), "fdgpu_resolve")

    def solve_noalloc(self, spec: TableSpec, d_time: int) -> dict:
        """fdgpu_solve_noalloc: one table into the caller's device buffer at
        address d_time (par nx*ny*nz floats), no device allocation; returns
        the table's counters."""
        d = CTable()
        keep = []
        self._fill(d, spec, keep)
        s = np.zeros(len(STAT_NAMES), dtype=np.int32)
        _check(lib().fdgpu_solve_noalloc(self._h, C.byref(d), C.c_void_p(d_time), s.ctypes.data, len(STAT_NAMES)),
               "fdgpu_solve_noalloc")
        return dict(zip(STAT_NAMES, s.tolist()))

    def set_schedule(self, change_budget: int):
        """Per-face schedule knob (fdgpu_set_schedule); results do not depend on it."""
        _check(lib().fdgpu_set_schedule(self._h, int(change_budget)), "fdgpu_set_schedule")

    def set_workgroup(self, lanes: int):
        """Workgroup shape of batched launches: 512 (one table per CU), 256 (two
        per CU) or 0 (automatic); results do not depend on it."""
        _check(lib().fdgpu_set_workgroup(self._h, int(lanes)), "fdgpu_set_workgroup")

    def set_wave(self, flags: int):
        """Wavefront passes (fdgpu_set_wave): bit0 the first pass, bit1 reverse
        passes; results do not depend on it."""
        _check(lib().fdgpu_set_wave(self._h, int(flags)), "fdgpu_set_wave")

    def get_wave(self) -> int:
        b = C.c_int()
        _check(lib().fdgpu_get_wave(self._h, C.byref(b)), "fdgpu_get_wave")
        return b.value

    def set_team(self, size: int, share_pct: int = 0):
        """Teams (fdgpu_set_team): the wavefront passes of the largest batched
        tables shared by several workgroups -- size 0: planned per table
        (share_pct: the planner's load %, 0 for 100); 2..16: that size for
        tables of at least share_pct % (0 for 200) of the mean nodes per
        workgroup; 1: never.  Results do not depend on it."""
        _check(lib().fdgpu_set_team(self._h, int(size), int(share_pct)), "fdgpu_set_team")

    def get_team(self) -> tuple[int, int]:
        a, b = C.c_int(), C.c_int()
        _check(lib().fdgpu_get_team(self._h, C.byref(a), C.byref(b)), "fdgpu_get_team")
        return a.value, b.value

    def get_schedule(self) -> int:
        b = C.c_int()
        _check(lib().fdgpu_get_schedule(self._h, C.byref(b)), "fdgpu_get_schedule")
        return b.value

    def resolved_again(self) -> int:
        """Tables solved a second time after a failed kernel consistency check."""
        c = (C.c_longlong * 2)()
        _check(lib().fdgpu_context_counters(self._h, c, 2), "fdgpu_context_counters")
        return int(c[0])

    def wave_again(self) -> int:
        """Tables solved a second time, face by face, after a wavefront pass
        that could not finish."""
        c = (C.c_longlong * 2)()
        _check(lib().fdgpu_context_counters(self._h, c, 2), "fdgpu_context_counters")
        return int(c[1])

    def last_kernel_ms(self) -> float:
        return float(lib().fdgpu_last_kernel_ms(self._h))

    def read(self, table: int, shape: tuple) -> np.ndarray:
        out = np.empty(shape, dtype=np.float32)
        _check(lib().fdgpu_table_read(self._h, table, out.ctypes.data, out.size), "fdgpu_table_read")
        return out

    def stats(self, table: int) -> dict:
        s = np.zeros(len(STAT_NAMES), dtype=np.int32)
        _check(lib().fdgpu_table_stats(self._h, table, s.ctypes.data, len(STAT_NAMES)), "fdgpu_table_stats")
        return dict(zip(STAT_NAMES, s.tolist()))

    def write(self, table: int, path: str, vel_header: bytes | None):
        hb = None if vel_header is None else C.create_string_buffer(bytes(vel_header), 232)
        _check(lib().fdgpu_table_write(self._h, table, path.encode(), hb), "fdgpu_table_write")

    def release(self, table: int):
        _check(lib().fdgpu_table_release(self._h, table), "fdgpu_table_release")

    # serving (continuous batching)
    def serve_start(self, max_dims: tuple, arena_bytes: int = 0, staging_bytes: int = 0,
                    queue_capacity: int = 0, workgroups: int = 0):
        cfg = CServeConfig()
        lib().fdgpu_serve_config_defaults(C.byref(cfg))
        cfg.max_nx, cfg.max_ny, cfg.max_nz = max_dims
        if arena_bytes:
            cfg.arena_bytes = arena_bytes
        if staging_bytes:
            cfg.staging_bytes = staging_bytes
        if queue_capacity:
            cfg.queue_capacity = queue_capacity
        cfg.workgroups = workgroups
        _check(lib().fdgpu_serve_start(self._h, C.byref(cfg)), "fdgpu_serve_start")

    def serve_submit(self, spec: TableSpec, velocity: np.ndarray | None = None) -> int:
        """Admits one request; raises FdgpuError with code EBUSY when memory or
        the queue is full."""
        d = CTable()
        keep = []
        self._fill(d, spec, keep)
        v = None if velocity is None else np.ascontiguousarray(velocity, dtype=np.float32)
        rid = C.c_int()
        _check(lib().fdgpu_serve_submit(self._h, C.byref(d), None if v is None else v.ctypes.data, C.byref(rid)),
               "fdgpu_serve_submit")
        return rid.value

    def serve_submit_coarse(self, spec: TableSpec, coarse: Coarse, phase: int) -> int:
        """Admits one request whose velocities are phase `phase` (0 P, 1 S) of a
        coarse model, interpolated on the device by the workgroup that takes it."""
        d = CTable()
        keep = []
        self._fill(d, spec, keep)
        c = coarse.to_c(keep)
        rid = C.c_int()
        _check(lib().fdgpu_serve_submit_coarse(self._h, C.byref(d), C.byref(c), int(phase), C.byref(rid)),
               "fdgpu_serve_submit_coarse")
        return rid.value

    def serve_poll(self, req: int) -> bool:
        done = C.c_int()
        _check(lib().fdgpu_serve_poll(self._h, req, C.byref(done)), "fdgpu_serve_poll")
        return bool(done.value)

    def serve_wait(self, req: int, timeout_s: float = -1.0):
        _check(lib().fdgpu_serve_wait(self._h, req, timeout_s), "fdgpu_serve_wait")

    def serve_read(self, req: int, shape: tuple) -> np.ndarray:
        out = np.empty(shape, dtype=np.float32)
        _check(lib().fdgpu_serve_read(self._h, req, out.ctypes.data, out.size), "fdgpu_serve_read")
        return out

    def serve_stats(self, req: int) -> dict:
        s = np.zeros(len(STAT_NAMES), dtype=np.int32)
        _check(lib().fdgpu_serve_stats(self._h, req, s.ctypes.data, len(STAT_NAMES)), "fdgpu_serve_stats")
        return dict(zip(STAT_NAMES, s.tolist()))

    def serve_write(self, req: int, path: str, vel_header: bytes | None):
        hb = None if vel_header is None else C.create_string_buffer(bytes(vel_header), 232)
        _check(lib().fdgpu_serve_write(self._h, req, path.encode(), hb), "fdgpu_serve_write")

    def serve_release(self, req: int):
        _check(lib().fdgpu_serve_release(self._h, req), "fdgpu_serve_release")

    def serve_info(self) -> dict:
        v = np.zeros(len(SERVE_INFO), dtype=np.uint64)
        _check(lib().fdgpu_serve_info(self._h, v.ctypes.data, len(SERVE_INFO)), "fdgpu_serve_info")
        return dict(zip(SERVE_INFO, [int(x) for x in v]))

    def serve_stop(self):
        _check(lib().fdgpu_serve_stop(self._h), "fdgpu_serve_stop")

    def device_ptr(self, table: int) -> int:
        p = C.c_void_p()
        _check(lib().fdgpu_table_device_ptr(self._h, table, C.byref(p)), "fdgpu_table_device_ptr")
        return int(p.value)


def run_spec(spec_path: str) -> None:
    """The drop-in `sphfd` run of a spec file (src/sphfd.c:200-257), in process."""
    _check(lib().fdgpu_run_spec(spec_path.encode()), "fdgpu_run_spec")


def sphfd(parline: str, output_path: str) -> int:
    """The reference's in-process entry point `sphfd(2, {"sphfd", parline},
    output_path)` (src/sphfd.c:260): solves one parlist line and writes its
    table under output_path.  Returns the C return code (0 or FDGPU_E*).  The
    ctypes call releases the GIL, so calls from several Python threads run
    concurrently and are gathered into shared launches, as OpenMP threads'
    calls are."""
    av = (C.c_char_p * 2)(b"sphfd", parline.encode())
    return int(lib().sphfd(2, av, output_path.encode()))


def sphfd_info() -> dict:
    """Counters of the in-process entry point: calls and the launches they
    were solved in."""
    info = (C.c_ulonglong * 2)()
    _check(lib().fdgpu_sphfd_info(info, 2), "fdgpu_sphfd_info")
    return {"calls": int(info[0]), "launches": int(info[1])}
