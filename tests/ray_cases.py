"""Seeded sphrayderv cases: a coarse/fine grid, P and S velocity models,
stations on the top surface with their travel-time tables (solved by the
sphfd oracle), and events with P and S picks at every station -- written as
the reference's input files (spec, station list, event file, table
directory) for the reference program, and as rays for the oracle and the
device.  Test infrastructure only."""
from __future__ import annotations

import dataclasses
import hashlib
import os

import numpy as np

import oracle_bind as ob
from fdtomoc_amd import rayderv, synth, tables


@dataclasses.dataclass
class RayCase:
    seed: int
    spec: dict                 # coarse grid values of the spec file
    grid: rayderv.RayGrid
    g: synth.GridSpec          # the fine grid as the tables' par files state it
    vp: np.ndarray
    vs: np.ndarray
    stations: list             # (name, lat, lon, elev_m, tcor_p, tcor_s)
    events: list               # (year, day, hour, minute, sec, lat, lon, depth, evid, picks)
                               # picks: (station, year, day, hour, minute, sec, phase, rwt)

    def table_par(self, st) -> dict:
        return dict(nx=self.g.nx, ny=self.g.ny, nz=self.g.nz, h=self.g.h, x0=self.g.x0, y0=self.g.y0,
                    z0=self.g.z0, fxs=st[2], fys=st[1], fzs=-st[3] / 1000.0, reverse=4)


def ray_case(seed: int, nxc: int = 9, nyc: int = 9, nzc: int = 6, nsta: int = 4, nev: int = 8,
             ido1d: int = 0) -> RayCase:
    rng = np.random.default_rng(seed)
    igx = rng.integers(3, 6, size=nxc - 1)
    igy = rng.integers(3, 6, size=nyc - 1)
    igz = rng.integers(2, 5, size=nzc - 1)
    h = 1.0
    nx, ny, nz = 1 + int(igx.sum()), 1 + int(igy.sum()), 1 + int(igz.sum())
    g = synth.centred_grid(nx, ny, nz, h, lat=23.0 + seed % 5, lon=121.0)
    spec = dict(nxc=nxc, nyc=nyc, nzc=nzc, igridx=igx, igridy=igy, igridz=igz, x0=g.x0, y0=g.y0, z0=g.z0, h=h)
    grid = rayderv.grid_from_spec(nxc, nyc, nzc, igx, igy, igz, g.x0, g.y0, g.z0, h, ido1d=ido1d)
    assert (grid.nx, grid.ny, grid.nz) == (nx, ny, nz)
    vp = synth.gradient_velocity(g, 4.5, 7.5, 30.0) * synth.gaussian_anomalies(g, rng, 3, 0.1, (3.0, 8.0))
    vp = np.ascontiguousarray(vp, dtype=np.float32)
    vs = (vp / np.float32(1.73)).astype(np.float32)
    stations = []
    for s in range(nsta):
        i, j = int(rng.integers(2, nx - 2)), int(rng.integers(2, ny - 2))
        lon, lat = g.node_lonlat(i, j)
        stations.append((f"ST{s:02d}", round(lat, 5), round(lon, 5), 0, round(float(rng.uniform(-0.2, 0.2)), 3),
                         round(float(rng.uniform(-0.3, 0.3)), 3)))
    events = []
    for e in range(nev):
        fi, fj, fk = rng.uniform(0.5, nx - 2.5), rng.uniform(0.5, ny - 2.5), rng.uniform(0.5, nz - 2.5)
        ex, ey = grid.x0 + fi * grid.df, grid.y0 + fj * grid.dq
        lon = round(ex / rayderv.DEGRAD, 5)
        lat = round(rayderv.glatinv(rayderv.HPI - ey) / rayderv.DEGRAD, 5)
        dep = round(grid.z0 + fk * grid.h, 3)
        sec0 = round(float(rng.uniform(0, 30)), 2)
        picks = []
        for st in stations:
            for ph in "PS":
                tt = float(rng.uniform(2.0, 12.0)) * (1.0 if ph == "P" else 1.7)
                rwt = float(rng.choice([1.0, 0.5, 0.25]))
                picks.append((st[0], 2021, 45, 6, 10, round(sec0 + tt, 2), ph, rwt))
        events.append((2021, 45, 6, 10, sec0, lat, lon, dep, f"EV{seed:03d}{e:02d}", picks))
    return RayCase(seed, spec, grid, g, vp, vs, stations, events)


def case_tables(case: RayCase) -> dict:
    """(station, phase) -> table solved by the sphfd oracle."""
    out = {}
    for st in case.stations:
        par = ob.Par(**case.table_par(st))
        for ph, v in (("P", case.vp), ("S", case.vs)):
            t, _ = ob.solve(par, ob.slowness(v))
            out[(st[0], ph)] = t
    return out


def table_hash(t: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(t, dtype="<f4").tobytes()).hexdigest()


def case_rays(case: RayCase) -> list:
    """Every pick as a ray, in the reference's order (events, then picks):
    dicts with the ray's (ex, ey, ez, xs, ys, zs), its table key, the
    arrival time after the origin (float, as obstime), the station
    correction and the weight 1/rwt^2."""
    sta = {s[0]: rayderv.station(s[0], s[1], s[2], s[3], (s[4], s[5])) for s in case.stations}
    out = []
    for ev in case.events:
        year, day, hour, minute, sec, lat, lon, dep, evid, picks = ev
        ex, ey, ez = rayderv.event_position(lat, lon, dep)
        dpot = rayderv.epoch(year, day, hour, minute, float(np.float32(sec)))
        for p, (name, y, d, hh, mm, s, ph, rwt) in enumerate(picks):
            st = sta[name]
            tarr = rayderv.epoch(y, d, hh, mm, float(np.float32(s)))
            out.append(dict(event=evid, pick=p, station=name, phase=ph, ray=(ex, ey, ez, st.lon, st.colat, st.z),
                            key=(name, ph), obstime=np.float32(tarr - dpot),
                            tc=np.float32(st.tcor[0 if ph == "P" else 1]),
                            pwt=np.float32(np.float32(1.0) / np.float32(np.float32(rwt) * np.float32(rwt)))))
    return out


# ------------------------------------------------ the reference's input files

def write_inputs(case: RayCase, wd: str, tabs: dict) -> str:
    """Writes the spec, station list, event file and table directory of a
    case for the reference program; returns the spec file's name."""
    os.makedirs(os.path.join(wd, "ttimes"), exist_ok=True)
    g = case.g
    for (name, ph), t in tabs.items():
        head = tables.make_header(g.nx, g.ny, g.nz, g.x0, g.y0, g.z0, g.df_deg(), g.dq_deg(), g.h)
        tables.write_grid(os.path.join(wd, "ttimes", f"{name}.{'p' if ph == 'P' else 's'}times"), head, t)
    with open(os.path.join(wd, "stations.txt"), "w") as f:
        # sty stx ielev name slat slon tcorP tcorS (sphrayderv.c:867)
        f.write("\n".join(f"{s[1]:.5f} {s[2]:.5f} {s[3]} {s[0]} {s[1]:.5f} {s[2]:.5f} {s[4]:.3f} {s[5]:.3f}"
                          for s in case.stations))
        f.write("\n")
    with open(os.path.join(wd, "events.txt"), "w") as f:
        for year, day, hour, minute, sec, lat, lon, dep, evid, picks in case.events:
            f.write(f"{year} {day} {hour} {minute} {sec:.2f} {lat:.5f} {lon:.5f} {dep:.3f} {evid}\n")
            for name, y, d, hh, mm, s, ph, rwt in picks:
                f.write(f"{name} {y} {d} {hh} {mm} {s:.2f} {ph} {rwt}\n")
            f.write("\n")
    sp = case.spec

    def row(key, arr):
        return key + " " + " ".join(str(int(v)) for v in list(arr) + [0]) + "\n"
    text = (f"nxc {sp['nxc']}\nnyc {sp['nyc']}\nnzc {sp['nzc']}\nh {sp['h']!r}\n"
            f"x0 {sp['x0']!r}\ny0 {sp['y0']!r}\nz0 {sp['z0']!r}\n"
            + row("igridx", sp["igridx"]) + row("igridy", sp["igridy"]) + row("igridz", sp["igridz"])
            + "stafile stations.txt\nlocdfil events.txt\noldvfil none.mod\ntelrerr telrerr.txt\n"
              "dtdsfil dtds.bin\nresfile res.bin\nhitfile hit.txt\ndtdhfil dtdh.txt\nbookfil book.txt\n"
              "sclefil scle.txt\ndotfile dot.txt\nheadfil head.txt\ntimedir ttimes\n")
    with open(os.path.join(wd, "spec.txt"), "w") as f:
        f.write(text)
    return "spec.txt"


def read_outputs(wd: str) -> dict:
    """The reference's derivative files: per written ray (in its order) the
    variables and weighted derivatives (dtdsfil), the weighted residual
    (resfile) and the four hypocentre derivatives (dtdhfil, printed with
    six decimals); the good picks of the data file (dotfile) with residual,
    azimuth and incidence; and the rays written per event (bookfil)."""
    raw = np.fromfile(os.path.join(wd, "dtds.bin"), dtype="<i4")
    rays, p = [], 0
    while raw[p] != -1:
        ja = int(raw[p])
        pairs = raw[p + 1:p + 1 + 2 * ja].reshape(ja, 2)
        rays.append((pairs[:, 0].copy(), pairs[:, 1].copy().view(np.float32)))
        p += 1 + 2 * ja
    mbl = int(raw[p + 1])
    jndx = raw[p + 2:p + 2 + mbl].copy()
    assert p + 2 + mbl == len(raw)
    res = np.fromfile(os.path.join(wd, "res.bin"), dtype="<f4")
    dtdh = np.loadtxt(os.path.join(wd, "dtdh.txt"), dtype=np.float64).reshape(-1, 4)
    written = []
    for (pos, vals) in rays:
        j = jndx[pos - 1]
        order = np.argsort(j, kind="stable")
        written.append((j[order].astype(np.int32), vals[order]))
    good = []
    with open(os.path.join(wd, "dot.txt")) as f:
        ev = None
        for line in f:
            parts = line.split()
            if not parts:
                continue
            if len(parts) == 10 and parts[0].isdigit():
                ev = parts[8]
                continue
            # sta iyr jday ihr imn dsec phs rwts resmin [mark] azs ais
            mark = len(parts) == 12
            good.append(dict(event=ev, station=parts[0], phase=parts[6], resmin=float(parts[8]),
                             az=float(parts[10 if mark else 9]), ai=float(parts[11 if mark else 10])))
    book = np.loadtxt(os.path.join(wd, "book.txt"), dtype=np.int64).reshape(-1, 2)
    assert len(written) == len(res) == len(dtdh)
    return dict(written=written, res=res, dtdh=dtdh, good=good, book=book)
