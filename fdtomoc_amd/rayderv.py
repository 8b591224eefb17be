"""Host side of sphrayderv's ray tracing over resident tables.

sphrayderv (src/sphrayderv.c) reads a spec file, a station list and an event
file, traces every pick's ray from the event back to the station through the
station's travel-time table, and writes the inversion's derivative files.
This module keeps the reference's set-up arithmetic -- the grid in radians
and geocentric colatitude, the coarse node coordinates, station and event
positions, the arrival times relative to the origin time -- so that the rays
handed to the device (`fdgpu_rays_trace`) start from the same doubles the
reference traces from, and it parses the reference's input files.

Units after set-up (sphrayderv.c:527-545): longitudes in radians, latitudes
as geocentric colatitudes in radians, depths in km positive down.
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np

# sphrayderv.c:95-97: float literals held in doubles (rearth stays a float)
HPI = float(np.float32(1.570796))
DEGRAD = float(np.float32(0.017453292))
REARTH = 6371.0


def glat(hlat: float) -> float:
    """Geographic -> geocentric latitude (lib/common/geographic_method.c:7-16),
    with the function's float constants."""
    halfpi, polfac, elfac = HPI, float(np.float32(0.010632)), float(np.float32(0.993277))
    if halfpi - abs(hlat) >= 0.05:
        return math.atan(elfac * math.sin(hlat) / math.cos(hlat))
    return hlat / elfac - abs(polfac) * (1 if hlat >= 0 else -1)


def glatinv(hlat: float) -> float:
    """Geocentric -> geographic latitude (lib/common/geographic_method.c:22-35)."""
    if 1.570796 - abs(hlat) >= 0.05:
        return math.atan(math.sin(hlat) / math.cos(hlat) / 0.993277)
    return (hlat + 0.010632) * 0.993277 if hlat > 0 else (hlat - 0.010632) * 0.993277


@dataclasses.dataclass
class RayGrid:
    """The fine grid of the tables and the coarse inversion grid, in
    sphrayderv's units (the fields of fdgpu_ray_grid)."""
    nx: int
    ny: int
    nz: int
    x0: float
    y0: float
    z0: float
    df: float
    dq: float
    h: float
    gx: np.ndarray      # float32 coarse node coordinates
    gy: np.ndarray
    gz: np.ndarray
    ido1d: int = 0
    max_segments: int = 0

    @property
    def nxc(self) -> int:
        return len(self.gx)

    @property
    def nyc(self) -> int:
        return len(self.gy)

    @property
    def nzc(self) -> int:
        return len(self.gz)

    def segment_limit(self) -> int:
        return self.max_segments or 8 * (self.nx + self.ny + self.nz)


def grid_from_spec(nxc: int, nyc: int, nzc: int, igridx, igridy, igridz, x0_deg: float, y0_deg: float,
                   z0: float, h: float, dq_deg: float = 0.0, df_deg: float = 0.0, ido1d: int = 0) -> RayGrid:
    """sphrayderv's grid set-up from its spec values (sphrayderv.c:483-545,
    814-826): fine dimensions 1 + the sum of the igrid entries, origin to
    radians and geocentric colatitude, dq/df from h when below 1e-4, coarse
    node coordinates accumulated in float."""
    igridx, igridy, igridz = (np.asarray(a, dtype=np.int64) for a in (igridx, igridy, igridz))
    nx = 1 + int(igridx[:nxc - 1].sum())
    ny = 1 + int(igridy[:nyc - 1].sum())
    nz = 1 + int(igridz[:nzc - 1].sum())
    y0 = HPI - glat(y0_deg * DEGRAD)
    x0 = x0_deg * DEGRAD
    dq = dq_deg * DEGRAD
    df = df_deg * DEGRAD
    if abs(dq) < 0.0001:
        dq = h / REARTH
    if abs(df) < 0.0001:
        df = abs(h / (REARTH * math.sin(y0)))

    def coarse(origin, step, ig, n):
        g = np.zeros(n, dtype=np.float32)
        g[0] = np.float32(origin)
        for i in range(1, n):
            g[i] = np.float32(float(g[i - 1]) + step * int(ig[i - 1]))
        return g
    return RayGrid(nx, ny, nz, x0, y0, float(z0), df, dq, float(h), coarse(x0, df, igridx, nxc),
                   coarse(y0, dq, igridy, nyc), coarse(z0, h, igridz, nzc), ido1d)


@dataclasses.dataclass
class Station:
    name: str
    lon: float      # stlon: radians, held as float (sphrayderv.c:874)
    colat: float    # stlat: geocentric colatitude, float
    z: float        # stz: km (-elevation)
    tcor: tuple = (0.0, 0.0)


def station(name: str, lat_deg: float, lon_deg: float, elev_m: int, tcor=(0.0, 0.0)) -> Station:
    """A station-list line's position as sphrayderv stores it
    (sphrayderv.c:867-876): float longitude and colatitude in radians and
    the depth -(elevation/1000) in float."""
    slat, slon = float(np.float32(lat_deg)), float(np.float32(lon_deg))
    return Station(name, float(np.float32(slon * DEGRAD)), float(np.float32(HPI - glat(slat * DEGRAD))),
                   float(-np.float32(np.float32(int(elev_m)) / np.float32(1000.0))),
                   (float(np.float32(tcor[0])), float(np.float32(tcor[1]))))


def event_position(lat_deg: float, lon_deg: float, depth: float) -> tuple[float, float, float]:
    """An event header's position (sphrayderv.c:949-951): the depth is read
    as a float, latitude and longitude as doubles."""
    return lon_deg * DEGRAD, HPI - glat(lat_deg * DEGRAD), float(np.float32(depth))


def epoch(year: int, day: int, hour: int, minute: int, sec: float) -> float:
    """htoe (lib/common/time_process.c:85-109): seconds since 1970."""
    def leap(y):
        return (y % 4 == 0 and y % 100 != 0) or y % 400 == 0
    days = 0
    if year > 1970:
        for y in range(1970, year):
            days += 365 + leap(y)
    elif year < 1970:
        for y in range(year, 1970):
            days -= 365 + leap(y)
    days += day - 1
    return float(days * 86400) + 3600.0 * hour + 60.0 * minute + sec


def date_of(epoch_s: float) -> tuple:
    """etoh (lib/common/time_process.c:11-84): (year, day of year, hour,
    minute, seconds) of an epoch time."""
    doy = int(epoch_s / 86400.0)
    secleft = math.fmod(epoch_s, 86400.0)
    hour = minute = 0
    second = 0.0
    if secleft != 0.0:
        if secleft < 0:
            doy -= 1
            secleft += 86400
        hour = int(secleft / 3600)
        secleft = math.fmod(secleft, 3600.0)
        minute = int(secleft / 60)
        second = math.fmod(secleft, 60.0)

    def leap(y):
        return (y % 4 == 0 and y % 100 != 0) or y % 400 == 0
    if doy >= 0:
        year = 1970
        while doy >= 365 + leap(year):
            doy -= 365 + leap(year)
            year += 1
    else:
        year = 1969
        while doy < 0:
            doy += 365 + leap(year)
            year -= 1
    return year, doy + 1, hour, minute, second


# --------------------------------------------------- sphrayderv's input files

def read_spec(path: str) -> dict:
    """The spec keys sphrayderv reads (sphrayderv.c:198-480): the first
    token of a line is the key, the next its value; igridx/igridy/igridz
    carry one value per coarse node."""
    vals, rows = {}, {}
    with open(path) as f:
        for line in f:
            tok = line.split()
            if len(tok) < 2:
                continue
            if tok[0] in ("igridx", "igridy", "igridz"):
                rows.setdefault(tok[0], [int(t) for t in tok[1:]])
            else:
                vals.setdefault(tok[0], tok[1])
    vals.update(rows)
    return vals


def read_station_list(path: str) -> list:
    """Station-list lines `sty stx ielev name slat slon tcorP tcorS`
    (sphrayderv.c:853-880): Station records (slat/slon are the positions
    used) in file order."""
    out = []
    with open(path) as f:
        for line in f:
            tok = line.split()
            if len(tok) < 8:
                continue
            out.append(station(tok[3], float(tok[4]), float(tok[5]), int(tok[2]), (float(tok[6]), float(tok[7]))))
    return out


def read_events(path: str) -> list:
    """Event blocks (sphrayderv.c:914-1044): a header `year day hour minute
    sec lat lon depth evid`, pick lines `sta year day hour minute sec phase
    rwt` (a leading '*' before the phase is dropped), a blank line."""
    events, cur = [], None
    with open(path) as f:
        for line in f:
            tok = line.split()
            if not tok:
                if cur is not None:
                    events.append(cur)
                    cur = None
                continue
            if cur is None:
                cur = dict(time=(int(tok[0]), int(tok[1]), int(tok[2]), int(tok[3]), float(np.float32(tok[4]))),
                           lat=float(tok[5]), lon=float(tok[6]), depth=float(np.float32(tok[7])), evid=tok[8],
                           picks=[])
                continue
            rest = line.split(None, 6)[6] if len(line.split(None, 6)) > 6 else ""
            rest = rest.strip()
            if rest.startswith("*"):
                rest = rest[1:]
            rt = rest.split()
            cur["picks"].append(dict(sta=tok[0], time=(int(tok[1]), int(tok[2]), int(tok[3]), int(tok[4]),
                                                       float(np.float32(tok[5]))),
                                     phase=rt[0][0] if rt else " ", rwt=float(np.float32(rt[1])) if len(rt) > 1 else 0.0))
    if cur is not None:
        events.append(cur)
    return events


# ------------------------------------------------------- the program's files

def run(spec_path: str, trace, workdir: str = ".") -> dict:
    """sphrayderv's local-earthquake run (src/sphrayderv.c main with its
    defaults: local events, no shots or teleseisms, ivpvs 0, matrices and
    data files written) with the rays traced by `trace` and the reference's
    bookkeeping and files: dtdsfil (binary rows of weighted derivatives),
    resfile (weighted residuals), dtdhfil (hypocentre partials), bookfil,
    hitfile, sclefil, dotfile and headfil.  The error and log text files
    (statistics and warnings) are not written.

    trace(grid, keys, rays) -> list of dicts (status, tev, dtdx, dtdy, dtdz,
    ai, az, idx, val): keys[r] = (station, phase letter of its table file),
    rays[r] = (ex, ey, ez, xs, ys, zs).  Returns counters."""
    import os
    import struct
    sp = read_spec(os.path.join(workdir, spec_path))
    for key, default in (("dotel", 0), ("doshot", 0), ("ivpvs", 0), ("nomat", 0), ("iray", 0), ("iraystat", 0),
                         ("istacor", 0)):
        if int(sp.get(key, default)) not in (0,):
            raise NotImplementedError(f"{key} = 1 (only local earthquakes with the default outputs)")
    idatout = int(sp.get("idatout", 1)) == 1
    ivs = 0 if int(sp.get("ivs", 1)) != 1 else 1
    resflag = float(np.float32(sp.get("resflag", 1.0)))
    nxc, nyc, nzc = int(sp["nxc"]), int(sp["nyc"]), int(sp["nzc"])
    grid = grid_from_spec(nxc, nyc, nzc, sp["igridx"], sp["igridy"], sp["igridz"], float(sp.get("x0", 0.0)),
                          float(sp.get("y0", 0.0)), float(sp.get("z0", 0.0)), float(sp["h"]),
                          float(sp.get("dq", 0.0)), float(sp.get("df", 0.0)))
    maxvarc = nxc * nyc * nzc
    sta = read_station_list(os.path.join(workdir, sp["stafile"]))
    by_name = {}
    for kn, s in enumerate(sta):
        by_name.setdefault(s.name, kn)
    events = read_events(os.path.join(workdir, sp["locdfil"]))
    f32 = np.float32
    # trace every pick of every in-grid event at once
    keys, rays, owner = [], [], []
    for e, ev in enumerate(events):
        ex, ey, ez = event_position(ev["lat"], ev["lon"], ev["depth"])
        ies, jes, kes = int((ex - grid.x0) / grid.df), int((ey - grid.y0) / grid.dq), int((ez - grid.z0) / grid.h)
        ev["pos"] = (ex, ey, ez)
        ev["inside"] = not (ex < grid.x0 or ies >= grid.nx - 1 or ey < grid.y0 or jes >= grid.ny - 1 or ez < grid.z0
                            or kes >= grid.nz - 1)
        if not ev["inside"]:
            continue
        for p, pk in enumerate(ev["picks"]):
            kn = by_name.get(pk["sta"])
            if kn is None:
                continue
            s = sta[kn]
            keys.append((pk["sta"], "S" if (ivs and pk["phase"] == "S") else "P"))
            rays.append((ex, ey, ez, s.lon, s.colat, s.z))
            owner.append((e, p))
    traced = dict(zip(owner, trace(grid, keys, rays))) if rays else {}

    nhit, indx, jndx = {}, {}, []
    kbl, mhit, mndex, mndm = 0, set(), {}, []
    files = {k: open(os.path.join(workdir, sp[k]), "wb") for k in ("dtdsfil", "resfile", "dtdhfil", "bookfil")}
    dot = open(os.path.join(workdir, sp["dotfile"]), "w") if idatout else None
    hed = open(os.path.join(workdir, sp["headfil"]), "w") if idatout else None
    nev = nwritten = 0
    for e, ev in enumerate(events):
        nev += 1
        if not ev["inside"]:
            continue
        ex, ey, ez = ev["pos"]
        dpot = epoch(*ev["time"])
        rows, avresev, facsev = [], f32(0), f32(0)
        for p, pk in enumerate(ev["picks"]):
            pwt = f32(f32(1.0) / f32(f32(pk["rwt"]) * f32(pk["rwt"])))
            obstime = f32(epoch(*pk["time"]) - dpot)
            row = dict(pick=pk, good=False, obstime=obstime, pwt=pwt)
            rows.append(row)
            kn = by_name.get(pk["sta"])
            r = traced.get((e, p))
            if kn is None or r is None or r["status"] == 0:
                continue
            iph = 1 if pk["phase"] == "S" else 0
            tc = f32(sta[kn].tcor[iph])
            dt = float(obstime) - ((r["tev"] + 0.0) + float(tc))
            wet = np.sqrt(pwt, dtype=np.float32)
            res1 = f32(dt * float(pwt))
            avresev = f32(avresev + res1)
            facsev = f32(facsev + pwt)
            row.update(resmin=f32(dt), dat=f32(dt * float(wet)),
                       am=[wet, f32(f32(r["dtdx"]) * wet), f32(f32(r["dtdy"]) * wet), f32(f32(r["dtdz"]) * wet)],
                       az=r["az"], ai=r["ai"])
            if r["status"] != 1:
                continue
            row["good"] = True
            if p == 0:
                kbl, mhit, mndex, mndm = 0, set(), {}, []
            vm = []
            for jb, du in zip(r["idx"], r["val"]):
                j = int(jb) + (maxvarc if iph == 1 else 0)
                vm.append((j, f32(f32(du) * wet)))
                if j not in mhit:
                    mhit.add(j)
                    mndex[j] = kbl
                    mndm.append(j)
                    kbl += 1
                if nhit.get(j, 0) == 0:
                    indx[j] = len(jndx)
                    jndx.append(j)
                nhit[j] = nhit.get(j, 0) + 1
            row["vm"] = vm
        avresev = f32(avresev / facsev) if facsev != 0 else f32(np.nan)
        nwrt = 0
        for row in rows:
            if not row["good"]:
                continue
            vmp = {}
            for j, v in row["vm"]:
                vmp[mndex[j]] = f32(vmp.get(mndex[j], f32(0)) + v)
            cols = [i1 for i1 in range(kbl) if vmp.get(i1, 0) != 0]
            if not cols:
                continue
            out = [struct.pack("<i", len(cols))]
            for i1 in cols:
                out.append(struct.pack("<i", indx[mndm[i1]] + 1) + np.float32(vmp[i1]).tobytes())
            files["dtdsfil"].write(b"".join(out))
            files["resfile"].write(np.float32(row["dat"]).tobytes())
            files["dtdhfil"].write("".join("%f\n" % float(a) for a in row["am"]).encode())
            nwrt += 1
        files["bookfil"].write(f"{nev} {nwrt}\n".encode())
        nwritten += nwrt
        if idatout:
            iyr, jday, ihr, imn, dsec = date_of(dpot)
            sec = float(f32(dsec))
            xlon = ex / DEGRAD
            xlat = glatinv(HPI - ey) / DEGRAD
            head = "%4d %3d %2d %2d %8.4f %9.5f %10.5f %8.4f %12s %13.3f\n" % (
                iyr, jday, ihr, imn, sec, xlat, xlon, ez, ev["evid"], float(avresev))
            hed.write(head)
            dot.write(head)
            for row in rows:
                if not row["good"]:
                    continue
                pk = row["pick"]
                iyr, jday, ihr, imn, dsec = date_of(float(row["obstime"]) + dpot)
                mark = "*" if abs(float(row["resmin"])) > resflag else " "
                dot.write("%s  %4d %3d %2d %2d %7.3f %c %15.3f %10.3f %c %f %f\n" % (
                    pk["sta"], iyr, jday, ihr, imn, dsec, pk["phase"], pk["rwt"], float(row["resmin"]), mark,
                    float(row["az"]), float(row["ai"])))
            dot.write("\n")
    files["dtdsfil"].write(struct.pack("<ii", -1, len(jndx)) + np.asarray(jndx, dtype="<i4").tobytes())
    files["bookfil"].write(b"-2 -2\n")
    for f in files.values():
        f.close()
    with open(os.path.join(workdir, sp["sclefil"]), "w") as f:
        f.write(f"{len(jndx)}\n" + "".join(f"{j}\n" for j in jndx))
    with open(os.path.join(workdir, sp["hitfile"]), "w") as f:
        f.write(f"{len(jndx)}" + "".join(f"{j} {nhit[j]}\n" for j in jndx))
    if idatout:
        dot.close()
        hed.close()
    return dict(events=nev, rays=len(rays), written=nwritten, variables=len(jndx))


def device_trace(ctx, workdir: str, timedir: str):
    """A trace function for `run` on the device: each station-phase table
    file of the run loaded once as a resident table (fdgpu_table_load, read
    as the reference reads it, sphrayderv.c:1163-1217), every ray traced in
    one fdgpu_rays_trace call; returns the table ids it loaded too."""
    import os
    from . import fdgpu
    ids = {}

    def trace(grid, keys, rays):
        for k in sorted(set(keys)):
            if k not in ids:
                path = os.path.join(workdir, timedir, f"{k[0]}.{'p' if k[1] == 'P' else 's'}times")
                ids[k] = ctx.load_table(path, (grid.nz, grid.ny, grid.nx))
        res, idx, val = ctx.trace_rays(grid, fdgpu.ray_array([ids[k] for k in keys], rays))
        return [dict(status=int(d["status"]), tev=float(d["tev"]), dtdx=float(d["dtdx"]), dtdy=float(d["dtdy"]),
                     dtdz=float(d["dtdz"]), ai=float(d["ai"]), az=float(d["az"]),
                     idx=idx[d["offset"]:d["offset"] + d["ndu"]], val=val[d["offset"]:d["offset"] + d["ndu"]])
                for d in res]
    return trace, ids


def main() -> int:
    """`python -m fdtomoc_amd.rayderv` reads the spec file's name on stdin, as
    the reference program does (sphrayderv.c:191-196), and runs it with the
    rays on the device."""
    import sys
    from . import fdgpu
    spec = sys.stdin.readline().strip()
    if not spec:
        print("sphrayderv: no spec file named on stdin", file=sys.stderr)
        return 1
    timedir = read_spec(spec).get("timedir", "./ttimes")
    ctx = fdgpu.Context(0)
    try:
        trace, ids = device_trace(ctx, ".", timedir)
        info = run(spec, trace, ".")
    finally:
        ctx.close()
    print(f"sphrayderv: {info['events']} events, {info['rays']} rays, {info['written']} rows, "
          f"{info['variables']} variables")
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
