"""One step of FDtomoC's location loop on the device, as run.sh chains its
programs (c2f -> sphfd -> sphfdloc, run.sh:19-31): a coarse model becomes
resident P and S models (c2f, src/c2f.c), each station's P and S tables are
solved on them (sphfd), and events are located over the resident tables
(sphfdloc's grid search, src/sphfdloc.c:752-1130).  Nothing goes through
files; every step is the library's C ABI (fdgpu.py)."""
from __future__ import annotations

import dataclasses

from . import fdgpu


@dataclasses.dataclass
class Pick:
    station: int        # index into the stations list
    phase: str          # "P" or "S"
    obstime: float      # arrival time minus the event's reference time, s
    pwt: float = 1.0    # weight 1/rwt^2
    good: int = 1       # 0: flagged unusable ('*')


def locate_with_coarse_model(ctx: fdgpu.Context, coarse: fdgpu.Coarse, station_pars: list, events: list,
                             loc: fdgpu.CLocParams, keep: bool = False):
    """station_pars: one par dict per station (fdgpu_par fields, the fine grid
    of `coarse`); events: lists of Pick.  Returns (locations, ids) where
    locations are fdgpu.Context.locate's dicts and ids = (model_p, model_s,
    [(table_p, table_s) per station]) when keep is set (the caller releases
    them), else None."""
    mp, ms = ctx.models_from_coarse(coarse)
    ids = ctx.solve([fdgpu.TableSpec(p, m) for p in station_pars for m in (mp, ms)], fdgpu.BATCHED)
    pairs = [(ids[2 * s], ids[2 * s + 1]) for s in range(len(station_pars))]
    evs = []
    for picks in events:
        evs.append(dict(table=[pairs[k.station][0 if k.phase == "P" else 1] for k in picks],
                        obstime=[k.obstime for k in picks], pwt=[k.pwt for k in picks],
                        phase="".join(k.phase for k in picks), isgood=[k.good for k in picks]))
    out = ctx.locate(loc, evs)
    if keep:
        return out, (mp, ms, pairs)
    for i in ids:
        ctx.release(i)
    ctx.release_model(mp)
    ctx.release_model(ms)
    return out, None
