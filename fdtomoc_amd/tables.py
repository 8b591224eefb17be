"""Velocity-model and travel-time-table files in the FDtomoC binary layout.

Layout (include/common/vhead.h:4-11, read at src/sphfd.c:611, written at
src/sphfd.c:5800-5801): a packed 232-byte header -- ``char header[120]``,
twelve little-endian doubles ``fxs fys fzs clat clon cz x0 y0 z0 dx dy dz``,
a float ``az`` and three ints ``nx ny nz`` -- followed by ``nx*ny*nz``
float32 values, x (east) fastest, then y (north to south), then z (top down).

The header text is five 4-character fields (HEAD, type, system, quantity,
flattening) and a comment, space padded (lib/common/string_process.c:104-122).
A travel-time table carries quantity ``PTIM`` at bytes 12-15 and the source in
geocentric form (lon rad, colatitude rad, radius km) in ``fxs fys fzs``
(src/sphfd.c:5744-5769).
"""
from __future__ import annotations

import dataclasses
import os

import numpy as np

VHEAD_DTYPE = np.dtype([
    ("header", "S120"),
    ("fxs", "<f8"), ("fys", "<f8"), ("fzs", "<f8"),
    ("clat", "<f8"), ("clon", "<f8"), ("cz", "<f8"),
    ("x0", "<f8"), ("y0", "<f8"), ("z0", "<f8"),
    ("dx", "<f8"), ("dy", "<f8"), ("dz", "<f8"),
    ("az", "<f4"), ("nx", "<i4"), ("ny", "<i4"), ("nz", "<i4"),
])
assert VHEAD_DTYPE.itemsize == 232


def header_text(kind: str = "FINE", system: str = "SPHR", quantity: str = "BMOD",
                flatten: str = "NOFL", comment: str = "") -> bytes:
    """The 120-byte header text, as hdr_appender builds it (string_process.c:104-122)."""
    txt = "%4s%4s%4s%4s%4s%s" % ("HEAD", kind, system, quantity, flatten, comment)
    if len(txt) > 120:
        raise ValueError("header comment too long")
    return txt.ljust(120).encode("ascii")


@dataclasses.dataclass
class Grid3D:
    """A header plus a (nz, ny, nx) float32 array."""
    head: np.ndarray          # one VHEAD_DTYPE record
    data: np.ndarray          # float32, shape (nz, ny, nx)

    @property
    def shape(self):
        return self.data.shape


def make_header(nx: int, ny: int, nz: int, x0: float, y0: float, z0: float,
                dx: float, dy: float, dz: float, text: bytes | None = None,
                clat: float | None = None, clon: float | None = None,
                cz: float = 0.0) -> np.ndarray:
    h = np.zeros((), dtype=VHEAD_DTYPE)
    h["header"] = text if text is not None else header_text()
    h["x0"], h["y0"], h["z0"] = x0, y0, z0
    h["dx"], h["dy"], h["dz"] = dx, dy, dz
    h["clat"] = y0 if clat is None else clat
    h["clon"] = x0 if clon is None else clon
    h["cz"] = cz
    h["nx"], h["ny"], h["nz"] = nx, ny, nz
    return h


def write_grid(path: str | os.PathLike, head: np.ndarray, data: np.ndarray) -> None:
    data = np.ascontiguousarray(data, dtype="<f4")
    with open(path, "wb") as f:
        f.write(np.asarray(head, dtype=VHEAD_DTYPE).tobytes())
        f.write(data.tobytes())


def read_grid(path: str | os.PathLike, count: int | None = None) -> Grid3D:
    """Read a model or table.  ``count`` limits the floats read (a P+S model
    file holds two grids back to back, as c2f writes them)."""
    raw = np.fromfile(path, dtype=np.uint8)
    head = raw[:232].view(VHEAD_DTYPE)[0].copy()
    nx, ny, nz = int(head["nx"]), int(head["ny"]), int(head["nz"])
    n = nx * ny * nz if count is None else count
    vals = raw[232:232 + 4 * n].view("<f4")
    if vals.size != n:
        raise ValueError(f"{path}: expected {n} floats, found {vals.size}")
    data = vals.reshape(nz, ny, nx) if n == nx * ny * nz else vals
    return Grid3D(head=head, data=data)
