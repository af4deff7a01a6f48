"""GDSII stream files: a reader and a writer of the parts a photonic layout uses.

The reference reads and writes GDS through gdspy (beamz/design/core.py:445-553),
which this environment does not have; this module implements the GDSII stream
format itself (Calma GDSII Stream Format release 6):

* a file is a sequence of records: 2-byte big-endian length (header included),
  1-byte record type, 1-byte data type, data;
* data types: 0 none, 1 bit array (2 bytes), 2 int16, 3 int32, 5 8-byte real,
  6 ASCII string (padded with NUL to an even length);
* 8-byte reals are sign (1 bit), excess-64 base-16 exponent (7 bits) and a
  56-bit mantissa m: value = (-1)^sign * (m / 2^56) * 16^(exponent - 64);
* coordinates are int32 multiples of the database unit given by UNITS.

The writer emits one library with one structure ("main" by default) of
BOUNDARY elements, in micron user units with a nanometre database unit as
gdspy does by default for the reference's export (unit 1e-6, precision 1e-9).
A polygon with holes is written as one keyhole boundary: each hole is joined to
the outline by a zero-width cut, since a GDSII boundary has no holes.

The reader returns, for every structure of the file, its polygons by (layer,
datatype) in user units, with BOUNDARY and BOX elements of the structure itself
and of every structure it references (SREF, AREF: reflection, magnification,
rotation and translation applied), which is what gdspy's
``cell.get_polygons(by_spec=True)`` gives the reference's import. PATH and
TEXT elements are counted and skipped.
"""
from __future__ import annotations

import math
import struct
from fractions import Fraction

import numpy as np

# record types
HEADER, BGNLIB, LIBNAME, UNITS, ENDLIB, BGNSTR, STRNAME, ENDSTR = 0x00, 0x01, 0x02, 0x03, 0x04, 0x05, 0x06, 0x07
BOUNDARY, PATH, SREF, AREF, TEXT, LAYER, DATATYPE, WIDTH, XY, ENDEL = (0x08, 0x09, 0x0A, 0x0B, 0x0C, 0x0D, 0x0E,
                                                                      0x0F, 0x10, 0x11)
SNAME, COLROW, STRANS, MAG, ANGLE, BOX, BOXTYPE = 0x12, 0x13, 0x1A, 0x1B, 0x1C, 0x2D, 0x2E
# data types
NODATA, BITARRAY, INT16, INT32, REAL8, ASCII = 0, 1, 2, 3, 5, 6

MAX_XY = 8191          # points per boundary record, closing point included


def encode_real8(v: float) -> bytes:
    """GDSII 8-byte real of v. A double's 53-bit significand fits the 56-bit mantissa at every base-16
    exponent, so the encoding is exact (computed in rational arithmetic; rounding to nearest only guards
    inputs that are not doubles)."""
    if v == 0.0:
        return b"\0" * 8
    sign = 0x80 if v < 0 else 0
    a = Fraction(abs(v))
    e = 64 + math.floor(math.log(float(a), 16)) + 1       # a = m * 16^(e-64), m in [1/16, 1)
    while a >= Fraction(16) ** (e - 64):
        e += 1
    while a < Fraction(16) ** (e - 65):
        e -= 1
    mant = round(a / Fraction(16) ** (e - 64) * 2 ** 56)
    if mant >= 2 ** 56:
        mant >>= 4
        e += 1
    if not 0 <= e <= 127:
        raise ValueError(f"{v} is out of the GDSII real range")
    return bytes([sign | e]) + mant.to_bytes(7, "big")


def decode_real8(b: bytes) -> float:
    """The double nearest to a GDSII 8-byte real (exact rational arithmetic, then one rounding)."""
    e = b[0] & 0x7F
    v = Fraction(int.from_bytes(b[1:8], "big"), 2 ** 56) * Fraction(16) ** (e - 64)
    return float(-v if b[0] & 0x80 else v)


def _record(rtype, dtype=NODATA, data=b""):
    if len(data) % 2:
        data += b"\0"
    n = len(data) + 4
    if n > 65535:
        raise ValueError("GDSII record longer than 65535 bytes")
    return struct.pack(">HBB", n, rtype, dtype) + data


def _ints16(*v):
    return struct.pack(f">{len(v)}h", *v)


def _ints32(v):
    v = np.asarray(v, dtype=">i4")
    return v.tobytes()


def _string(s):
    b = s.encode("ascii")
    return b + (b"\0" if len(b) % 2 else b"")


def _signed_area(pts):
    return 0.5 * sum(pts[i - 1][0] * pts[i][1] - pts[i][0] * pts[i - 1][1] for i in range(len(pts)))


def keyhole(outline, holes):
    """One closed point list for an outline with holes: every hole, traversed against the outline's
    orientation, is entered and left through a zero-width cut from its left-most vertex to the nearest vertex
    of the list built so far (holes taken from left to right)."""
    pts = [tuple(p) for p in outline]
    turn = _signed_area(pts) > 0
    for h in sorted(holes, key=lambda h: min(p[0] for p in h)):
        h = [tuple(p) for p in h]
        if (_signed_area(h) > 0) == turn:
            h = h[::-1]                                  # a hole runs against the outline
        k = min(range(len(h)), key=lambda q: (h[q][0], h[q][1]))
        j = min(range(len(pts)), key=lambda q: (pts[q][0] - h[k][0]) ** 2 + (pts[q][1] - h[k][1]) ** 2)
        ring = h[k:] + h[:k] + [h[k]]
        pts = pts[:j + 1] + ring + pts[j:]
    return pts


def write_gds(path, polygons, unit=1e-6, precision=1e-9, libname="BEAMZ", cellname="main"):
    """Write polygons [(layer, datatype, points, holes)] (points and holes in user units of `unit` metres)
    as BOUNDARY elements of one structure. Coordinates are rounded to the database unit `precision`."""
    scale = unit / precision
    date = (2000, 1, 1, 0, 0, 0) * 2
    out = [_record(HEADER, INT16, _ints16(600)), _record(BGNLIB, INT16, _ints16(*date)),
           _record(LIBNAME, ASCII, _string(libname)),
           _record(UNITS, REAL8, encode_real8(precision / unit) + encode_real8(precision)),
           _record(BGNSTR, INT16, _ints16(*date)), _record(STRNAME, ASCII, _string(cellname))]
    for layer, datatype, points, holes in polygons:
        pts = keyhole(points, holes) if holes else [tuple(p) for p in points]
        if len(pts) < 3:
            continue
        if len(pts) + 1 > MAX_XY:
            raise ValueError(f"a polygon of {len(pts)} points exceeds the GDSII limit of {MAX_XY - 1}")
        xy = np.rint(np.asarray(pts + [pts[0]], dtype=np.float64) * scale).astype(np.int64)
        if np.abs(xy).max() >= 2 ** 31:
            raise ValueError("coordinates exceed the 32-bit GDSII range at this precision")
        out += [_record(BOUNDARY), _record(LAYER, INT16, _ints16(layer)), _record(DATATYPE, INT16, _ints16(datatype)),
                _record(XY, INT32, _ints32(xy.ravel())), _record(ENDEL)]
    out += [_record(ENDSTR), _record(ENDLIB)]
    with open(path, "wb") as f:
        f.write(b"".join(out))


def _records(data):
    p = 0
    while p + 4 <= len(data):
        n, rtype, dtype = struct.unpack(">HBB", data[p:p + 4])
        if n < 4:
            break               # zero padding after ENDLIB
        yield rtype, dtype, data[p + 4:p + n]
        p += n


def read_gds(path):
    """Parse a GDSII file. Returns (user_unit_in_metres, {structure name: {(layer, datatype): [points]}}, skipped)
    with points (n, 2) float arrays in user units (the closing point dropped), references flattened, and the
    number of PATH/TEXT elements skipped."""
    with open(path, "rb") as f:
        data = f.read()
    db_in_user, db_in_m = 1e-3, 1e-9
    cells = {}
    cur = None
    el = None
    skipped = 0
    for rtype, dtype, d in _records(data):
        if rtype == UNITS:
            db_in_user, db_in_m = decode_real8(d[:8]), decode_real8(d[8:16])
        elif rtype == STRNAME:
            cur = {"polys": [], "refs": []}
            cells[d.rstrip(b"\0").decode("ascii")] = cur
        elif rtype in (BOUNDARY, BOX, SREF, AREF, PATH, TEXT):
            el = {"type": rtype, "layer": 0, "datatype": 0, "strans": 0, "mag": 1.0, "angle": 0.0}
        elif el is not None:
            if rtype == LAYER:
                el["layer"] = struct.unpack(">h", d[:2])[0]
            elif rtype in (DATATYPE, BOXTYPE):
                el["datatype"] = struct.unpack(">h", d[:2])[0]
            elif rtype == XY:
                el["xy"] = np.frombuffer(d, dtype=">i4").astype(np.float64).reshape(-1, 2) * db_in_user
            elif rtype == SNAME:
                el["sname"] = d.rstrip(b"\0").decode("ascii")
            elif rtype == COLROW:
                el["colrow"] = struct.unpack(">hh", d[:4])
            elif rtype == STRANS:
                el["strans"] = struct.unpack(">H", d[:2])[0]
            elif rtype == MAG:
                el["mag"] = decode_real8(d[:8])
            elif rtype == ANGLE:
                el["angle"] = decode_real8(d[:8])
            elif rtype == ENDEL:
                if el["type"] in (BOUNDARY, BOX):
                    pts = el["xy"]
                    if len(pts) > 1 and np.array_equal(pts[0], pts[-1]):
                        pts = pts[:-1]
                    cur["polys"].append(((el["layer"], el["datatype"]), pts))
                elif el["type"] in (SREF, AREF):
                    cur["refs"].append(el)
                else:
                    skipped += 1
                el = None
    flat = {name: _flatten(name, cells, ()) for name in cells}
    return db_in_m / db_in_user, flat, skipped


def _transform(pts, el, origin):
    """Reference transform: reflection about x (STRANS bit 15), magnification, rotation, then translation."""
    p = np.array(pts, dtype=np.float64)
    if el["strans"] & 0x8000:
        p[:, 1] = -p[:, 1]
    p *= el["mag"]
    a = math.radians(el["angle"])
    c, s = math.cos(a), math.sin(a)
    p = np.stack([c * p[:, 0] - s * p[:, 1], s * p[:, 0] + c * p[:, 1]], axis=1)
    return p + origin


def _flatten(name, cells, stack):
    if name in stack:
        raise ValueError(f"GDSII structure {name} references itself")
    cell = cells[name]
    out = {}
    for spec, pts in cell["polys"]:
        out.setdefault(spec, []).append(pts)
    for el in cell["refs"]:
        sub = _flatten(el["sname"], cells, stack + (name,)) if el.get("sname") in cells else {}
        xy = el["xy"]
        if el["type"] == SREF:
            origins = [xy[0]]
        else:
            cols, rows = el["colrow"]
            dc = (xy[1] - xy[0]) / cols
            dr = (xy[2] - xy[0]) / rows
            origins = [xy[0] + i * dc + j * dr for j in range(rows) for i in range(cols)]
        for o in origins:
            for spec, polys in sub.items():
                out.setdefault(spec, []).extend(_transform(p, el, o) for p in polys)
    return out
