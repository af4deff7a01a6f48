"""GDSII reader/writer and Design.export_gds / import_gds (beamz/design/core.py:445-553).

The reference goes through gdspy, which is absent here, and ships no GDS file, so the
format is pinned by what the GDSII specification fixes: exactly representable reals,
the UNITS bytes that micron/nanometre files carry, record bytes written out by hand,
and reference transforms whose results are worked out by hand below.
"""
import struct

import numpy as np
import pytest

import beamz_amd as bz
from beamz_amd.design import gds


def test_real8_known_values():
    assert gds.encode_real8(0.0) == bytes(8)
    assert gds.encode_real8(1.0).hex() == "4110000000000000"        # 1/16 * 16^1
    assert gds.encode_real8(0.5).hex() == "4080000000000000"
    assert gds.encode_real8(-2.0).hex() == "c120000000000000"
    assert gds.encode_real8(1.0 / 16).hex() == "4010000000000000"
    assert gds.encode_real8(255.0).hex() == "42ff000000000000"
    # UNITS of a micron / nanometre library as common layout tools write them
    assert gds.decode_real8(bytes.fromhex("3E4189374BC6A7EF")) == 1e-3
    assert gds.decode_real8(bytes.fromhex("3944B82FA09B5A51")) == pytest.approx(1e-9, rel=3e-16)
    assert gds.decode_real8(bytes.fromhex("4110000000000000")) == 1.0
    assert gds.decode_real8(bytes.fromhex("C120000000000000")) == -2.0


def test_real8_round_trip_is_exact():
    rng = np.random.default_rng(3)
    vals = rng.standard_normal(2000) * 10.0 ** rng.integers(-70, 70, 2000)
    for v in vals:
        b = gds.encode_real8(float(v))
        assert 1 <= b[0] & 0x7F <= 127 and b[1] >> 4 != 0          # normalised: leading hex digit non-zero
        assert gds.decode_real8(b) == v


def _parse(path):
    data = open(path, "rb").read()
    recs, p = [], 0
    while p < len(data):
        n, rt, dt = struct.unpack(">HBB", data[p:p + 4])
        recs.append((rt, dt, data[p + 4:p + n]))
        p += n
    return recs


def test_record_layout(tmp_path):
    f = tmp_path / "tri.gds"
    gds.write_gds(f, [(3, 1, [(0.0, 0.0), (1.5, 0.0), (0.25, 2.0)], [])])
    raw = open(f, "rb").read()
    assert raw[:6] == bytes.fromhex("000600020258")                 # HEADER, int16, version 600
    recs = _parse(f)
    kinds = [r[0] for r in recs]
    assert kinds == [0x00, 0x01, 0x02, 0x03, 0x05, 0x06, 0x08, 0x0D, 0x0E, 0x10, 0x11, 0x07, 0x04]
    units = recs[3][2]
    assert gds.decode_real8(units[:8]) == 1e-3 and gds.decode_real8(units[8:]) == 1e-9
    assert struct.unpack(">h", recs[7][2])[0] == 3 and struct.unpack(">h", recs[8][2])[0] == 1
    xy = struct.unpack(">8i", recs[9][2])                           # closed: 4 points, database unit 1 nm
    assert xy == (0, 0, 1500, 0, 250, 2000, 0, 0)
    assert recs[5][2].rstrip(b"\0") == b"main"


def _area(pts):
    p = np.asarray(pts, dtype=np.float64)
    return 0.5 * abs(np.dot(p[:, 0], np.roll(p[:, 1], -1)) - np.dot(p[:, 1], np.roll(p[:, 0], -1)))


def test_keyhole_holes(tmp_path):
    outer = [(0, 0), (10, 0), (10, 10), (0, 10)]
    holes = [[(6, 6), (8, 6), (8, 8), (6, 8)], [(2, 2), (2, 4), (4, 4), (4, 2)]]
    kh = gds.keyhole(outer, holes)
    assert _area(kh) == pytest.approx(100 - 4 - 4)
    f = tmp_path / "holes.gds"
    gds.write_gds(f, [(0, 0, outer, holes)])
    _, cells, skipped = gds.read_gds(f)
    (pts,) = cells["main"][(0, 0)]
    assert skipped == 0 and _area(pts) == pytest.approx(92.0)
    for h in holes:
        for v in h:
            assert any(np.allclose(v, q) for q in pts)


def _lib(cells):
    """A library written record by record: cells = {name: [element record lists]}."""
    r = gds._record
    out = [r(0x00, 2, gds._ints16(600)), r(0x01, 2, gds._ints16(*([2000, 1, 1, 0, 0, 0] * 2))),
           r(0x02, 6, gds._string("LIB")), r(0x03, 5, gds.encode_real8(1e-3) + gds.encode_real8(1e-9))]
    for name, els in cells.items():
        out += [r(0x05, 2, gds._ints16(*([2000, 1, 1, 0, 0, 0] * 2))), r(0x06, 6, gds._string(name))]
        for el in els:
            out += el + [r(0x11)]
        out.append(r(0x07))
    out.append(r(0x04))
    return b"".join(out)


def test_sref_aref_flattening(tmp_path):
    r = gds._record
    square = [r(0x08), r(0x0D, 2, gds._ints16(5)), r(0x0E, 2, gds._ints16(0)),
              r(0x10, 3, gds._ints32([0, 0, 1000, 0, 1000, 1000, 0, 1000, 0, 0]))]
    # reflected about x, magnified 2x, rotated 90 degrees, placed at (10, 20) um:
    # (x, y) -> (x, -y) -> (2x, -2y) -> (2y, 2x) -> (10 + 2y, 20 + 2x)
    sref = [r(0x0A), r(0x12, 6, gds._string("A")), r(0x1A, 1, struct.pack(">H", 0x8000)),
            r(0x1B, 5, gds.encode_real8(2.0)), r(0x1C, 5, gds.encode_real8(90.0)),
            r(0x10, 3, gds._ints32([10000, 20000]))]
    # 2 columns at a 3 um pitch, 3 rows at a 4 um pitch, from (-50, 0) um
    aref = [r(0x0B), r(0x12, 6, gds._string("A")), r(0x13, 2, gds._ints16(2, 3)),
            r(0x10, 3, gds._ints32([-50000, 0, -44000, 0, -50000, 12000]))]
    path = [r(0x09), r(0x0D, 2, gds._ints16(1)), r(0x0E, 2, gds._ints16(0)), r(0x0F, 3, gds._ints32([100])),
            r(0x10, 3, gds._ints32([0, 0, 5000, 0]))]
    f = tmp_path / "refs.gds"
    f.write_bytes(_lib({"A": [square], "TOP": [sref, aref, path]}))
    unit, cells, skipped = gds.read_gds(f)
    assert unit == pytest.approx(1e-6, rel=1e-15) and skipped == 1
    assert len(cells["A"][(5, 0)]) == 1
    polys = cells["TOP"][(5, 0)]
    assert len(polys) == 7
    want = {(10, 20), (10, 22), (12, 22), (12, 20)}
    assert {tuple(np.round(p).astype(int)) for p in polys[0]} == want
    corners = sorted(tuple(np.round(p.min(axis=0), 9)) for p in polys[1:])
    assert corners == sorted((-50.0 + 3 * i, 4.0 * j) for i in range(2) for j in range(3))
    for p in polys[1:]:
        assert _area(p) == pytest.approx(1.0)

    loop = tmp_path / "loop.gds"
    loop.write_bytes(_lib({"B": [[r(0x0A), r(0x12, 6, gds._string("B")), r(0x10, 3, gds._ints32([0, 0]))]]}))
    with pytest.raises(ValueError, match="references itself"):
        gds.read_gds(loop)


def _mmi():
    """The multimode-interference layout of the reference's design-export example (examples/5_design_export.py)."""
    um = 1e-6
    X, Y, W, H, off, wg = 20 * um, 10 * um, 9 * um, 3.5 * um, 1.05 * um, 0.565 * um
    core, clad = bz.Material(2.04 ** 2), bz.Material(1.444 ** 2)
    d = bz.Design(width=X, height=Y, material=clad, pml_size=1.55 * um)
    d += bz.Rectangle(position=(0, Y / 2 - wg / 2), width=X / 2, height=wg, material=core)
    d += bz.Rectangle(position=(X / 2, Y / 2 + off - wg / 2), width=X / 2, height=wg, material=core)
    d += bz.Rectangle(position=(X / 2, Y / 2 - off - wg / 2), width=X / 2, height=wg, material=core)
    d += bz.Rectangle(position=(X / 2 - W / 2, Y / 2 - H / 2), width=W, height=H, material=core)
    return d


def test_design_export_import(tmp_path, capsys):
    d = _mmi()
    d += bz.Monitor(start=(2e-6, 3e-6), end=(2e-6, 7e-6))
    f = tmp_path / "mmi.gds"
    d.export_gds(str(f))
    assert "with 2 material-based layers" in capsys.readouterr().out
    g = bz.Design.import_gds(str(f))
    assert sorted(g.layers) == [0, 1]
    assert len(g.layers[0]) == 1 and len(g.layers[1]) == 4          # background; four core rectangles
    assert g.is_3d and g.depth == 1e-6 and len(g.structures) >= 5
    src = [s for s in d.structures if not getattr(s, "is_pml", False) and isinstance(s, bz.Rectangle)]
    for got, ref in zip(g.layers[0] + g.layers[1], src):
        a = np.asarray(got.vertices)[:, :2]
        b = np.asarray(ref.vertices)[:, :2]
        np.testing.assert_allclose(a, b, rtol=0, atol=0.5e-9 + 1e-15)   # rounded to the 1 nm database unit
        assert got.depth == 1e-6


def test_unify_polygons_needs_shapely():
    assert _mmi().unify_polygons() is False
