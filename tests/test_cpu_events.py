"""The oracle of the marx2fits event coordinates (oracle/orc_event.c,
oracle/subpix.py), pinned by what the geometry and the tables fix:

- the subpixel interpolation reproduces the CALDB table at its nodes, halves
  between them and extends the end segments (acis_subpix.c:283-322);
- the tiled system lays the chips out as acis_geom.c:104-109 describes
  (ACIS-I 01/23 with chips 0, 2 turned one way and 1, 3 the other; ACIS-S
  side by side), each chip onto a disjoint 1024 x 1024 tile, and the three
  HRC-S segments one after the other along the tiled y axis (AXAF-HRC-2.7S);
- a focal-plane pixel of the chip pixel of a simulated photon is the
  projection of that photon's own position through the mirror node;
- undithering puts the events of a dithered on-axis source back on the sky
  position of the source (the centre of each detector's focal-plane
  system) while their detector pixels spread over the dither pattern, for
  ACIS, HRC-S and HRC-I, and the undithered sky pixels of a zero-amplitude dither
  equal the roll rotation of the detector pixels (both sky formulas of
  marx2fits.c:3869-3910 agree);
- the dither state is held over a frame: the result equals per-row dither
  columns replaced by the first row of each frame;
- every flight grade is one of the four of its 3x3 subpixel.
No reference output exists for this path (the reference's CI runs marx2fits
without comparing outputs)."""
import os

import numpy as np
import pytest

from marx_amd import Tracer
from marx_amd._lib import DATA_DIR
from marx_amd.sim import Photons
from oracle.oracle import Oracle, subpix
from oracle.subpix import read_bintables, subpix_tables
from tests.cfgview import ConfigView

ON_AXIS = dict(SourceRA=250.2134679741175, SourceDEC=-53.75743813458669)
SUBPIX = os.path.join(DATA_DIR, "caldb", "acisD1999-07-22subpixN0001_marx.fits")
GRADES = [[10, 11, 138, 139], [2, 34, 130, 162], [18, 22, 50, 54], [8, 12, 136, 140], [0, 0, 0, 0],
          [16, 17, 48, 49], [72, 76, 104, 108], [64, 65, 68, 69], [80, 81, 208, 209]]


def _trace(det="ACIS-S", dither="INTERNAL", n=40000, seed=5, emin=0.5, emax=7.0, **kw):
    t = Tracer(seed=seed, source=True, GratingType="NONE", DetectorType=det, SourceType="POINT",
               SpectrumType="FLAT", MinEnergy=emin, MaxEnergy=emax, DitherModel=dither, **ON_AXIS, **kw)
    t.init_source(SourceFlux=n / (2e4 * t.geometric_area))
    o = Oracle(t, seed=seed)
    p = Photons(n)
    o.create_photons(p, n)
    o.reflect(p)
    o.diffract(p)
    o.detect(p)
    r = p.active
    return t, o, r[(r["flags"] & 0xFF) == 0].copy()


def _cols(r):
    c = {"TIME": r["arrival_time"].astype(np.float32), "TAG": r["tag"], "CCDID": r["ccd_num"],
         "CHIPX": r["y_pixel"], "CHIPY": r["z_pixel"], "B_ENERGY": r["pi"]}
    for k, name in enumerate(("RA", "DEC", "ROLL", "DET_DY", "DET_DZ", "DET_THETA")):
        c[name] = r["dither"][:, k]
    return c


@pytest.fixture
def dithered():
    # per test: the library holds one configuration, which other tests'
    # tracers replace
    return _trace()


@pytest.fixture(scope="module")
def tables():
    return subpix_tables(SUBPIX)


def test_subpix_file_layout(tables):
    tabs = read_bintables(SUBPIX)
    assert set(tabs) >= {"MARX_ACIS_SUBPIX_FI", "MARX_ACIS_SUBPIX_BI"}
    for ext in ("MARX_ACIS_SUBPIX_FI", "MARX_ACIS_SUBPIX_BI"):
        c = tabs[ext]
        assert len(c["FLTGRADE"]) == 33
        for row in range(33):
            n = int(c["NPOINTS"][row])
            if n > 0:
                assert np.all(np.diff(c["ENERGY"][row, :n]) > 0)
    subtab, off, n = tables
    # grade 0 (the centre subpixel) moves nothing
    for t in (0, 1):
        k = int(n[256 * t])
        assert k >= 2 and not subtab[off[256 * t] + k:off[256 * t] + 3 * k].any()


def test_subpix_interpolation_is_linear_in_energy(tables):
    subtab, off, num = tables
    checked = 0
    for key in np.nonzero(num)[0]:
        n = int(num[key])
        tab = subtab[off[key]:off[key] + 3 * n]
        e, dx, dy = tab[:n], tab[n:2 * n], tab[2 * n:]
        for j in range(n):
            assert subpix(tab, n, e[j]) == (pytest.approx(dx[j], abs=1e-7), pytest.approx(dy[j], abs=1e-7))
        for j in range(n - 1):
            m = np.float32(0.5 * (float(e[j]) + float(e[j + 1])))
            w = (float(m) - float(e[j])) / float(np.float32(e[j + 1] - e[j]))
            got = subpix(tab, n, m)
            assert got[0] == pytest.approx((1 - w) * dx[j] + w * dx[j + 1], abs=1e-6)
            assert got[1] == pytest.approx((1 - w) * dy[j] + w * dy[j + 1], abs=1e-6)
        # beyond the ends: the end segments extended
        for (a, b, x) in ((0, 1, float(e[0]) - 0.3), (n - 2, n - 1, float(e[-1]) + 2.0)):
            s = (x - float(e[a])) / float(np.float32(e[b] - e[a]))
            got = subpix(tab, n, x)
            assert got[0] == pytest.approx(dx[a] + s * (dx[b] - dx[a]), abs=1e-5)
        checked += 1
    assert checked == 66


def _grid_rows(ids, pix):
    """rows at chip pixels pix (k, 2) of every chip in ids."""
    ids = np.repeat(np.asarray(ids), len(pix))
    xy = np.tile(np.asarray(pix, np.float32), (len(ids) // len(pix), 1))
    n = len(ids)
    return {"TIME": np.zeros(n, np.float32), "TAG": np.arange(n, dtype=np.uint32), "CCDID": ids.astype(np.int8),
            "CHIPX": xy[:, 0], "CHIPY": xy[:, 1], "B_ENERGY": np.ones(n, np.float32)}


@pytest.mark.parametrize("det,ids", [("ACIS-S", range(4, 10)), ("ACIS-I", range(4)), ("HRC-S", range(1, 4))])
def test_tiled_pixels(det, ids):
    t = Tracer(GratingType="NONE", DetectorType=det, DitherModel="NONE")
    o = Oracle(t, seed=1)
    nx, ny = (4096, 16384) if det == "HRC-S" else (1024, 1024)
    corners = [(12.2, 12.2), (nx - 0.3, 12.2), (12.2, ny - 0.3), (nx - 0.3, ny - 0.3), (100.5, 300.5),
               (101.5, 300.5), (100.5, 301.5)]
    out = o.event_coords(t.event_cfg("EXACT"), _grid_rows(list(ids), corners))
    tx = out["tdetx"].reshape(len(ids), -1)
    ty = out["tdety"].reshape(len(ids), -1)
    boxes = {}
    for k, cid in enumerate(ids):
        # one chip pixel step moves one tiled pixel along one axis
        d1 = (tx[k, 5] - tx[k, 4], ty[k, 5] - ty[k, 4])
        d2 = (tx[k, 6] - tx[k, 4], ty[k, 6] - ty[k, 4])
        if det in ("ACIS-S", "HRC-S"):
            assert d1 == (1, 0) and d2 == (0, 1)
        elif cid in (0, 2):
            # (0,0) at the upper left, chip x running down, chip y right
            assert d1 == (0, -1) and d2 == (1, 0)
        else:
            assert d1 == (0, 1) and d2 == (-1, 0)
        boxes[cid] = (tx[k, :4].min(), tx[k, :4].max(), ty[k, :4].min(), ty[k, :4].max())
        span = sorted((boxes[cid][1] - boxes[cid][0], boxes[cid][3] - boxes[cid][2]))
        assert span == sorted((nx - 13, ny - 13))
    b = list(boxes.values())
    for i in range(len(b)):
        for j in range(i + 1, len(b)):
            assert b[i][1] < b[j][0] or b[j][1] < b[i][0] or b[i][3] < b[j][2] or b[j][3] < b[i][2]
    if det == "ACIS-I":
        # 01 over 23
        assert boxes[0][1] < boxes[1][0] and boxes[2][1] < boxes[3][0]
        assert boxes[0][2] > boxes[2][3] and boxes[1][2] > boxes[3][3]
    elif det == "HRC-S":
        # the segments follow one another along y, the same x range
        assert boxes[1][3] < boxes[2][2] and boxes[2][3] < boxes[3][2]
        assert len({boxes[c][0] for c in ids}) == 1
    else:
        xs = [boxes[c][0] for c in ids]
        assert xs == sorted(xs) and len({boxes[c][2] for c in ids}) == 1


def test_detector_pixel_is_the_projection_of_the_photon(dithered):
    """EXACT focal-plane pixels from the chip pixels equal the projection
    through the mirror node of each photon's simulated position (INTERNAL
    dither moves no detector): tan = y / (F - x) with F the parameter file's
    focal length and 0.492 arcsec pixels, x taken from the chip's corner
    plane, as marx_init_chip_to_mnc puts the chip origin at x = 0
    (pixlib.c:158 drops x_ll.x).  The chip axes of the CALDB corners are
    orthogonal only to |xhat . yhat| (4e-5 for two chips), which the chip
    pixels, projections onto those axes, carry over."""
    t, o, r = dithered
    ec = t.event_cfg("EXACT")
    out = o.event_coords(ec, _cols(r))
    d = ConfigView(t).cfg.det
    chips = {d.chip[k].id: d.chip[k] for k in range(d.num_chips)}
    x = r["x"]
    xll = np.array([chips[int(c)].x_ll[0] for c in r["ccd_num"]])
    eps = np.array([abs(np.dot(chips[int(c)].xhat[:], chips[int(c)].yhat[:])) for c in r["ccd_num"]])
    F = ec.focal_length
    delta = 0.492 * np.pi / (180 * 3600)
    detx = 4096.5 - (x[:, 1] / (x[:, 0] - xll - F)) / delta
    dety = 4096.5 + (x[:, 2] / (x[:, 0] - xll - F)) / delta
    # the chip pixel is a float: 1e-4 pixel
    tol = 2e-3 + 1100 * eps
    assert (np.abs(out["detx"] - detx) < tol).all()
    assert (np.abs(out["dety"] - dety) < tol).all()
    # and the corner's x offset matters where the chip is tilted out of x = 0
    far = np.abs(xll) > 0.01
    assert far.any() and (np.abs(4096.5 - (x[:, 1] / (x[:, 0] - F)) / delta - out["detx"])[far] > 1e-3).any()


def test_sky_pixels_undo_the_dither(dithered):
    t, o, r = dithered
    for mode in ("EXACT", "EDSER"):
        out = o.event_coords(t.event_cfg(mode), _cols(r))
        sx, sy = out["x"], out["y"]
        assert np.std(out["detx"]) > 8 and np.std(out["dety"]) > 8
        # the on-axis source at the nominal pointing: the sky position of the
        # optical axis, to the PSF core
        assert abs(np.median(sx) - 4096.5) < 0.5 and abs(np.median(sy) - 4096.5) < 0.5
        core = np.hypot(sx - np.median(sx), sy - np.median(sy))
        assert np.percentile(core, 50) < 1.5


@pytest.mark.parametrize("det,centre,pix", [("HRC-S", 32768.5, 0.1318), ("HRC-I", 16384.5, 0.13175)])
def test_hrc_sky_pixels_undo_the_dither(det, centre, pix):
    """The same for HRC, in its own focal-plane system (detpix.c:50-61):
    0.13 arcsec pixels about the centre pixel, the 16 arcsec dither being
    about 120 of them."""
    t, o, r = _trace(det=det, n=60000, emin=0.5, emax=2.0)
    assert len(r) > 2000
    for mode in ("EXACT", "EDSER"):
        out = o.event_coords(t.event_cfg(mode), _cols(r))
        assert (out["fltgrade"] == -1).all()
        assert np.std(out["detx"]) > 8 / pix * 0.5 and np.std(out["dety"]) > 8 / pix * 0.5
        sx, sy = out["x"], out["y"]
        assert abs(np.median(sx) - centre) < 2 and abs(np.median(sy) - centre) < 2
        core = np.hypot(sx - np.median(sx), sy - np.median(sy))
        # the HRMA core and the HRC blur: about half an arcsec
        assert np.percentile(core, 50) < 0.6 / pix


def test_both_sky_formulas_agree(dithered):
    """With zero pointing offsets the undithering path reduces to the roll
    rotation that marx2fits applies to undithered simulations (a sign or
    axis slip in either shows as pixels of difference)."""
    t, o, r = dithered
    c = _cols(r)
    for k in ("RA", "DEC", "DET_DY", "DET_DZ", "DET_THETA"):
        c[k] = np.zeros_like(c[k])
    roll_deg = t.par.get_double("Roll_Nom")
    c["ROLL"] = np.full_like(c["ROLL"], np.float32(np.radians(roll_deg)))
    ec = t.event_cfg("EXACT")
    a = o.event_coords(ec, c)
    ec.dither_used = 0
    ec.roll_nom = float(np.degrees(np.float32(np.radians(roll_deg))))
    b = o.event_coords(ec, c)
    assert np.array_equal(a["detx"], b["detx"])
    # the rotation acts on tangent-plane pixels, the undithering on angles:
    # they part at third order in the off-axis angle; and marx_mnc_to_ra_dec
    # takes the angles as acos of numbers near 1, which resolves them to
    # about 1.5e-8 rad (0.006 pixel) near the axis
    delta = 0.492 * np.pi / (180 * 3600)
    th = np.hypot(a["detx"] - 4096.5, a["dety"] - 4096.5) * delta
    tol = 0.01 + th ** 3 / delta
    assert (np.abs(a["x"] - b["x"]) < tol).all() and (np.abs(a["y"] - b["y"]) < tol).all()
    assert (th > 1e-3).any()


def test_dither_is_held_over_a_frame(dithered):
    t, o, r = dithered
    c = _cols(r)
    ec = t.event_cfg("NONE")
    held = o.event_coords(ec, c)
    # the same with each row's dither columns replaced by its frame's first
    # row and a frame per row
    frame = (c["TIME"].astype(np.float64) / (ec.exposure_time + ec.frame_transfer_time)).astype(np.int64)
    first = np.r_[0, np.nonzero(np.diff(frame))[0] + 1]
    head = first[np.searchsorted(first, np.arange(len(frame)), side="right") - 1]
    assert len(first) > 100 and (np.diff(first) > 1).any()
    c2 = dict(c)
    for k in ("RA", "DEC", "ROLL", "DET_DY", "DET_DZ", "DET_THETA"):
        c2[k] = c[k][head]
    ec.exposure_time, ec.frame_transfer_time = 0.0, 0.0
    own = o.event_coords(ec, c2)
    for k in ("detx", "dety", "x", "y", "tdetx", "fltgrade"):
        assert np.array_equal(held[k], own[k]), k
    # and per-row dither differs
    assert not np.array_equal(held["x"], o.event_coords(ec, c)["x"])


def test_flight_grades_follow_the_subpixel(dithered):
    t, o, r = dithered
    out = o.event_coords(t.event_cfg("EDSER"), _cols(r))
    cx = r["y_pixel"] + np.float32(1)
    cy = r["z_pixel"] + np.float32(1)
    sub = 3 * (3 * (cy - np.floor(cy))).astype(int) + (3 * (cx - np.floor(cx))).astype(int)
    g = out["fltgrade"]
    for s in range(9):
        sel = sub == s
        assert np.isin(g[sel], GRADES[s]).all()
        if s != 4 and sel.sum() > 400:
            counts = np.array([(g[sel] == v).sum() for v in GRADES[s]])
            assert counts.min() > 0.15 * sel.sum()
