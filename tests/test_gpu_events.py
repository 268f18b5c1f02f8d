"""Event coordinates on the GPU (mxg_event_coords through the C-ABI) of ACIS
and HRC simulations against
the oracle of marx2fits (oracle/orc_event.c) run on the column files that
mxg_write_device wrote for the same trace, the way marx2fits reads them.

Bars: tags, chip ids, tiled pixels and flight grades bit-exact; focal-plane
pixels to 1e-9 pixel (the same IEEE arithmetic on both sides, no
transcendental function on that path but the cosine and sine of the dither
roll); sky pixels to 1e-7 pixel plus the conditioning of
marx_mnc_to_ra_dec, which takes both angles as acos of numbers near 1: an
ulp of difference in a cosine or sine of the undithering (GPU and CPU
libraries may round those differently) becomes up to eps / |angle| radians,
so the bar is 1e-7 + 4e-16 / (|angle| delta) pixel for the angle of each
axis, capped at 0.01 pixel."""
import os

import numpy as np
import pytest

from marx_amd import Tracer
from marx_amd._lib import DATA_DIR
from oracle import dumpfile as D
from oracle.oracle import Oracle
from oracle.subpix import subpix_tables

ON_AXIS = dict(SourceRA=250.2134679741175, SourceDEC=-53.75743813458669)
ARCSEC = np.pi / (180 * 3600)
# focal-plane systems (detpix.c:43-61): pixel size and centre
FP = {"ACIS-S": (0.492 * ARCSEC, 4096.5), "ACIS-I": (0.492 * ARCSEC, 4096.5),
      "HRC-S": (0.1318 * ARCSEC, 32768.5), "HRC-I": (0.13175 * ARCSEC, 16384.5)}
TABLES = os.path.join(DATA_DIR, "caldb", "acisD1999-07-22subpixN0001_marx.fits")
DITHER_FILES = (("RA", "sky_ra.dat"), ("DEC", "sky_dec.dat"), ("ROLL", "sky_roll.dat"),
                ("DET_DY", "det_dy.dat"), ("DET_DZ", "det_dz.dat"), ("DET_THETA", "det_theta.dat"))


def _columns(d):
    col = lambda f: D.read_column(open(os.path.join(d, f), "rb").read())[4]
    c = {"TIME": col("time.dat"), "TAG": col("tag.dat"), "CCDID": col("detector.dat"),
         "CHIPX": col("xpixel.dat"), "CHIPY": col("ypixel.dat")}
    # HRC writes no PI column (hrc-s.c:245-248); only EDSER, an ACIS mode, reads it
    c["B_ENERGY"] = (col("b_energy.dat") if os.path.exists(os.path.join(d, "b_energy.dat"))
                     else np.zeros(len(c["TAG"]), np.float32))
    for k, f in DITHER_FILES:
        if os.path.exists(os.path.join(d, f)):
            c[k] = col(f)
    return c


def _sky_tol(v, fp):
    delta, centre = fp
    ang = np.maximum(np.abs(v - centre) * delta, 1e-14)
    return 1e-7 + np.minimum(0.01, 4e-16 / (ang * delta))


def _compare(g, e, fp):
    for k in ("tag", "ccd_id", "tdetx", "tdety", "fltgrade"):
        assert np.array_equal(g[k], e[k]), k
    for k in ("detx", "dety"):
        np.testing.assert_allclose(g[k], e[k], rtol=0, atol=1e-9, err_msg=k)
    for k in ("x", "y"):
        bad = np.abs(g[k] - e[k]) > _sky_tol(e[k], fp)
        assert not bad.any(), (k, np.nonzero(bad)[0][:5], g[k][bad][:5], e[k][bad][:5])


def _run(t, tmp_path, n, modes, total_time=0.0, seed=7):
    _, ndet, _ = t.trace_device(n)
    rows = t.write_device(str(tmp_path), total_time=total_time)
    assert rows == ndet > 1000
    cols = _columns(str(tmp_path))
    o = Oracle(t)
    tables = subpix_tables(TABLES)
    for mode in modes:
        print(f"event coords {mode} rows={rows}", flush=True)
        g = t.event_coords(rows + 10, mode, seed=seed, total_time=total_time)
        e = o.event_coords(t.event_cfg(mode, seed, total_time), cols, *tables)
        assert len(g["tag"]) == rows
        _compare(g, e, FP[t.par.get_string("DetectorType")])
    return g


@pytest.mark.gpu
@pytest.mark.parametrize("mode", ["EDSER", "EXACT", "NONE", "RANDOMIZE"])
def test_hetg_acis_s_dithered(gpu, tmp_path, mode):
    t = Tracer(seed=3, GratingType="HETG", DetectorType="ACIS-S", SourceType="POINT", SpectrumType="FLAT",
               MinEnergy=0.3, MaxEnergy=9.0, DitherModel="INTERNAL", **ON_AXIS)
    t.init_source(SourceFlux=200000 / (1e4 * t.geometric_area))
    g = _run(t, tmp_path, 200000, [mode], total_time=1234.5)
    # dispersed orders spread over the chips of the array
    assert len(np.unique(g["ccd_id"])) >= 4


@pytest.mark.gpu
def test_acis_i_undithered(gpu, tmp_path):
    t = Tracer(seed=4, GratingType="NONE", DetectorType="ACIS-I", SourceType="DISK", SpectrumType="FLAT",
               MinEnergy=0.5, MaxEnergy=7.0, DitherModel="NONE",
               **{"S-DiskTheta0": 0.0, "S-DiskTheta1": 600.0}, **ON_AXIS)
    t.init_source(SourceFlux=100000 / (1e4 * t.geometric_area))
    g = _run(t, tmp_path, 100000, ["EDSER", "EXACT"])
    assert set(np.unique(g["ccd_id"])) == {0, 1, 2, 3}


@pytest.mark.gpu
def test_letg_hrc_s_dithered(gpu, tmp_path):
    """HRC-S: the three segments, AXAF-FP-2.3, dither re-read on every row,
    EDSER taken as RANDOMIZE; no flight grades."""
    t = Tracer(seed=6, GratingType="LETG", DetectorType="HRC-S", SourceType="DISK", SpectrumType="FLAT",
               MinEnergy=0.07, MaxEnergy=10.0, DitherModel="INTERNAL",
               **{"S-DiskTheta0": 0.0, "S-DiskTheta1": 120.0}, **ON_AXIS)
    t.init_source(SourceFlux=400000 / (1e4 * t.geometric_area))
    g = _run(t, tmp_path, 400000, ["EDSER", "EXACT", "NONE"], total_time=77.0)
    assert set(np.unique(g["ccd_id"])) == {1, 2, 3}
    assert (g["fltgrade"] == -1).all()


@pytest.mark.gpu
def test_hrc_i_dithered(gpu, tmp_path):
    t = Tracer(seed=8, GratingType="NONE", DetectorType="HRC-I", SourceType="POINT", SpectrumType="FLAT",
               MinEnergy=0.5, MaxEnergy=2.0, DitherModel="INTERNAL", **ON_AXIS)
    t.init_source(SourceFlux=200000 / (1e4 * t.geometric_area))
    g = _run(t, tmp_path, 200000, ["EDSER", "EXACT"])
    assert set(np.unique(g["ccd_id"])) == {0}


@pytest.mark.gpu
def test_rows_beyond_capacity_fail(gpu, tmp_path):
    t = Tracer(seed=5, GratingType="NONE", DetectorType="ACIS-S", SourceType="POINT", SpectrumType="FLAT",
               MinEnergy=1.0, MaxEnergy=2.0, DitherModel="INTERNAL", **ON_AXIS)
    t.init_source()
    _, ndet, _ = t.trace_device(20000)
    with pytest.raises(RuntimeError, match="capacity"):
        t.event_coords(ndet - 1)
    assert len(t.event_coords(ndet)["tag"]) == ndet
