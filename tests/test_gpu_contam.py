"""ACIS contamination on the GPU (k_detect's contam_transmission) against the
oracle, with synthetic contamination files (the reference's CALDB file is
not available): HETG + ACIS-S with the analytic row profile and with an
FXY table, ACIS-I with the radial profile.  Categorical results must agree
(a uniform within rounding of exp(-v) QE may flip, covered by the flip
allowance), positions to 1e-9 mm, pixels to the ulp."""
import numpy as np
import pytest

from marx_amd import Photons, Tracer, on_axis_photons
from oracle.oracle import Oracle
from tests.contamgen import write_contam
from tests.parity import compare


@pytest.mark.gpu
@pytest.mark.parametrize("grating,det,blocking", [("HETG", "ACIS-S", 0), ("HETG", "ACIS-S", 4),
                                                  ("NONE", "ACIS-I", 0)])
def test_contam_parity(gpu, tmp_path, grating, det, blocking):
    path = write_contam(str(tmp_path / "contam.fits"), blocking=blocking)
    t = Tracer(seed=61, GratingType=grating, DetectorType=det, DitherModel="NONE", ACISContamFile=path,
               TStart=2015.0, DetOffsetY=2.0)
    rng = np.random.default_rng(8)
    n = 200000
    rec = on_axis_photons(n, rng.uniform(0.2, 6.0, n), time_span=500.0)
    pg, po = Photons(n), Photons(n)
    pg.load(rec)
    po.load(rec)
    t.reflect(pg)
    t.diffract(pg)
    t.detect(pg)
    o = Oracle(t)
    o.reflect(po)
    o.diffract(po)
    o.detect(po)
    a, b = pg.active, po.active
    compare(a, b, max_flips=3, check_pixels=True, label=f"contam {det}/{grating}/{blocking}")
    det_ok = (a["flags"] & 0xFF) == 0
    assert det_ok.sum() > 1000
