"""ACIS quantum efficiency (acis-s.c:138-176 over the CALDB ACISQE file,
acis_qe reading as caldb.c does), pinned by the table itself, read here
with the oracle's own FITS reader: photons of one energy of the table's
grid that hit a chip are detected with probability QE(E) of that chip's
extension AXAF_QE<ccd + 1> (binomial; the CALDB QE includes the filter,
and contamination is not applied)."""
import os

import numpy as np
import pytest

from marx_amd import Tracer
from marx_amd._lib import DATA_DIR
from oracle.subpix import read_bintables
from tests.test_cpu_fef import _detect, _hits

QE = os.path.join(DATA_DIR, "caldb", "acisD1997-04-17qeN0006.fits")


@pytest.mark.parametrize("det,ccd,target", [("ACIS-S", 7, 0.6), ("ACIS-S", 5, 1.8), ("ACIS-I", 2, 4.0)])
def test_detection_fraction_is_the_chip_qe(det, ccd, target):
    tab = read_bintables(QE)[f"AXAF_QE{ccd + 1}"]
    e = tab["ENERGY"].astype(np.float64)
    row = int(np.argmin(np.abs(e - target)))
    energy = float(np.float32(tab["ENERGY"][row]))
    qe = float(tab["QE"][row])
    assert 0.05 < qe < 0.98
    t = Tracer(seed=37, GratingType="NONE", DetectorType=det, DitherModel="NONE")
    rng = np.random.default_rng(ccd)
    n = 50000
    a = _detect(t, _hits(t, ccd, rng.uniform(20, 1000, n), rng.uniform(20, 1000, n), energy))
    assert (a["ccd_num"] == ccd).all()
    k = ((a["flags"] & 0xFF) == 0).sum()
    assert abs(k - n * qe) < 5 * np.sqrt(n * qe * (1 - qe)), (k, n * qe)
