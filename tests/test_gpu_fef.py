"""The ACIS FEF on the GPU (k_detect's apply_fef) against the oracle, with a
synthetic FEF file (the reference's CALDB FEF is not available): HETG +
ACIS-S and ACIS-I, QE on.  The PHA of every photon the two sides agree on
must be equal and its PI within one float ulp (the erf of the mixture
normalisation may round differently on the two sides; a photon whose
uniform lands within that difference of a component boundary can take
another component, which the flip allowance covers)."""
import numpy as np
import pytest

from marx_amd import Photons, Tracer, on_axis_photons
from oracle.oracle import Oracle
from tests.fefgen import write_fef
from tests.parity import compare


@pytest.mark.gpu
@pytest.mark.parametrize("grating,det,ccds", [("HETG", "ACIS-S", range(4, 10)), ("NONE", "ACIS-I", range(0, 4))])
def test_fef_parity(gpu, tmp_path, grating, det, ccds):
    path = write_fef(str(tmp_path / "fef.fits"), ccds, split=256)
    t = Tracer(seed=23, GratingType=grating, DetectorType=det, DitherModel="NONE", ACISFEFFile=path,
               DetOffsetY=1.5)
    rng = np.random.default_rng(4)
    n = 200000
    rec = on_axis_photons(n, rng.uniform(0.3, 9.0, n), time_span=500.0)
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
    compare(a, b, max_flips=3, check_pixels=True, label=f"FEF {det}/{grating}")
    same = a["flags"] == b["flags"]
    det_ok = same & ((a["flags"] & 0xFF) == 0)
    assert det_ok.sum() > 1000
    assert np.array_equal(a["pulse_height"][same], b["pulse_height"][same])
    ulp = np.abs(a["pi"][det_ok].view(np.int32).astype(np.int64) - b["pi"][det_ok].view(np.int32).astype(np.int64))
    assert ulp.max() <= 1
    # the FEF rejected some photons (PI below the table) and gave PHA > 0 to the rest
    assert (a["pulse_height"][det_ok] >= 0).all() and a["pulse_height"][det_ok].max() > 100
