"""The grating facet-sector tables (diffract.c:1900-1948; the file format is
documented in the files' own headers): per shell, sectors from a start to
an end azimuth (degrees from +Z of the unrotated torus), an alignment error
a + da g (arc minutes) and a period scatter p + dp g.  The library's tables
must hold each file's rows in order, converted to radians, as numpy reads
the same text."""
import os

import numpy as np
import pytest

from marx_amd import Tracer
from marx_amd._lib import DATA_DIR
from tests.cfgview import ConfigView

SHELLS = "1346"


# the LETG facet files that marx.par names are not part of the reference's
# data (Use_LETG_Sector_Files is off by default): HETG only
@pytest.mark.parametrize("grating", ["HETG"])
def test_sector_tables(grating):
    t = Tracer(GratingType=grating, DetectorType="NONE", DitherModel="NONE",
               **{f"Use_{grating}_Sector_Files": True})
    c = ConfigView(t)
    for i, s in enumerate(SHELLS):
        name = t.par.get_string(f"{grating}_Sector{s}_File")
        rows = np.loadtxt(os.path.join(DATA_DIR, name), comments="#", ndmin=2)
        sec = c.cfg.grating.sectors[i]
        assert sec.n == len(rows) > 0
        got = c.d(sec.off, 6 * sec.n).reshape(6, sec.n)
        scale = [np.pi / 180, np.pi / 180, np.pi / 180 / 60, np.pi / 180 / 60, 1.0, 1.0]
        for k in range(6):
            np.testing.assert_allclose(got[k], rows[:, k] * scale[k], rtol=1e-15, atol=0,
                                       err_msg=f"{grating} shell {s} column {k}")
        # sectors are ordered, non-overlapping azimuth ranges
        assert np.all(got[1] > got[0])
        assert np.all(got[0][1:] >= got[1][:-1] - 1e-12)
