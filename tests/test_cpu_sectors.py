"""The grating facet-sector tables (diffract.c:1900-1948; the file format is
documented in the files' own headers): per shell, sectors from a start to
an end azimuth (degrees from +Z of the unrotated torus), an alignment error
a + da g (arc minutes) and a period scatter p + dp g.  The library's tables
must hold each file's rows in order, converted to radians, as numpy reads
the same text; the diffracted directions must carry, per row, clocking and
period errors distributed as that row says."""
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


def _rot_x(v, a):
    c, s = np.cos(a), np.sin(a)
    return np.stack([v[:, 0], c * v[:, 1] - s * v[:, 2], s * v[:, 1] + c * v[:, 2]], axis=1)


def test_sector_alignment_and_period_draws():
    """How the trace uses the tables (diffract.c:1900-1948 with the diffraction
    of diffract.c): a photon at azimuth atan2(y, z) of its Rowland-torus hit
    point, in the grating frame (turned by the shell's dispersion angle about
    x), takes the row whose [start, end) holds it; its facet is clocked by
    -(a + da g) and its period scaled by 1 + p + dp g.  Both are read back
    exactly from each diffracted direction: the rotated facet's l axis
    carries no change of direction, which fixes the clocking, and the change
    along its d axis is m lambda / d (1 - dp/p).  Per row the clocking
    errors are N(a, da) and the period errors N(p, dp) (KS on the pooled
    z-scores; the row means within 5 sigma); assigning each photon the
    neighbouring row fails."""
    from scipy import stats
    from marx_amd import Photons, on_axis_photons
    from oracle.oracle import Oracle
    t = Tracer(seed=17, GratingType="HETG", DetectorType="NONE", DitherModel="NONE",
               Use_HETG_Sector_Files=True)
    c = ConfigView(t)
    n, energy = 500000, 1.0
    ph = Photons(n)
    ph.load(on_axis_photons(n, energy))
    o = Oracle(t)
    o.reflect(ph)
    before = ph.active.copy()
    o.diffract(ph)
    a = ph.active
    before, a = before[np.argsort(before["tag"])], a[np.argsort(a["tag"])]
    assert np.array_equal(before["tag"], a["tag"])
    ok = ((a["flags"] & 0xFF) == 0) & (a["order"] != 0)
    hbar_c = 1.973269631e-4                     # keV um
    z_th, z_p, z_shift = [], [], []
    for i in range(4):
        g = c.cfg.grating.g[i]
        sel = ok & (a["mirror_shell"] == i)
        assert sel.sum() > 1500
        assert 0.1 < g.period < 0.5                  # um: 2000 and 4000 Angstrom
        alpha = g.dispersion_angle
        x = _rot_x(np.asarray(a["x"][sel]), -alpha)
        dp = _rot_x(np.asarray(a["p"][sel]), -alpha) - _rot_x(np.asarray(before["p"][sel]), -alpha)
        nn = -x / np.linalg.norm(x, axis=1)[:, None]
        ll = np.stack([nn[:, 2], np.zeros(len(x)), -nn[:, 0]], axis=1)
        ll /= np.linalg.norm(ll, axis=1)[:, None]
        dd = np.cross(nn, ll)
        dtheta = np.arctan(np.sum(dp * ll, axis=1) / np.sum(dp * dd, axis=1))
        d_rot = np.sin(dtheta)[:, None] * ll + np.cos(dtheta)[:, None] * dd
        m_lambda_d = a["order"][sel] * (2 * np.pi * hbar_c / energy) / g.period
        dp_over_p = -(np.sum(dp * d_rot, axis=1) - m_lambda_d) / m_lambda_d
        sec = c.cfg.grating.sectors[i]
        lo, hi, th, thb, pp, ppb = c.d(sec.off, 6 * sec.n).reshape(6, sec.n)
        az = np.mod(np.arctan2(x[:, 1], x[:, 2]), 2 * np.pi)
        k = np.searchsorted(lo, az, "right") - 1
        inside = (k >= 0) & (az < hi[np.maximum(k, 0)])
        assert inside.mean() > 0.99
        k, dtheta, dp_over_p = k[inside], dtheta[inside], dp_over_p[inside]
        for r in np.unique(k):
            s = k == r
            if s.sum() >= 200:
                assert abs(dtheta[s].mean() - th[r]) < 5 * thb[r] / np.sqrt(s.sum()), (i, r)
                assert abs(dp_over_p[s].mean() - pp[r]) < 5 * ppb[r] / np.sqrt(s.sum()), (i, r)
        z_th.append((dtheta - th[k]) / thb[k])
        z_p.append((dp_over_p - pp[k]) / ppb[k])
        kn = (k + 1) % sec.n
        z_shift.append((dtheta - th[kn]) / thb[kn])
    p_th = stats.kstest(np.concatenate(z_th), "norm").pvalue
    p_p = stats.kstest(np.concatenate(z_p), "norm").pvalue
    p_shift = stats.kstest(np.concatenate(z_shift), "norm").pvalue
    assert p_th > 1e-3 and p_p > 1e-3 and p_shift < 1e-10, (p_th, p_p, p_shift)
