"""HRC-S chip pixels (hrc_s_geom.c:344-394, hrc-s.c): a photon that hits
segment k at in-plane offset (dx, dy) mm from the segment's lower-left
active corner has u = U_Active[0] + dx / U_Pixel_Size, v = V_Active[2k] +
dy / V_Pixel_Size and chip pixel (CX0 + u - U0, CY0 + v - V0) with the
segment's *_UV and *_CXCY, all read here from hrc_s_geom.txt.  Without the
blur (DetIdeal) the detector must give exactly that map."""
import os

import numpy as np

from marx_amd import Tracer
from marx_amd._lib import DATA_DIR
from tests.test_cpu_fef import _detect, _hits


def _geom():
    vals = {}
    for line in open(os.path.join(DATA_DIR, "hrc", "hrc_s_geom.txt")):
        w = line.split()
        if w and not w[0].startswith("#"):
            nums = []
            for x in w[1:]:
                try:
                    nums.append(float(x))
                except ValueError:
                    break
            vals[w[0]] = nums
    return vals


def test_hrc_s_pixels_are_the_segment_map():
    g = _geom()
    upix, vpix = g["U_Pixel_Size"][0], g["V_Pixel_Size"][0]
    t = Tracer(seed=5, GratingType="NONE", DetectorType="HRC-S", DitherModel="NONE", DetIdeal=True,
               **{"HRC-HESF": False})
    rng = np.random.default_rng(2)
    n = 2000
    for k, name in enumerate(("Left", "Middle", "Right")):
        u0, v0 = g[name + "_UV"]
        cx0, cy0 = g[name + "_CXCY"]
        ulo = g["U_Active"][0]
        vlo, vhi = g["V_Active"][2 * k], g["V_Active"][2 * k + 1]
        # _hits places photons at chip pixel (px, py) of the facet in units
        # of its 6.429 um pixels from the corner
        px = rng.uniform(10, (g["U_Active"][1] - ulo) * upix / 6.429e-3 - 10, n)
        py = rng.uniform(10, (vhi - vlo) * vpix / 6.429e-3 - 10, n)
        a = _detect(t, _hits(t, k + 1, px, py, 1.0))
        ok = (a["flags"] & 0xFF) == 0
        assert ok.mean() > 0.99 and (a["ccd_num"][ok] == k + 1).all()
        dx, dy = px * 6.429e-3, py * 6.429e-3
        u = ulo + dx / upix
        v = vlo + dy / vpix
        np.testing.assert_allclose(a["u_pixel"][ok], u[ok], rtol=0, atol=2e-3)
        np.testing.assert_allclose(a["v_pixel"][ok], v[ok], rtol=0, atol=5e-3)
        np.testing.assert_allclose(a["y_pixel"][ok], (cx0 + u - u0)[ok], rtol=0, atol=2e-3)
        np.testing.assert_allclose(a["z_pixel"][ok], (cy0 + v - v0)[ok], rtol=0, atol=5e-3)


def test_hrc_i_pixels_are_the_plate_map():
    """HRC-I (hrc_i_geom.c:158-170): pixel = LL_CXCY + offset / pixel size,
    LL_CXCY from the centre and aimpoint pixels and the plate's size."""
    vals = {}
    for line in open(os.path.join(DATA_DIR, "hrc", "hrc_i_geom.txt")):
        w = line.split()
        if w and not w[0].startswith("#"):
            try:
                vals[w[0]] = [float(x) for x in w[1:4] if x[0] in "-0123456789."]
            except ValueError:
                pass
    t = Tracer(seed=5, GratingType="NONE", DetectorType="HRC-I", DitherModel="NONE", DetIdeal=True)
    from tests.cfgview import ConfigView
    d = ConfigView(t).cfg.det
    ll = (d.chip[0].xpixel_offset, d.chip[0].ypixel_offset)
    xp, yp = vals["HRC_I_X_Pixel_Size"][0], vals["HRC_I_Y_Pixel_Size"][0]
    rng = np.random.default_rng(3)
    n = 2000
    px, py = rng.uniform(100, 9000, n), rng.uniform(100, 9000, n)
    a = _detect(t, _hits(t, 0, px, py, 1.0))
    ok = (a["flags"] & 0xFF) == 0
    assert ok.mean() > 0.99
    np.testing.assert_allclose(a["y_pixel"][ok], (ll[0] + px * 6.429e-3 / xp)[ok], rtol=0, atol=5e-3)
    np.testing.assert_allclose(a["z_pixel"][ok], (ll[1] + py * 6.429e-3 / yp)[ok], rtol=0, atol=5e-3)
    # the corner pixel: the centre pixel less half the plate's extent
    cx = vals["HRC_I_Center_CXCY"]
    assert ll[0] < cx[0] < ll[0] + 16384 and ll[1] < cx[1] < ll[1] + 16384
