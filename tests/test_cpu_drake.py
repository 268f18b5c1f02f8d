"""The HESF ("Drake flat", drake.c:170-260 and 268-369) ahead of HRC-S
detection, pinned from what its parameters mean rather than from its code:

- plate k (HESFHeightk, HESFThetak) is a strip of the plane through the
  corner (x_k, +-gap, z_k) from the HRC geometric centre (the midpoint of the
  middle MCP, hrc-s.c:440-444), x_k stacking the heights upward from
  HESFOffsetX and z_k falling by h tan(theta) per plate from HESFOffsetZ, over
  HESFLength in |y| beyond the gap.  A ray along -x at height z therefore
  hits the one plate whose z band (z_k - h_k tan theta_k, z_k] holds z, at
  x = x_k + (z_k - z) / tan theta_k;
- a reflected ray leaves along (-cos 2 theta_k, 0, sin 2 theta_k) from that
  point (mirror reflection off a plane tilted by theta_k);
- it is reflected with the Fresnel probability at grazing angle theta_k of
  the carbon coating, or of the chromium strip within HESFCrWidth of the
  gap edge, with beta and delta interpolated in energy from the optical
  constant files; blocked photons stay at the hit point;
- rays between the plates, or above the top plate, are not touched."""
import os

import numpy as np

from marx_amd import Tracer
from marx_amd._lib import ATTR_DTYPE, DATA_DIR
from tests.cfgview import ConfigView
from tests.test_cpu_fef import _detect
from tests.test_cpu_math import _fresnel
from tests.test_cpu_support_gratings import _bdat

BLOCKED, REFLECTED = 0x20, 0x100


def _tracer():
    t = Tracer(seed=13, GratingType="NONE", DetectorType="HRC-S", DitherModel="NONE", DetIdeal=True)
    assert int(t.par.get_double("HESFN")) == 2
    return t


def _centre(t):
    d = ConfigView(t).cfg.det
    m = [d.chip[k] for k in range(d.num_chips) if d.chip[k].id == 2][0]
    return (np.array(m.x_ll) + 0.5 * (m.xlen * np.array(m.xhat) + m.ylen * np.array(m.yhat))
            + np.array(d.offset))


def _plates(t):
    """(x_k, z_k, h_k, theta_k) of each plate from the parameters."""
    g = t.par.get_double
    x, z, out = g("HESFOffsetX"), g("HESFOffsetZ"), []
    for k in range(1, int(t.par.get_double("HESFN")) + 1):
        h, th = g(f"HESFHeight{k}"), np.radians(g(f"HESFTheta{k}"))
        out.append((x, z, h, th))
        x, z = x + h, z - h * np.tan(th)
    return out


def _rays(c, y, z, energy, tag0=0):
    n = len(y)
    rec = np.zeros(n, ATTR_DTYPE)
    rec["x"] = np.stack([np.full(n, c[0] + 400.0), c[1] + y, c[2] + z], axis=1)
    rec["p"] = (-1.0, 0.0, 0.0)
    rec["energy"] = energy
    rec["tag"] = np.arange(tag0, tag0 + n)
    return rec


def test_plates_reflect_along_their_mirror_direction():
    t = _tracer()
    c, g = _centre(t), t.par.get_double
    length, crw = g("HESFLength"), g("HESFCrWidth")
    rng = np.random.default_rng(4)
    n = 3000
    for k, (xk, zk, h, th) in enumerate(_plates(t)):
        for side, gap in ((1, g("HESFGapY1")), (-1, g("HESFGapY2"))):
            u = rng.uniform(0.02, 0.98, n)
            s = rng.uniform(0.5, length - 0.5, n)
            z = zk - u * h * np.tan(th)
            a = _detect(t, _rays(c, side * (gap + s), z, 0.1))
            hit = np.stack([c[0] + xk + (zk - z) / np.tan(th), c[1] + side * (gap + s), c[2] + z], axis=1)
            refl = (a["flags"] & REFLECTED) != 0
            blk = (a["flags"] & BLOCKED) != 0
            assert not (refl & blk).any() and (refl | blk).all(), (k, side)
            assert 0.05 < refl.mean() < 0.95, (k, side, refl.mean())
            out = np.array([-np.cos(2 * th), 0.0, np.sin(2 * th)])
            np.testing.assert_allclose(a["p"][refl], np.broadcast_to(out, (refl.sum(), 3)), rtol=0, atol=1e-12)
            np.testing.assert_allclose(a["p"][blk], np.broadcast_to([-1.0, 0, 0], (blk.sum(), 3)), atol=0)
            np.testing.assert_allclose(a["x"][blk], hit[blk], rtol=0, atol=1e-9)
            # reflected photons lie on the reflected ray through the hit point
            d = a["x"][refl] - hit[refl]
            assert np.abs(np.cross(d, out)).max() < 1e-9 * max(1.0, np.abs(d).max())
            # the chromium strip reflects with its own constants; checked below
            assert ((s < crw) & refl).any() and ((s >= crw) & refl).any()


def test_rays_off_the_plates_are_untouched():
    t = _tracer()
    c, g = _centre(t), t.par.get_double
    pl = _plates(t)
    rng = np.random.default_rng(5)
    n = 2000
    gap = min(g("HESFGapY1"), g("HESFGapY2"))
    # between the plates (|y| < gap) at plate heights, and above the top plate
    y = np.concatenate([rng.uniform(-gap + 0.5, gap - 0.5, n),
                        rng.uniform(g("HESFGapY1") + 1, g("HESFGapY1") + g("HESFLength") - 1, n)])
    zlo = pl[-1][1] - pl[-1][2] * np.tan(pl[-1][3])
    z = np.concatenate([rng.uniform(zlo, pl[0][1], n), rng.uniform(pl[0][1] + 0.1, pl[0][1] + 3.0, n)])
    a = _detect(t, _rays(c, y, z, 0.1))
    assert (a["flags"] & (BLOCKED | REFLECTED) == 0).all()
    np.testing.assert_array_equal(a["p"], np.broadcast_to([-1.0, 0, 0], (2 * n, 3)))


def test_reflected_fraction_is_the_coating_reflectivity():
    """Carbon beyond HESFCrWidth from the gap edge, chromium within it; binomial
    counts within 5 sigma of the Fresnel reflectivity at theta_k."""
    t = _tracer()
    c, g = _centre(t), t.par.get_double
    crw, gap = g("HESFCrWidth"), g("HESFGapY1")
    tabs = {m: _bdat(os.path.join(DATA_DIR, "hrc", m + ".dat")) for m in ("carbon", "chromium")}
    rng = np.random.default_rng(6)
    n, tag, mid, separated = 20000, 0, 0, 0
    for k, (xk, zk, h, th) in enumerate(_plates(t)):
        for energy in (0.06, 0.1, 0.15, 0.25, 0.4, 0.7):
            e = float(np.float32(energy))
            r = {}
            for m, tab in tabs.items():
                beta = float(np.float32(np.interp(e, tab[:, 0], tab[:, 1])))
                delta = float(np.float32(np.interp(e, tab[:, 0], tab[:, 2])))
                r[m] = _fresnel(np.sin(th), beta, delta)
            sd = np.sqrt(max(r["carbon"] * (1 - r["carbon"]), 1e-4) / n)
            separated += abs(r["carbon"] - r["chromium"]) > 10 * sd
            for m, (slo, shi) in (("chromium", (0.01, crw - 0.01)), ("carbon", (crw + 0.01, 150.0))):
                z = zk - rng.uniform(0.05, 0.95, n) * h * np.tan(th)
                a = _detect(t, _rays(c, gap + rng.uniform(slo, shi, n), z, energy, tag))
                tag += n
                frac = ((a["flags"] & REFLECTED) != 0).mean()
                sig = np.sqrt(max(r[m] * (1 - r[m]), 1e-4) / n)
                assert abs(frac - r[m]) < 5 * sig + 2e-4, (k, energy, m, frac, r[m])
                mid += 0.1 < r[m] < 0.9
    # the cases have power: reflectivities away from 0 and 1, and the two
    # coatings told apart at several energies
    assert mid >= 6 and separated >= 3
