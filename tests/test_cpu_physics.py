"""The oracle pinned by the physics its passes must obey, on small photon
lists (CPU only):

- the HRMA aperture area and shell sampling follow from the shell radii;
- an ideal Wolter-I mirror brings on-axis rays to the focus;
- diffracted photons obey the grating equation: at the focal plane the
  dispersion distance is m lambda / d times the Rowland diameter;
- detected photons lie on their chip plane and their pixel coordinates are
  the in-plane offsets divided by the pixel size;
- the op-counting build of the oracle computes exactly what the plain build
  computes (its counts are of the same arithmetic)."""
import ctypes

import numpy as np
import pytest

from marx_amd import Photons, Tracer, on_axis_photons
from oracle.oracle import Oracle, OrcCtx
from cfgview import ConfigView

HC_KEV_ANGSTROM = 12.398419843320026


def _focal_plane(x, p):
    t = -x[:, 0] / p[:, 0]
    return x[:, 1] + t * p[:, 1], x[:, 2] + t * p[:, 2]


def test_aperture_area_and_shell_sampling():
    t = Tracer(GratingType="NONE", DetectorType="NONE")
    h = ConfigView(t).cfg.hrma
    areas = np.array([h.shell[k].num_open_shutters / 4.0 * np.pi
                      * (h.shell[k].max_radius ** 2 - h.shell[k].min_radius ** 2) for k in range(4)])
    assert t.geometric_area == pytest.approx(areas.sum() / 100.0, rel=1e-12)     # mm^2 -> cm^2
    frac = [h.shell[k].area_fraction for k in range(4)]
    np.testing.assert_allclose(frac, np.cumsum(areas) / areas.sum(), rtol=1e-12)

    n = 40000
    ph = Photons(n)
    ph.load(on_axis_photons(n, 1.49))
    Oracle(t).reflect(ph)
    a = ph.active
    sampled = (a["flags"] & 0x10) == 0            # not vignetted before the aperture
    counts = np.bincount(a["mirror_shell"][sampled], minlength=4)
    expect = areas / areas.sum() * sampled.sum()
    assert np.all(np.abs(counts - expect) < 4.0 * np.sqrt(expect))


def test_ideal_mirror_focuses_on_axis_rays():
    t = Tracer(GratingType="NONE", DetectorType="NONE", HRMA_Ideal="yes", HRMA_Use_Blur="no",
               HRMA_Use_WFold="no")
    n = 20000
    ph = Photons(n)
    ph.load(on_axis_photons(n, 1.49))
    Oracle(t).reflect(ph)
    a = ph.active
    ok = (a["flags"] & 0xFF) == 0
    assert ok.sum() > 0.3 * n
    x, p = a["x"][ok], a["p"][ok]
    np.testing.assert_allclose(np.linalg.norm(p, axis=1), 1.0, atol=1e-14)
    assert np.all(p[:, 0] < 0) and np.all(x[:, 0] > 0)       # travelling towards the focus
    y, z = _focal_plane(x, p)
    r = np.hypot(y, z)
    # the residual is the shell misalignment of the geometry file, a few um
    assert np.sqrt(np.mean(r ** 2)) < 0.03 and r.max() < 0.1


def test_grating_equation_at_the_focal_plane():
    energy = 1.0
    t = Tracer(GratingType="HETG", DetectorType="NONE", HRMA_Ideal="yes", HRMA_Use_Blur="no",
               HRMA_Use_WFold="no")
    v = ConfigView(t)
    n = 60000
    ph = Photons(n)
    ph.load(on_axis_photons(n, energy))
    o = Oracle(t)
    o.reflect(ph)
    o.diffract(ph)
    a = ph.active
    ok = (a["flags"] & 0xFF) == 0
    y, z = _focal_plane(a["x"][ok], a["p"][ok])
    order = a["order"][ok].astype(int)
    shell = a["mirror_shell"][ok].astype(int)
    lam = HC_KEV_ANGSTROM / energy
    assert set(np.unique(order)) >= {-1, 0, 1}
    for s in range(4):
        g = v.cfg.grating.g[s]
        sel0 = (shell == s) & (order == 0)
        y0, z0 = y[sel0].mean(), z[sel0].mean()
        for m in (-3, -1, 1, 3):
            sel = (shell == s) & (order == m)
            if sel.sum() < 20:
                continue
            dist = np.hypot(y[sel] - y0, z[sel] - z0)
            sin_beta = m * lam / (g.period * 1e4)           # period in um
            want = v.cfg.grating.rowland[s] * abs(sin_beta)
            assert np.median(dist) == pytest.approx(want, rel=5e-3)
            # dispersion direction: the grating's clocking angle
            ang = np.arctan2(np.median(z[sel] - z0), np.median(y[sel] - y0))
            assert abs(np.sin(ang - g.dispersion_angle)) < 0.02


def test_detected_photons_lie_on_their_chip():
    t = Tracer(GratingType="NONE", DetectorType="ACIS-S", DitherModel="NONE")
    d = ConfigView(t).cfg.det
    chips = {d.chip[k].id: d.chip[k] for k in range(d.num_chips)}
    assert sorted(chips) == [4, 5, 6, 7, 8, 9]
    for g in chips.values():
        xh, yh, nn = (np.array(g.xhat), np.array(g.yhat), np.array(g.normal))
        assert np.dot(xh, xh) == pytest.approx(1) and np.dot(yh, yh) == pytest.approx(1)
        # axes from the CALDB corner coordinates (not exactly orthogonal);
        # the normal is perpendicular to both
        assert abs(np.dot(xh, yh)) < 1e-3
        assert abs(np.dot(xh, nn)) < 1e-12 and abs(np.dot(yh, nn)) < 1e-12
        assert g.x_pixel_size == pytest.approx(0.024, rel=1e-3)
        assert g.xlen == pytest.approx(1024 * g.x_pixel_size, rel=1e-12)
    n = 20000
    ph = Photons(n)
    ph.load(on_axis_photons(n, 1.49))
    o = Oracle(t)
    o.reflect(ph)
    before = ph.active.copy()
    o.detect(ph)
    a = ph.active
    det = (a["flags"] & 0xFF) == 0
    assert det.sum() > 1000
    m = np.array(d.matrix).reshape(3, 3)
    off = np.array(d.offset)
    for i in np.nonzero(det)[0][:500]:
        g = chips[int(a["ccd_num"][i])]
        xd = m @ (a["x"][i] - off)                     # hit point in the detector frame
        r = xd - np.array(g.x_ll)
        assert abs(np.dot(r, g.normal)) < 1e-9
        assert a["y_pixel"][i] == pytest.approx(np.dot(r, g.xhat) / g.x_pixel_size, abs=1e-3)
        assert a["z_pixel"][i] == pytest.approx(np.dot(r, g.yhat) / g.y_pixel_size, abs=1e-3)
        # the hit is on the incoming ray, unless the frame-transfer streak
        # (acis-i.c:60-92) moved the photon to a random row
        if a["flags"][i] & 0x200:
            continue
        dx = a["x"][i] - before["x"][i]
        c = np.cross(dx, before["p"][i])
        assert np.linalg.norm(c) < 1e-9 * max(1.0, np.linalg.norm(dx))


def test_opcount_build_computes_what_the_oracle_computes():
    from oracle.opcount.opcount import load_opcount
    t = Tracer(GratingType="HETG", DetectorType="ACIS-S", DitherModel="NONE")
    cfg, fa, _, da, _ = t.config_pointers()
    ctx = OrcCtx(cfg, fa, da, 17, 0)
    plain = Oracle(t, seed=17)
    opc = load_opcount()
    rec = on_axis_photons(5000, np.random.default_rng(0).uniform(0.4, 8.0, 5000))
    a, b = Photons(5000), Photons(5000)
    a.load(rec)
    b.load(rec)
    plain.reflect(a)
    plain.diffract(a)
    plain.detect(a)
    for fn in (opc.orc_mirror_reflect, opc.orc_grating_diffract, opc.orc_detect):
        assert fn(ctypes.byref(ctx), ctypes.cast(b.ptr, ctypes.c_void_p)) == 0
    assert a.active.tobytes() == b.active.tobytes()


def test_parallel_oracle_is_independent_of_thread_count():
    """The CPU baseline's OpenMP traces (chunks, and one task per source of a
    batch) give the same detections on 1 and 8 threads: nothing is shared
    between concurrently traced chunks."""
    import tempfile
    from marx_amd.workloads import batched_acis_i, hetg_acis_s
    with tempfile.TemporaryDirectory() as tmp:
        t = hetg_acis_s(tmp, photons=1e6)
        o = Oracle(t)
        assert o.trace_parallel(200000, 20000, 1) == o.trace_parallel(200000, 20000, 8)
        t, dirs, counts = batched_acis_i(tmp, num_sources=64, max_count=3000)
        o = Oracle(t)
        one = o.trace_batch_parallel(dirs, counts, 0, 1.0e4, 1)
        assert one > 0 and one == o.trace_batch_parallel(dirs, counts, 0, 1.0e4, 8)
