"""Pins of the oracle's mirror pass that do not come from the oracle's own
formulas (CPU only):

- struts: every photon is flagged blocked exactly when the ray it carries to
  a strut set crosses one of that set's bands, the bands being six lines
  through the axis at multiples of 30 degrees with the half-widths and
  planes of marx/libsrc/hrma.c:907-925 (written here as azimuth bands, not
  as the reference's rotations); the pre-collimator share of each shell is
  the bands' share of its aperture annulus, a closed form;
- normal blur (hrma.c:1054-1093): with one surface blurred and nothing else
  random, the angle between blurred and unblurred reflections is twice a
  Gaussian tilt of the normal about a random in-plane axis, 2 |B g| (cos^2 a
  cos^2 psi + sin^2 a)^(1/2) for the grazing angle a (KS against that
  law);
- shell tilts (hrma.c:842-897): the rotation matrices are the rotations the
  reference's comment documents (az about -Z, then el negatively about the
  new Y), built here with scipy; and a tilt of one surface moves that
  shell's image by +-F delta in the direction reflection geometry gives
  (derived in the test's docstring, independent of the code).
"""
import os

import numpy as np
import pytest
from scipy import stats
from scipy.spatial.transform import Rotation

from marx_amd import Photons, Tracer, on_axis_photons
from oracle.oracle import Oracle
from cfgview import ConfigView

VBLOCKED, UNREFLECTED = 0x10, 0x02
INCH = 25.4
# marx/libsrc/hrma.c:907-925: strut planes (x offset from the cap) and half
# widths, in mm
STRUT_SETS = {
    "pre": ([1492.060, 942.266], 0.5 * 0.5 * INCH),
    "cap": ([0.5 * 1.965 * INCH, -0.5 * 1.965 * INCH], 0.5 * 0.75 * INCH),
    "post": ([-1050.353, -1271.333], 0.5 * 0.5 * INCH),
}
ARCSEC = np.pi / 180.0 / 3600.0
DATA = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "marx_amd", "data")


def _in_bands(y, z, half_width):
    """The point (y, z) lies within half_width of one of the six lines
    through the axis at azimuths k * 30 degrees."""
    r = np.hypot(y, z)
    phi = np.mod(np.arctan2(z, y), np.pi / 6.0)
    off = np.minimum(phi, np.pi / 6.0 - phi)          # angle to the nearest line
    return r * np.sin(off) < half_width


def _crosses_set(ray, cap, name):
    """ray (n, 6): x, p.  True where the ray's line crosses a band of one of
    the set's planes."""
    planes, hw = STRUT_SETS[name]
    hit = np.zeros(len(ray), dtype=bool)
    for xp in planes:
        t = (cap + xp - ray[:, 0]) / ray[:, 3]
        hit |= _in_bands(ray[:, 1] + t * ray[:, 4], ray[:, 2] + t * ray[:, 5], hw)
    return hit


def test_strut_shadows_follow_the_strut_bands():
    t = Tracer(seed=5, GratingType="NONE", DetectorType="NONE", HRMA_Ideal="yes")
    h = ConfigView(t).cfg.hrma
    n = 60000
    ph = Photons(n)
    ph.load(on_axis_photons(n, 1.49))
    taps = Oracle(t).reflect_with_strut_taps(ph)[:n]
    a = ph.attrs[:n]
    blocked = (a["flags"] & VBLOCKED) != 0

    # the taps are the path: each ray passes through the next one's point,
    # and an unblocked photon leaves with the last ray
    pre, cap, post = taps[:, 0], taps[:, 1], taps[:, 2]
    assert not np.isnan(pre).any()
    np.testing.assert_array_equal(pre[:, 3:], np.tile([-1.0, 0.0, 0.0], (n, 1)))
    for r0, r1 in ((pre, cap), (cap, post)):
        m = ~np.isnan(r1[:, 0])
        d = r1[m, :3] - r0[m, :3]
        miss = np.linalg.norm(np.cross(d, r0[m, 3:]), axis=1)
        assert miss.max() < 1e-6, miss.max()
    done = ~blocked & ((a["flags"] & UNREFLECTED) == 0)
    np.testing.assert_array_equal(a["x"][done], post[done, :3])

    expect = _crosses_set(pre, h.cap_position, "pre")
    m = ~np.isnan(cap[:, 0])
    expect[m] |= _crosses_set(cap[m], h.cap_position, "cap")
    m = ~np.isnan(post[:, 0])
    expect[m] |= _crosses_set(post[m], h.cap_position, "post")
    assert expect.sum() > 0.05 * n
    np.testing.assert_array_equal(blocked, expect)

    # the pre-collimator bands' share of each shell's annulus, r uniform on
    # [rmin, rmax] (hrma.c:1005-1006): 12 asin(w / r) / pi per radius
    hw = STRUT_SETS["pre"][1]
    pre_hit = _crosses_set(pre, h.cap_position, "pre")
    for k in range(4):
        s = h.shell[k]
        rr = np.linspace(s.min_radius, s.max_radius, 20001)
        share = np.trapezoid(12.0 * np.arcsin(hw / rr) / np.pi, rr) / (s.max_radius - s.min_radius)
        mine = a["mirror_shell"] == k
        got, m_k = pre_hit[mine].sum(), mine.sum()
        sd = np.sqrt(m_k * share * (1.0 - share))
        assert abs(got - m_k * share) < 5.0 * sd, (k, got, m_k * share)


def _blur_run(blur, surface):
    par = {f"{s}{k}Blur": 0.0 for s in "PH" for k in (1, 3, 4, 6)}
    if blur:
        par.update({f"{surface}{k}Blur": blur for k in (1, 3, 4, 6)})
    # HRMA_Ideal switches the blur off (hrma.c:1106-1108): the reflectivity
    # draw stays, and both runs draw the same variates
    t = Tracer(seed=17, GratingType="NONE", DetectorType="NONE", HRMA_Use_Blur="yes",
               HRMA_Use_WFold="no", HRMA_Use_Struts="no", **par)
    n = 30000
    ph = Photons(n)
    ph.load(on_axis_photons(n, 1.49))
    Oracle(t).reflect(ph)
    return ph.attrs[:n].copy()


@pytest.mark.parametrize("surface", ["P", "H"])
def test_mirror_blur_is_a_gaussian_tilt_of_the_normal(surface):
    """A normal tilted by theta about an in-plane axis at angle psi to the
    plane of incidence turns the reflection by 2 theta (cos^2 a cos^2 psi +
    sin^2 a)^(1/2) (a the grazing angle), theta = B g, g ~ N(0, 1), psi
    uniform.  The H blur is the last turn of the ray: its angles follow that
    law exactly.  A P blur's turn is seen after the H reflection, whose hit
    point it moves: the angle at the focus is scaled by about 1 - d / F
    (d the P-H distance, ~0.6-0.9 m of F = 10 m), so its law is fitted with
    a scale in [0.88, 1.0]; a blur taken 1.25 times or twice, or as a
    FWHM, fails either way."""
    B = 20.0
    a = _blur_run(B, surface)
    b = _blur_run(0.0, surface)
    ok = (((a["flags"] | b["flags"]) & 0xFF) == 0) & (a["mirror_shell"] == b["mirror_shell"])
    assert ok.sum() > 8000
    cosang = np.clip(np.einsum("ij,ij->i", a["p"][ok], b["p"][ok]), -1.0, 1.0)
    ang = np.arccos(cosang)
    # an on-axis ray leaves a Wolter-I pair at four grazing angles to the axis
    graze = np.arccos(np.clip(-b["p"][ok][:, 0], -1.0, 1.0)) / 4.0
    rng = np.random.default_rng(3)
    m = 400000
    ga = rng.choice(graze, m)
    ref = 2.0 * np.abs(rng.standard_normal(m)) * np.sqrt(np.cos(ga) ** 2 * np.cos(rng.uniform(0, 2 * np.pi, m)) ** 2
                                                         + np.sin(ga) ** 2)
    x = ang / (B * ARCSEC)
    scales = [1.0] if surface == "H" else np.arange(0.88, 1.0001, 0.01)
    best = max(stats.ks_2samp(x / s, ref).pvalue for s in scales)
    assert best > 1e-3, best
    # a blur 1.25 times as large is rejected at every scale of the fit
    assert max(stats.ks_2samp(x / (1.25 * s), ref).pvalue for s in scales) < 1e-6


def _documented_rotation(az_deg, el_deg):
    """The mirror's attitude as hrma.c:853-864 documents it: first a
    positive rotation by osac_az about the SAOSAC +Y axis, which is MARX -Z;
    then a negative rotation by osac_el about the SAOSAC X' axis, the new
    MARX Y axis.  Returns (mirror rotation M, MARX -> mirror frame)."""
    first = Rotation.from_rotvec(np.radians(az_deg) * np.array([0.0, 0.0, -1.0]))
    new_y = first.apply([0.0, 1.0, 0.0])
    second = Rotation.from_rotvec(-np.radians(el_deg) * new_y)
    m = (second * first).as_matrix()
    return m, m.T


def test_tilt_matrices_are_the_documented_rotations():
    """The configured matrices against the documented attitude.  The
    reference composes the two rotations in the other order of axes
    (equal to first order); at the file's tilts (arcseconds) that is 1e-11,
    so the comparison is to 1e-9.  A sign slip in either angle moves an
    entry by 2 * tilt ~ 1e-5."""
    t = Tracer(GratingType="NONE", DetectorType="NONE")
    h = ConfigView(t).cfg.hrma
    for k in range(4):
        s = h.shell[k]
        for surf in range(2):
            az, el = s.tilt_deg[surf][0], s.tilt_deg[surf][1]
            mrot, fwd = _documented_rotation(az, el)
            got_f = np.array(s.fwd[surf][:]).reshape(3, 3)
            got_b = np.array(s.bwd[surf][:]).reshape(3, 3)
            np.testing.assert_allclose(got_f, fwd, atol=1e-9, rtol=0)
            np.testing.assert_allclose(got_b, mrot, atol=1e-9, rtol=0)
            np.testing.assert_allclose(got_f @ got_b, np.eye(3), atol=1e-15)
    # and tilts read from the file, in degrees (EKCHDOS06.rdb: h1 az_mis
    # 0.000672059134040231, el_mis 0.00123484664976509)
    assert h.shell[0].tilt_deg[1][0] == 0.000672059134040231
    assert h.shell[0].tilt_deg[1][1] == 0.00123484664976509


def _rdb_with(path, tilts):
    src = open(os.path.join(DATA, "hrma", "EKCHDOS06.rdb")).read().splitlines()
    out, hdr = [], None
    for line in src:
        if line.startswith("#") or not line.strip():
            out.append(line)
            continue
        f = line.split("\t")
        if hdr is None:
            hdr = f
        elif f[0][0] in "ph" and len(f) == len(hdr):
            f[hdr.index("az_mis")] = repr(tilts.get(f[0], (0.0, 0.0))[0])
            f[hdr.index("el_mis")] = repr(tilts.get(f[0], (0.0, 0.0))[1])
        out.append("\t".join(f))
    open(path, "w").write("\n".join(out) + "\n")


def _image_centroid(path, shell=0):
    t = Tracer(seed=3, GratingType="NONE", DetectorType="NONE", DitherModel="NONE", HRMA_Ideal="yes",
               HRMA_Use_WFold="no", HRMA_Use_Blur="no", HRMA_Use_Struts="no", HRMA_Geometry_File=path)
    n = 40000
    ph = Photons(n)
    ph.load(on_axis_photons(n, 1.49))
    Oracle(t).reflect(ph)
    a = ph.attrs[:n]
    ok = ((a["flags"] & 0xFF) == 0) & (a["mirror_shell"] == shell)
    x, p = a["x"][ok], a["p"][ok]
    s = -x[:, 0] / p[:, 0]
    return np.array([np.mean(x[:, 1] + s * p[:, 1]), np.mean(x[:, 2] + s * p[:, 2])])


@pytest.mark.parametrize("surface,angle,sign", [("p1", "az", (-1, 0)), ("p1", "el", (0, 1)),
                                                ("h1", "az", (1, 0)), ("h1", "el", (0, -1))])
def test_one_surface_tilt_moves_the_image(tmp_path, surface, angle, sign):
    """Reflection geometry, with the documented attitude.  Turning a mirror
    by delta about an axis across the optical axis turns the rays it
    reflects at the two azimuths where that tilt lies in their plane of
    incidence by 2 delta, and by ~0 at the two where it is across it: the
    image centroid moves by delta times the distance to the focus (F),
    along the direction the turned rays go.  The az attitude turns the
    mirror by -az about +Z (x toward y: the reflected rays of the mirror
    turn toward -y); the el attitude by -el about +Y.  The P reflection is
    followed by the H reflection, which mirrors an angular deviation in
    the plane of incidence (out = 2 beta - in), so a P tilt moves the image
    the other way from an H tilt: P az -> -y, P el -> +z, H az -> +y,
    H el -> -z, each by F delta to within the P-H distance (10 percent)."""
    base = str(tmp_path / "flat.rdb")
    _rdb_with(base, {})
    c0 = _image_centroid(base)
    delta = 30.0 / 3600.0
    tilted = str(tmp_path / "tilted.rdb")
    _rdb_with(tilted, {surface: (delta, 0.0) if angle == "az" else (0.0, delta)})
    d = _image_centroid(tilted) - c0
    f_delta = 10070.0 * np.radians(delta)
    along = d[0] if sign[0] else d[1]
    across = d[1] if sign[0] else d[0]
    s = sign[0] or sign[1]
    assert 0.9 * f_delta < s * along < 1.1 * f_delta, d
    assert abs(across) < 0.05 * f_delta, d
