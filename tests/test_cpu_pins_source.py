"""Pins of the oracle's source pass by the closed forms of the reference's
source models (CPU only):

- INTERNAL dither (marx/libsrc/dither.c:167-182, 551-581, 609-628): every
  photon's dither state is A sin(2 pi t / P + phase) per axis from the
  parameter file's amplitudes (arcsec), periods and phases, the roll adds
  Roll_Nom; an on-axis photon's direction is then (-cos dec cos ra,
  cos dec sin ra, sin dec) -- the rotation of -x that apply_dither performs,
  written out -- and with AspectBlur the recovered (ra, dec) scatter about
  the state by N(0, AspectBlur) each;
- FLAT spectrum (spectrum.c:138-147): energies uniform on [MinEnergy,
  MaxEnergy] (KS);
- FILE spectrum (spectrum.c:48-92, prob.c:46-58): energies follow the
  piecewise-linear cumulative through the table's rectangle sums
  y_i (x_i - x_(i-1)), computed here from the file (KS); the left-rectangle
  sum and the exact power law are rejected on the coarse table used;
- Poisson arrival times (source.c:313-342): gaps times flux times the
  aperture area are Exp(1) (KS);
- DISK (s-disk.c:63-127): the angle theta to the source direction has
  (theta^2 - theta0^2) / (theta1^2 - theta0^2) uniform, azimuths uniform.
"""
import numpy as np
import pytest
from scipy import stats

from marx_amd import Photons, Tracer
from oracle.oracle import Oracle
from cfgview import ConfigView

ON_AXIS = dict(SourceRA=250.2134679741175, SourceDEC=-53.75743813458669)
ARCSEC = np.pi / 180 / 3600


def _create(n, **kw):
    t = Tracer(seed=23, source=True, GratingType="NONE", DetectorType="NONE", **kw)
    p = Photons(n)
    got = Oracle(t).create_photons(p, n)
    assert got == n
    return t, p.active.copy()


DITHER = dict(DitherModel="INTERNAL", DitherAmp_RA=16.0, DitherAmp_Dec=11.0, DitherAmp_Roll=40.0,
              DitherPeriod_RA=1000.0, DitherPeriod_Dec=707.0, DitherPeriod_Roll=5000.0,
              DitherPhase_RA=0.3, DitherPhase_Dec=1.1, DitherPhase_Roll=0.7, Roll_Nom=30.0,
              SourceType="POINT", SpectrumType="FLAT", MinEnergy=1.0, MaxEnergy=2.0, SourceFlux=2.0e-4,
              **ON_AXIS)


def _lissajous(t, amp, period, phase):
    return amp * ARCSEC * np.sin(2.0 * np.pi * t / period + phase)


def test_internal_dither_is_the_lissajous_pattern():
    tr, a = _create(20000, AspectBlur=0.0, **DITHER)
    t = a["arrival_time"]
    assert t[-1] > 3 * 5000.0                     # several periods of every axis
    ds = a["dither"]
    ra = _lissajous(t, 16.0, 1000.0, 0.3)
    dec = _lissajous(t, 11.0, 707.0, 1.1)
    roll = np.radians(30.0) + _lissajous(t, 40.0, 5000.0, 0.7)
    for k, want in ((0, ra), (1, dec), (2, roll)):
        np.testing.assert_allclose(ds[:, k], want.astype(np.float32), rtol=2e-7, atol=1e-12)
    np.testing.assert_array_equal(ds[:, 3:], 0.0)
    # the direction: -x turned by (ra, dec) about the axis orthogonal to the
    # instantaneous pointing (the roll turns -x into itself)
    # (the angles the reference turns by are the stored float values:
    # dither.c:173 assigns through the float state)
    ra32, dec32 = ra.astype(np.float32).astype(np.float64), dec.astype(np.float32).astype(np.float64)
    p = a["p"]
    want = np.stack([-np.cos(dec32) * np.cos(ra32), np.cos(dec32) * np.sin(ra32), np.sin(dec32)], axis=1)
    np.testing.assert_allclose(p, want, atol=1e-11, rtol=0)


def test_aspect_blur_scatters_the_pointing():
    sigma = 0.5
    tr, a = _create(20000, AspectBlur=sigma, **DITHER)
    t = a["arrival_time"]
    ra = _lissajous(t, 16.0, 1000.0, 0.3)
    dec = _lissajous(t, 11.0, 707.0, 1.1)
    p = a["p"]
    got_dec = np.arcsin(p[:, 2])
    got_ra = np.arctan2(p[:, 1], -p[:, 0])
    for d in ((got_ra - ra) / (sigma * ARCSEC), (got_dec - dec) / (sigma * ARCSEC)):
        assert stats.kstest(d, "norm").pvalue > 1e-3
        assert stats.kstest(d / 2.0, "norm").pvalue < 1e-6    # twice the blur is rejected


def test_flat_spectrum_is_uniform():
    tr, a = _create(40000, SourceType="POINT", SpectrumType="FLAT", MinEnergy=0.3, MaxEnergy=9.0,
                    DitherModel="NONE", **ON_AXIS)
    assert stats.kstest(a["energy"], stats.uniform(0.3, 8.7).cdf).pvalue > 1e-3


def _coarse_power_law(path, gamma=2.0, emin=0.4, emax=8.0, n=25):
    # a linear grid: on a geometric one the left and right rectangle sums of
    # a power law are proportional, and normalised, equal
    e = np.linspace(emin, emax, n)
    f = e ** (-gamma)
    np.savetxt(path, np.stack([e, f], axis=1), header="coarse power law")
    return e, f


def _cdf_through(x, cum):
    return lambda e: np.interp(e, x, cum / cum[-1])


def test_file_spectrum_follows_its_rectangle_rule():
    """spectrum.c:59-76 sums y_i (x_i - x_(i-1)) for i >= 1 (the first term is
    0), and prob.c draws E by linear interpolation of that cumulative."""
    import tempfile, os
    d = tempfile.mkdtemp()
    path = os.path.join(d, "pl.dat")
    x, y = _coarse_power_law(path)
    tr, a = _create(60000, SourceType="POINT", SpectrumType="FILE", SpectrumFile=path, DitherModel="NONE",
                    **ON_AXIS)
    e = a["energy"]
    right = np.concatenate([[0.0], np.cumsum(y[1:] * np.diff(x))])
    assert stats.kstest(e, _cdf_through(x, right)).pvalue > 1e-3
    left = np.concatenate([[0.0], np.cumsum(y[:-1] * np.diff(x))])
    assert stats.kstest(e, _cdf_through(x, left)).pvalue < 1e-6
    exact = lambda v: (x[0] ** -1 - np.clip(v, x[0], x[-1]) ** -1) / (x[0] ** -1 - x[-1] ** -1)
    assert stats.kstest(e, exact).pvalue < 1e-6
    # the flux of a FILE spectrum with SourceFlux <= 0 is the table's sum
    c = ConfigView(tr).cfg.src
    assert c.total_flux > 0


def test_poisson_arrival_gaps():
    flux = 5.0e-3
    tr, a = _create(40000, SourceType="POINT", SpectrumType="FLAT", MinEnergy=1.0, MaxEnergy=2.0,
                    SourceFlux=flux, DitherModel="NONE", **ON_AXIS)
    t = a["arrival_time"]
    gaps = np.diff(np.concatenate([[0.0], t])) * flux * tr.geometric_area
    assert stats.kstest(gaps, "expon").pvalue > 1e-3
    assert stats.kstest(gaps * 2.0, "expon").pvalue < 1e-6   # a mean gap halved is rejected


@pytest.mark.parametrize("theta0,theta1", [(0.0, 120.0), (30.0, 90.0)])
def test_disk_profile(theta0, theta1):
    tr, a = _create(40000, SourceType="DISK", SpectrumType="FLAT", MinEnergy=1.0, MaxEnergy=2.0,
                    DitherModel="NONE", **{"S-DiskTheta0": theta0, "S-DiskTheta1": theta1}, **ON_AXIS)
    p0 = np.array([-1.0, 0.0, 0.0])
    p = a["p"]
    th = np.arctan2(np.linalg.norm(np.cross(p, p0), axis=1), p @ p0) / ARCSEC
    u = (th ** 2 - theta0 ** 2) / (theta1 ** 2 - theta0 ** 2)
    assert stats.kstest(u, "uniform").pvalue > 1e-3
    assert stats.kstest(np.sqrt(np.clip(u, 0, 1)), "uniform").pvalue < 1e-6   # uniform in theta is rejected
    az = np.arctan2(p[:, 2], p[:, 1])
    assert stats.kstest(az, stats.uniform(-np.pi, 2 * np.pi).cdf).pvalue > 1e-3
