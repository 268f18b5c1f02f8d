"""Named workloads (parameter sets) of the benchmark and the tests."""
import os

import numpy as np

from .sim import Tracer, offaxis_directions
from .spectra import write_power_law

# on-axis: the source at the nominal pointing of the standard parameter file
ON_AXIS = dict(SourceRA=250.2134679741175, SourceDEC=-53.75743813458669)


def hetg_acis_s(tmpdir, photons=10_000_000, exposure=50_000.0, seed=1):
    """HETG + ACIS-S: point source on axis, power law of photon index 2 over
    0.4-8 keV (tabulated FILE spectrum), HEG/MEG orders from the efficiency
    tables, INTERNAL dither of 16 arcsec amplitude and 1000 / 707 s periods.
    The flux is set so that `photons` photons arrive in `exposure` seconds."""
    spec = write_power_law(os.path.join(tmpdir, "powerlaw_g2_0.4_8.dat"), gamma=2.0, emin=0.4, emax=8.0)
    t = Tracer(seed=seed, GratingType="HETG", DetectorType="ACIS-S", SourceType="POINT",
               SpectrumType="FILE", SpectrumFile=spec, DitherModel="INTERNAL",
               DitherAmp_RA=16.0, DitherAmp_Dec=16.0, DitherPeriod_RA=1000.0, DitherPeriod_Dec=707.0,
               **ON_AXIS)
    flux = photons / (exposure * t.geometric_area)
    t.init_source(SourceFlux=flux)
    return t


def letg_hrc_s(tmpdir=None, photons=50_000_000, exposure=50_000.0, seed=1):
    """LETG + HRC-S: uniform disk of 2 arcmin radius on axis, flat spectrum
    over 0.07-10 keV, LETG orders from the efficiency tables with the fine
    and coarse support gratings, HRC-S with the HESF, MCP blur and pixel map,
    INTERNAL dither with the parameter file's amplitudes and periods.  The
    flux is set so that `photons` photons arrive in `exposure` seconds."""
    t = Tracer(seed=seed, GratingType="LETG", DetectorType="HRC-S", SourceType="DISK",
               SpectrumType="FLAT", MinEnergy=0.07, MaxEnergy=10.0, DitherModel="INTERNAL",
               **{"S-DiskTheta0": 0.0, "S-DiskTheta1": 120.0}, **ON_AXIS)
    flux = photons / (exposure * t.geometric_area)
    t.init_source(SourceFlux=flux)
    return t


def batched_acis_i(tmpdir, num_sources=4096, seed=1, layout_seed=0, min_count=100, max_count=100_000,
                   max_offaxis_arcmin=15.0):
    """Batched: num_sources point sources at off-axis angles uniform over
    0-max_offaxis_arcmin with random azimuths, photon counts log-uniform
    over [min_count, max_count], power law of photon index 2 over 0.3-10 keV,
    HRMA + ACIS-I, no grating, no dither.  Returns (tracer, dirs, counts)."""
    spec = write_power_law(os.path.join(tmpdir, "powerlaw_g2_0.3_10.dat"), gamma=2.0, emin=0.3, emax=10.0)
    t = Tracer(seed=seed, GratingType="NONE", DetectorType="ACIS-I", SourceType="POINT",
               SpectrumType="FILE", SpectrumFile=spec, DitherModel="NONE", **ON_AXIS)
    t.init_source()
    rng = np.random.default_rng(layout_seed)
    theta = np.radians(rng.uniform(0.0, max_offaxis_arcmin, num_sources) / 60.0)
    phi = rng.uniform(0.0, 2.0 * np.pi, num_sources)
    counts = np.round(10.0 ** rng.uniform(np.log10(min_count), np.log10(max_count), num_sources))
    return t, offaxis_directions(theta, phi), counts.astype(np.uint32)


def diff_mirror(tmpdir=None, seed=1, param_seed=0, tilt_sigma=1.0, shift_sigma=0.01):
    """Differentiable: on-axis point source at 1.49 keV, mirror only (no
    grating or detector, no dither), for the weighted-photon images of
    marx_amd.diff.  The 20 misalignment parameters (tilt az, tilt el in
    arcsec; decenter y, decenter z, despace x in mm; per shell pair) are drawn
    from N(0, tilt_sigma arcsec / shift_sigma mm): `params` where the
    gradient is taken, and `target_params` = params plus a second draw, whose
    image is the target.  Returns (tracer, params, target_params)."""
    t = Tracer(seed=seed, source=True, GratingType="NONE", DetectorType="NONE", SourceType="POINT",
               SpectrumType="FLAT", MinEnergy=1.49, MaxEnergy=1.49, DitherModel="NONE", **ON_AXIS)
    rng = np.random.default_rng(param_seed)

    def draw():
        p = np.zeros(20)
        for s in range(4):
            p[5 * s:5 * s + 2] = rng.normal(0.0, tilt_sigma, 2)
            p[5 * s + 2:5 * s + 5] = rng.normal(0.0, shift_sigma, 3)
        return p

    params = draw()
    return t, params, params + draw()


WORKLOADS = {"hetg-acis-s": hetg_acis_s, "letg-hrc-s": letg_hrc_s}


def plain_acis_s(photons=1_000_000, tmax=10_000.0, seed=1):
    """Plain: on-axis point source at 1.49 keV, arrival times uniform over
    10 ks, no dither, no grating, ACIS-S."""
    t = Tracer(seed=seed, GratingType="NONE", DetectorType="ACIS-S", SourceType="POINT",
               SpectrumType="FLAT", MinEnergy=1.49, MaxEnergy=1.49, DitherModel="NONE", **ON_AXIS)
    t.init_source(uniform_times=tmax)
    return t
