"""Spectrum tables in the two-column (energy keV, photons/cm^2/s/keV) text
format that SpectrumType=FILE reads (marx/libsrc/spectrum.c:214-272)."""
import numpy as np


def write_power_law(path, gamma=2.0, emin=0.4, emax=8.0, n=2000, norm=1.0):
    """Power law N(E) = norm * E^-gamma tabulated on n log-spaced energies."""
    e = np.geomspace(emin, emax, n)
    f = norm * e ** (-gamma)
    with open(path, "w") as fp:
        fp.write(f"# power law, photon index {gamma}, {emin}-{emax} keV\n")
        for a, b in zip(e, f):
            fp.write(f"{a:.9e} {b:.9e}\n")
    return path


def write_flat(path, emin=0.07, emax=10.0, n=2000):
    e = np.linspace(emin, emax, n)
    with open(path, "w") as fp:
        fp.write(f"# flat spectrum {emin}-{emax} keV\n")
        for a in e:
            fp.write(f"{a:.9e} 1.0\n")
    return path
