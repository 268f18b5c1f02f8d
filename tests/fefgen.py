"""Writes synthetic ACIS FEF files (test input): the FUNCTION binary table of
acis_fef.c (CCD_ID, CHIPX_LO/HI, CHIPY_LO/HI, REGNUM, ENERGY, CHANNEL,
G<k>_FWHM/POS/AMPL), rows grouped by region, energies ascending.  The
CALDB FEF of the reference is not available, so tests make their own."""
import numpy as np

ENERGIES = np.round(np.geomspace(0.2, 12.0, 24), 4)
EV_PER_CHANNEL = 14.6


def channel(e):
    return 1000.0 * e / EV_PER_CHANNEL


def gaussians(region, e):
    """[(fwhm, pos, ampl)] of a region at energy e (keV): a main peak at the
    channel of e, a shoulder, a negative-amplitude notch in region 1, a
    component centred below 0 (sampled from its tail) in region 2."""
    c = channel(e)
    fwhm = 2.355 * (4.0 + 1.6 * np.sqrt(c))
    g = [(fwhm, c * (1.0 + 0.01 * region), 1.0 / fwhm), (1.6 * fwhm, 0.8 * c, 0.25 / fwhm)]
    if region == 1:
        g.append((0.5 * fwhm, c + 0.6 * fwhm, -0.3 / fwhm))
    elif region == 2:
        g.append((30.0, -20.0, 0.5))
    else:
        g.append((fwhm, c, 0.0))
    return g


def write_fef(path, ccds, split=512):
    """Regions: per chip the four quadrants (region = quadrant index % 4)."""
    rows = []
    regnum = 0
    for ccd in ccds:
        for qx in range(0, 1024, split):
            for qy in range(0, 1024, split):
                region = (qx // split) * (1024 // split) + qy // split
                for e in ENERGIES:
                    gs = gaussians(region % 4, e)
                    rows.append((ccd, qx + 1, qx + split, qy + 1, qy + split, regnum, e, channel(e), gs))
                regnum += 1
    ng = len(rows[0][8])
    names = ["CCD_ID", "CHIPX_LO", "CHIPX_HI", "CHIPY_LO", "CHIPY_HI", "REGNUM", "ENERGY", "CHANNEL"]
    forms = ["J"] * 6 + ["E", "E"]
    for k in range(1, ng + 1):
        names += [f"G{k}_FWHM", f"G{k}_POS", f"G{k}_AMPL"]
        forms += ["E", "E", "E"]
    dt = np.dtype([(n, ">i4" if f == "J" else ">f4") for n, f in zip(names, forms)])
    rec = np.zeros(len(rows), dt)
    for i, r in enumerate(rows):
        vals = list(r[:8]) + [v for g in r[8] for v in g]
        rec[i] = tuple(vals)
    cards = [("XTENSION", "'BINTABLE'"), ("BITPIX", "8"), ("NAXIS", "2"), ("NAXIS1", str(dt.itemsize)),
             ("NAXIS2", str(len(rows))), ("PCOUNT", "0"), ("GCOUNT", "1"), ("TFIELDS", str(len(names)))]
    for k, (n, f) in enumerate(zip(names, forms), 1):
        cards += [(f"TTYPE{k}", f"'{n}'"), (f"TFORM{k}", f"'{f}'")]
    cards += [("EXTNAME", "'FUNCTION'")]

    def header(cs):
        h = "".join(f"{k:<8}= {v:>20}".ljust(80) for k, v in cs) + "END".ljust(80)
        return h.ljust((len(h) + 2879) // 2880 * 2880).encode("ascii")

    primary = header([("SIMPLE", "T"), ("BITPIX", "8"), ("NAXIS", "0"), ("EXTEND", "T")])
    body = rec.tobytes()
    with open(path, "wb") as f:
        f.write(primary + header(cards) + body + b"\0" * ((-len(body)) % 2880))
    return path
