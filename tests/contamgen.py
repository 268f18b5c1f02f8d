"""Writes synthetic ACIS contamination files (test input) in the layout that
marx/libsrc/aciscontam.c:218-465 reads: one binary table per chip
(EXTNAME ACIS<ccd>_CONTAM, or AXAF_CONTAM with a CCD_ID keyword), one row
per layer with COMPONENT, N_ENERGY, ENERGY[], MU[], N_TIME, TIME[],
TAU0[], TAU1[] and optionally FXY[(1024 / FXYBLK)^2].  The CALDB file of the
reference is not available, so tests make their own."""
import numpy as np

BLOCK = 2880
NE, NT = 40, 6
ENERGIES = np.round(np.geomspace(0.15, 12.0, NE), 5)
TIMES = np.linspace(0.0, 8.0e8, NT)          # seconds from MJDref


def mu(layer, e):
    """Absorption coefficients of a layer: a smooth power law, and for layer
    1 an absorption edge at 0.5 keV."""
    e = np.asarray(e, dtype=np.float64)
    m = 0.8 * (1.0 / e) ** 2.4
    if layer == 1:
        m = 0.3 * (1.0 / e) ** 2.0 * np.where(e > 0.5, 3.0, 1.0)
    return m


def taus(layer):
    """tau0(t), tau1(t) of a layer on TIMES: rising linearly."""
    t0 = (0.02 + 0.10 * layer) + 0.25 * TIMES / TIMES[-1]
    t1 = (0.05 + 0.03 * layer) + 0.60 * TIMES / TIMES[-1]
    return t0, t1


def fxy_table(ccd, layer, blocking):
    """A spatial table of (1024 / blocking)^2 values, row-major in chip y."""
    n = 1024 // blocking
    y, x = np.mgrid[0:n, 0:n].astype(np.float64) / n
    return (0.2 + layer * 0.1) + x ** 2 + 0.5 * (y - 0.3) ** 2 + 0.01 * ccd


def _card(k, v):
    return f"{k:<8}= {v:>20}".ljust(80)


def _header(cards):
    h = "".join(_card(k, v) for k, v in cards) + "END".ljust(80)
    return h.ljust((len(h) + BLOCK - 1) // BLOCK * BLOCK).encode("ascii")


def _table(ccd, layers, blocking, axaf, component=0, n_time=NT):
    cols = [("COMPONENT", "J", 1), ("N_ENERGY", "J", 1), ("ENERGY", "E", NE), ("MU", "E", NE),
            ("N_TIME", "J", 1), ("TIME", "D", NT), ("TAU0", "E", NT), ("TAU1", "E", NT)]
    if blocking:
        cols.append(("FXY", "E", (1024 // blocking) ** 2))
    code = {"J": ">i4", "E": ">f4", "D": ">f8"}
    dt = np.dtype([(n, code[f], (r,) if r > 1 else ()) for n, f, r in cols])
    rec = np.zeros(layers, dt)
    for L in range(layers):
        t0, t1 = taus(L)
        rec[L]["COMPONENT"] = component
        rec[L]["N_ENERGY"] = NE
        rec[L]["ENERGY"] = ENERGIES
        rec[L]["MU"] = mu(L, ENERGIES)
        rec[L]["N_TIME"] = n_time
        rec[L]["TIME"] = TIMES
        rec[L]["TAU0"] = t0
        rec[L]["TAU1"] = t1
        if blocking:
            rec[L]["FXY"] = fxy_table(ccd, L, blocking).reshape(-1)
    cards = [("XTENSION", "'BINTABLE'"), ("BITPIX", "8"), ("NAXIS", "2"), ("NAXIS1", str(dt.itemsize)),
             ("NAXIS2", str(layers)), ("PCOUNT", "0"), ("GCOUNT", "1"), ("TFIELDS", str(len(cols)))]
    for k, (n, f, r) in enumerate(cols, 1):
        cards += [(f"TTYPE{k}", f"'{n.lower()}'"), (f"TFORM{k}", f"'{r}{f}'")]
    if axaf:
        cards += [("EXTNAME", "'AXAF_CONTAM'"), ("CCD_ID", str(ccd))]
    else:
        cards += [("EXTNAME", f"'ACIS{ccd}_CONTAM'")]
    if blocking:
        cards += [("FXYBLK", str(blocking))]
    body = rec.tobytes()
    return _header(cards) + body + b"\0" * ((-len(body)) % BLOCK)


def write_contam(path, ccds=range(10), layers=2, blocking=0, axaf=False, component=0, n_time=NT):
    """blocking 0: no FXY column (the analytic spatial models); else the
    FXY table with that FXYBLK."""
    primary = _header([("SIMPLE", "T"), ("BITPIX", "8"), ("NAXIS", "0"), ("EXTEND", "T")])
    with open(path, "wb") as f:
        f.write(primary)
        for ccd in ccds:
            f.write(_table(ccd, layers, blocking, axaf, component, n_time))
    return path


def tau_at(layer, t):
    """The layer's taus linearly interpolated (extrapolated) at time t, as
    float columns read back."""
    t0, t1 = taus(layer)
    t0, t1 = t0.astype(np.float32).astype(np.float64), t1.astype(np.float32).astype(np.float64)
    k = int(np.clip(np.searchsorted(TIMES, t, side="right"), 1, NT - 1))
    w = (t - TIMES[k - 1]) / (TIMES[k] - TIMES[k - 1])
    return t0[k - 1] + (t0[k] - t0[k - 1]) * w, t1[k - 1] + (t1[k] - t1[k - 1]) * w
