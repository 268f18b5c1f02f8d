"""The bucket indices of the search tables (mxg_bucket, built at init: the
mirror's optical constants, area corrections and WFold e_alpha values, the
gratings' energy grids, the detectors' QE tables): a search that starts from the bucket holding the
value must
return the plain bisection's index (dev_common.h's search_f / search_w, which
restate JDMbinary_search_f, jdmath/src/binsrch.c) for every value -- each
table entry, its neighbouring floats, midpoints, values below and above the
table and random energies.  The bisection below is the textbook one; the
bucket counts are checked against a direct count of the table's keys.
(CPU only: the counts are the host's, the search a restatement of the
device's.)"""
import numpy as np
import pytest

from marx_amd import Tracer
from cfgview import ConfigView


def bisect(x, xp):
    """The bisection: entry index n0 < n1 while n1 > n0 + 1, returning an
    exact hit, else n1 when x >= xp[n0], else n0."""
    n0, n1 = 0, len(xp)
    while n1 > n0 + 1:
        n2 = (n0 + n1) >> 1
        v = xp[n2]
        if v >= x:
            if v == x:
                return n2
            n1 = n2
        else:
            n0 = n2
    return n1 if x >= xp[n0] else n0


def keys(v, double_keys):
    if double_keys:
        return np.asarray(v, np.float64).view(np.uint64).astype(object)
    return np.asarray(v, np.float32).view(np.uint32).astype(object)


def bucketed(x, xp, b, cnt, double_keys):
    x0 = xp[0]
    if not (x > x0):
        if x == x0:
            return 1
        return 0
    k = (int(keys([x], double_keys)[0]) - int(b.base)) >> int(b.shift)
    j = k if k < b.nb else b.nb - 1
    lo, hi = int(cnt[j]), int(cnt[j + 1])
    while lo < hi:
        mid = (lo + hi) >> 1
        if xp[mid] < x:
            lo = mid + 1
        else:
            hi = mid
    if b.ok == 2 and lo + 1 < len(xp) and xp[lo] == x and xp[lo + 1] == x:
        return bisect(x, xp)
    return lo


def _check_table(xp, b, cnt, double_keys, nbmax, rng, max_entries=4000):
    """Bucket counts against a direct count of the keys, and the bucketed
    search against the bisection for the entries (a random max_entries of
    them on longer tables), their neighbours, midpoints, values outside the
    table and random values inside it."""
    dt = np.float64 if double_keys else np.float32
    xp = np.asarray(xp, dt)
    repeats = bool(np.any(xp[1:] == xp[:-1]))
    assert b.ok == (2 if repeats else 1) and 1 <= b.nb <= nbmax
    k = keys(xp, double_keys)
    assert int(b.base) == int(k[0])
    # counts: entries whose key lies below each bucket's first key
    ku = np.array([int(kk) for kk in k], dtype=np.uint64)
    bounds = np.array([int(b.base) + (j << int(b.shift)) for j in range(b.nb)], dtype=np.uint64)
    want = [int(np.count_nonzero(ku < x)) for x in bounds] + [len(xp)]
    assert [int(c) for c in cnt] == want
    ent = xp if len(xp) <= max_entries else xp[np.sort(rng.choice(len(xp) - 1, max_entries, replace=False))]
    xs = list(ent) + list(np.nextafter(ent, np.inf).astype(dt)) + list(np.nextafter(ent, -np.inf).astype(dt))
    xs += list(((xp[1:].astype(np.float64) + xp[:-1]) / 2).astype(dt))[:max_entries]
    xs += [dt(xp[0] / 2), dt(xp[-1] * 2), dt(0.0), dt(-1.0)]
    xs += list(rng.uniform(float(xp[0]), float(xp[-1]), 2000).astype(dt))
    for x in xs:
        assert bucketed(x, xp, b, cnt, double_keys) == bisect(x, xp), (x, double_keys)


def test_mirror_search_tables_bucket_index():
    t = Tracer(seed=1, GratingType="NONE", DetectorType="NONE")
    cv = ConfigView(t)
    h = cv.cfg.hrma
    rng = np.random.default_rng(4)
    assert h.opt_n > 1000
    b = h.opt_bucket
    _check_table(cv.f(h.opt_e_off, h.opt_n), b, cv.f(b.off, b.nb + 1), False, 256, rng)
    for sh in range(4):
        S = h.shell[sh]
        b = S.corr_bucket
        assert S.corr_n > 100
        _check_table(cv.f(S.corr_e_off, S.corr_n), b, cv.f(b.off, b.nb + 1), False, 64, rng)
        for k in range(2):
            w, b = S.wfold[k], S.ea_bucket[k]
            assert w.n_arrays > 100
            ea = cv.d(w.e_alphas_off, w.n_arrays).astype(np.float32)
            _check_table(ea, b, cv.f(b.off, b.nb + 1), True, 32, rng)


@pytest.mark.parametrize("det", ["ACIS-S", "ACIS-I", "HRC-S", "HRC-I"])
def test_detector_qe_bucket_index(det):
    """The detector's QE tables: each ACIS chip's CALDB QE, the HRC MCP QE
    (2e4-4e4 rows) and UV/ion shield transmissions."""
    t = Tracer(seed=1, GratingType="NONE", DetectorType=det)
    cv = ConfigView(t)
    d = cv.cfg.det
    rng = np.random.default_rng(8)
    tables = []
    if det.startswith("ACIS"):
        tables = [(d.chip[k].qe_e_off, d.chip[k].qe_n, d.chip[k].qe_bucket) for k in range(d.num_chips)]
    else:
        tables = [(q.e_off, q.n, q.bucket) for q in list(d.mcp_qe) + list(d.filter_qe) if q.n > 0]
    assert tables and all(n > 1000 for _, n, _ in tables[:1])
    for off, n, b in tables:
        _check_table(cv.f(off, n), b, cv.f(b.off, b.nb + 1), False, 256, rng)


def test_unsorted_or_short_tables_are_searched_whole():
    """A table the bucket index cannot serve (not strictly increasing, or a
    single entry, as an ideal mirror has none) is marked ok = 0."""
    t = Tracer(seed=1, GratingType="NONE", DetectorType="NONE", HRMA_Ideal="yes")
    h = ConfigView(t).cfg.hrma
    assert h.opt_n == 0 and h.opt_bucket.ok == 0


def lower_bound_1(e, E):
    """draw_order's bracket search: the first j in [1, n - 1] with E[j] >= e,
    else n - 1 (bisection)."""
    lo, hi = 1, len(E) - 1
    while lo < hi:
        mid = (lo + hi) >> 1
        if e > E[mid]:
            lo = mid + 1
        else:
            hi = mid
    return lo


@pytest.mark.parametrize("grating", ["HETG", "LETG"])
def test_grating_energy_bucket_index(grating):
    t = Tracer(seed=1, GratingType=grating, DetectorType="NONE")
    cv = ConfigView(t)
    gc = cv.cfg.grating
    rng = np.random.default_rng(6)
    tables = [gc.g[k] for k in range(4)] + ([gc.fine, gc.coarse] if grating == "LETG" else [])
    for g in tables:
        E = cv.f(g.energies_off, g.num_energies)
        # the order draw bisects the orders when every energy's cumulative
        # efficiencies rise with the order index
        cum = cv.f(g.cum_eff_off, g.num_energies * g.num_orders).reshape(g.num_energies, g.num_orders)
        assert g.cum_sorted == int(bool(np.all(np.diff(cum, axis=1) >= 0)))
        assert g.cum_sorted == 1
        b = g.e_bucket
        cnt = cv.f(b.off, b.nb + 1)
        _check_table(E, b, cnt, False, 256, rng)
        xs = list(E) + list(np.nextafter(E, np.inf)) + list(rng.uniform(E[0] * 0.5, E[-1] * 1.5, 3000).astype(np.float32))
        for e in xs:
            e = np.float32(e)
            want = lower_bound_1(e, E)
            if e > E[0]:
                lb = bucketed(e, E, b, cnt, False)
                got = min(max(lb, 1), len(E) - 1)
                assert got == want, (e, got, want)


def test_spectrum_linear_bucket_index(tmp_path):
    """The FILE spectrum's cumulative distribution gets a linear bucket index
    (mxg_lbucket): the counts are the entries whose key min(nb - 1,
    int((x - x0) * scale)) is below each bucket, and the bucketed search
    (dev_common.h's search_dl, restated) returns the plain bisection's index
    for uniform variates, 0 and 1, every entry and its neighbours."""
    from marx_amd.spectra import write_power_law
    spec = write_power_law(str(tmp_path / "pl.dat"), gamma=2.0, emin=0.4, emax=8.0)
    t = Tracer(seed=1, source=True, GratingType="NONE", DetectorType="NONE", SpectrumType="FILE",
               SpectrumFile=spec)
    cv = ConfigView(t)
    s = cv.cfg.src
    b = s.file_bucket
    cum = cv.d(s.file_cum_off, s.file_n)
    assert s.file_n == 2000 and b.ok == 1 and b.nb == 2000
    assert cum[0] == 0.0 and cum[-1] == 1.0
    assert b.x0 == cum[0] and b.scale == b.nb / (cum[-1] - cum[0])

    def key(x):
        t_ = (np.float64(x) - np.float64(b.x0)) * np.float64(b.scale)
        return int(t_) if t_ < b.nb else b.nb - 1

    cnt = cv.d(b.off, b.nb + 1)
    keys = [key(x) for x in cum]
    assert all(k1 <= k2 for k1, k2 in zip(keys, keys[1:]))
    assert [int(c) for c in cnt] == [sum(1 for k in keys if k < j) for j in range(b.nb)] + [len(cum)]

    def search_dl(x):
        if not (x > cum[0]):
            return 1 if x == cum[0] else 0
        j = key(x)
        lo, hi = int(cnt[j]), int(cnt[j + 1])
        while lo < hi:
            mid = (lo + hi) >> 1
            if cum[mid] < x:
                lo = mid + 1
            else:
                hi = mid
        return lo

    rng = np.random.default_rng(9)
    u = rng.integers(0, 2**32, 20000, dtype=np.uint64).astype(np.float64) / float(0xFFFFFFFF)
    xs = list(cum) + list(np.nextafter(cum, np.inf)) + list(np.nextafter(cum, -np.inf)) + [0.0, 1.0] + list(u)
    for x in xs:
        assert search_dl(x) == bisect(x, cum), x
    # a flat spectrum has none
    t2 = Tracer(seed=1, source=True, GratingType="NONE", DetectorType="NONE", SpectrumType="FLAT",
                MinEnergy=1.0, MaxEnergy=2.0)
    assert ConfigView(t2).cfg.src.file_bucket.ok == 0
