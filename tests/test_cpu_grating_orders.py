"""Order sampling of the HETG pass (diffract.c:974-1130 over the CALDB
efficiency tables, diffract.c:1252-1300), pinned by the tables themselves,
read here with the oracle's own FITS reader: for photons of one energy of
the table's grid leaving shell s, a fraction 1 - vig (HETG_Shell<n>_Vig)
is vignetted by the grating support (binomial), and of the rest, order m is taken with the
probability EFF_m(E) of extension AXAF_GREFF(s+1) and none with
1 - sum_m EFF_m(E) (chi-square over the orders)."""
import os

import numpy as np
import pytest
from scipy import stats

from marx_amd import Photons, Tracer, on_axis_photons
from marx_amd._lib import DATA_DIR
from oracle.oracle import Oracle
from oracle.subpix import read_bintables

GREFF = os.path.join(DATA_DIR, "caldb", "hetgD1996-11-01greffpr001N0007.fits")
VBLOCKED, UNDIFFRACTED = 0x10, 0x04


@pytest.mark.parametrize("target_kev", [0.9, 2.5])
def test_hetg_orders_follow_the_efficiency_tables(target_kev):
    tabs = read_bintables(GREFF)
    vigs = {"1": 0.9, "3": 0.8, "4": 0.85, "6": 0.95}
    t = Tracer(seed=13, GratingType="HETG", DetectorType="NONE", DitherModel="NONE",
               **{f"HETG_Shell{k}_Vig": v for k, v in vigs.items()})
    e_grid = tabs["AXAF_GREFF1"]["ENERGY"]
    row = int(np.argmin(np.abs(e_grid - target_kev)))
    energy = float(np.float32(e_grid[row]))
    n = 400000
    p = Photons(n)
    p.load(on_axis_photons(n, energy))
    o = Oracle(t)
    o.reflect(p)
    before = (p.active["flags"] & 0xFF) == 0
    o.diffract(p)
    a = p.active
    checked = 0
    for s in range(4):
        tab = tabs[f"AXAF_GREFF{s + 1}"]
        assert np.float32(tab["ENERGY"][row]) == np.float32(energy)
        eff = tab["EFF"][row].astype(np.float64)
        norders = len(eff)
        orders = np.arange(norders) - norders // 2
        if eff.sum() > 1.0:
            eff = eff / eff.sum()
        sel = before & (a["mirror_shell"] == s)
        nsel = sel.sum()
        vig = vigs["1346"[s]]
        vb = (a["flags"][sel] & VBLOCKED) != 0
        k = vb.sum()
        assert abs(k - nsel * (1 - vig)) < 5 * np.sqrt(nsel * vig * (1 - vig)) + 1
        rest = sel & ((a["flags"] & VBLOCKED) == 0)
        m = rest.sum()
        und = ((a["flags"][rest] & UNDIFFRACTED) != 0).sum()
        got = np.array([((a["order"][rest] == o_) & ((a["flags"][rest] & UNDIFFRACTED) == 0)).sum()
                        for o_ in orders] + [und])
        exp = np.r_[eff, max(0.0, 1.0 - eff.sum())] * m
        keep = exp > 5
        # merge the rare categories into one
        cnt = np.r_[got[keep], got[~keep].sum()]
        ex = np.r_[exp[keep], exp[~keep].sum()]
        if ex[-1] < 5:
            cnt, ex = cnt[:-1], ex[:-1]
        chi2 = ((cnt - ex) ** 2 / ex).sum()
        assert stats.chi2.sf(chi2, len(ex) - 1) > 1e-4, (s, chi2, len(ex))
        checked += 1
    assert checked == 4
