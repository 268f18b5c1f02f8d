"""HETG / LETG diffraction: GPU kernel vs the FP64 CPU oracle through the
C-ABI (marx_grating_diffract, replacing marx/libsrc/grating.c:88-114 ->
diffract.c:974-1130).  The oracle runs the reference's shell-major
merge-walk interpolation over energy-sorted photons; the GPU interpolates
per photon."""
import numpy as np
import pytest

from marx_amd import Photons, Tracer, on_axis_photons
from oracle.oracle import Oracle
from tests.parity import compare


def _run(tracer, rec):
    n = len(rec)
    pg, po = Photons(n), Photons(n)
    pg.load(rec)
    po.load(rec)
    tracer.reflect(pg)
    tracer.diffract(pg)
    o = Oracle(tracer)
    o.reflect(po)
    o.diffract(po)
    return pg, po


@pytest.mark.gpu
@pytest.mark.parametrize("grid", [0, 8])
def test_hetg_parity(gpu, grid):
    """grid: the grating kernel's blocks per CU (0, the default: one round
    of the blocks a CU holds at once)."""
    t = Tracer(seed=11, GratingType="HETG", DetectorType="NONE")
    rng = np.random.default_rng(5)
    rec = on_axis_photons(200000, rng.uniform(0.4, 8.0, 200000))
    t.set_schedule(grating=grid)
    try:
        pg, po = _run(t, rec)
    finally:
        t.set_schedule(grating=0)
    compare(pg.active, po.active, max_flips=2, label=f"hetg, grid {grid}")
    a = pg.active[(pg.active["flags"] & 0xFF) == 0]
    # orders come from the full -11..+11 table; 0 and +-1 dominate
    frac = {o: np.mean(a["order"] == o) for o in (-1, 0, 1)}
    assert 0.4 < frac[0] < 0.7 and 0.1 < frac[1] < 0.3 and 0.1 < frac[-1] < 0.3


@pytest.mark.gpu
def test_hetg_no_sectors_computed_efficiencies(gpu):
    t = Tracer(seed=12, GratingType="HETG", DetectorType="NONE",
               Use_HETG_Sector_Files="no", UseGratingEffFiles="no")
    rng = np.random.default_rng(6)
    rec = on_axis_photons(100000, rng.uniform(0.5, 7.0, 100000))
    pg, po = _run(t, rec)
    compare(pg.active, po.active, max_flips=2, label="hetg computed")


@pytest.mark.gpu
@pytest.mark.parametrize("grid", [0, 8])
def test_letg_parity(gpu, grid):
    t = Tracer(seed=13, GratingType="LETG", DetectorType="NONE")
    rng = np.random.default_rng(7)
    rec = on_axis_photons(200000, rng.uniform(0.07, 10.0, 200000))
    t.set_schedule(grating=grid)
    try:
        pg, po = _run(t, rec)
    finally:
        t.set_schedule(grating=0)
    compare(pg.active, po.active, max_flips=2, label=f"letg, grid {grid}")
    a = pg.active[(pg.active["flags"] & 0xFF) == 0]
    assert (a["support_orders"][:, 0] != 0).any()


@pytest.mark.gpu
def test_hetg_use_this_order(gpu):
    t = Tracer(seed=14, GratingType="HETG", DetectorType="NONE", Use_This_Order=1)
    rec = on_axis_photons(50000, 1.5)
    pg, po = _run(t, rec)
    compare(pg.active, po.active, max_flips=2, label="hetg order 1")
    a = pg.active[(pg.active["flags"] & 0xFF) == 0]
    assert set(np.unique(a["order"]).tolist()) <= {-1, 1}
