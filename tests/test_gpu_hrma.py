"""HRMA reflection: GPU kernel vs the FP64 CPU oracle, through the C-ABI
(marx_mirror_reflect, replacing marx/libsrc/mirror.c:93-108 ->
hrma.c:1161-1341)."""
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
    Oracle(tracer).reflect(po)
    return pg, po


@pytest.mark.gpu
@pytest.mark.parametrize("seed", [1, 20241202])
def test_hrma_monochromatic_parity(gpu, seed):
    # Plain configuration: on-axis point source at 1.49 keV.
    t = Tracer(seed=seed, GratingType="NONE", DetectorType="NONE")
    rec = on_axis_photons(200000, 1.49, tag_start=1000)
    pg, po = _run(t, rec)
    s = compare(pg.active, po.active, max_flips=2, label=f"hrma seed={seed}")
    good = (pg.active["flags"] & 0xFF) == 0
    assert 0.30 < good.mean() < 0.37
    assert s["n"] == 200000


@pytest.mark.gpu
def test_hrma_broadband_offaxis_parity(gpu):
    # 0.3-10 keV, 5 arcmin off axis, struts and scattering on.
    t = Tracer(seed=7, GratingType="NONE", DetectorType="NONE")
    rng = np.random.default_rng(3)
    n = 100000
    rec = on_axis_photons(n, rng.uniform(0.3, 10.0, n))
    th = 5.0 / 60.0 * np.pi / 180.0
    rec["p"] = (-np.cos(th), np.sin(th) * 0.6, np.sin(th) * 0.8)
    pg, po = _run(t, rec)
    compare(pg.active, po.active, max_flips=2, label="hrma broadband")


@pytest.mark.gpu
def test_hrma_ideal_and_shutters(gpu):
    # HRMA_Ideal (no reflectivity draw, no blur, no scatter) with shells 3
    # and 6 partly shuttered: exercises the shutter rejection loop.
    t = Tracer(seed=5, GratingType="NONE", DetectorType="NONE", HRMA_Ideal="yes",
               Shutters3="1010", Shutters6="0111")
    rec = on_axis_photons(50000, 2.0)
    pg, po = _run(t, rec)
    compare(pg.active, po.active, max_flips=0, label="hrma ideal+shutters")


@pytest.mark.gpu
def test_hrma_empty_and_history(gpu):
    t = Tracer(seed=5, GratingType="NONE", DetectorType="NONE")
    pg = Photons(10)
    pg.load(on_axis_photons(0, 1.0))
    t.reflect(pg)
    assert pg.pt.num_sorted == 0
    # a second call on the same list is a no-op (history bit already set)
    rec = on_axis_photons(1000, 1.49)
    pg = Photons(1000)
    pg.load(rec)
    t.reflect(pg)
    snap = pg.active.copy()
    t.reflect(pg)
    assert np.array_equal(snap, pg.active)
