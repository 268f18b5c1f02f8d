"""Photon creation: GPU source kernels vs the CPU oracle through the C-ABI
(marx_create_photons, replacing marx/libsrc/source.c:268-384 with
s-point.c, s-disk.c, s-gauss.c, s-beta.c, s-line.c, s-image.c, spectrum.c and the
INTERNAL dither of dither.c), and
the whole chain source -> HRMA -> HETG -> ACIS-S, through the AoS entry
points and through the device-resident mxg_trace_device."""
import numpy as np
import pytest

from marx_amd import Photons, Tracer
from marx_amd.spectra import write_power_law
from oracle.oracle import Oracle
from tests.fitsimg import write_image
from tests.parity import compare

ON_AXIS = dict(SourceRA=250.2134679741175, SourceDEC=-53.75743813458669)


def _both(t, num, exposure=None):
    pg, po = Photons(num), Photons(num)
    ng = t.create_photons(pg, num, exposure)
    no = Oracle(t).create_photons(po, num, exposure)
    assert ng == no
    assert pg.pt.total_time == po.pt.total_time
    assert pg.pt.tag_start == po.pt.tag_start
    assert np.array_equal(pg.sorted_index(), po.sorted_index())
    return pg, po


@pytest.mark.gpu
@pytest.mark.parametrize("kind", ["flat-point-dither", "file-point", "flat-disk-dither", "uniform-times",
                                  "gauss-dither", "beta-offaxis", "line-file", "image-offaxis-dither"])
def test_source_parity(gpu, tmp_path, kind):
    extra = dict(ON_AXIS)
    utimes = 0.0
    if kind == "flat-point-dither":
        extra.update(SpectrumType="FLAT", MinEnergy=0.4, MaxEnergy=8.0, DitherModel="INTERNAL")
    elif kind == "file-point":
        f = write_power_law(str(tmp_path / "pl.dat"))
        extra.update(SpectrumType="FILE", SpectrumFile=f, DitherModel="NONE")
    elif kind == "flat-disk-dither":
        extra.update(SourceType="DISK", **{"S-DiskTheta0": 10.0, "S-DiskTheta1": 120.0},
                     SpectrumType="FLAT", MinEnergy=0.07, MaxEnergy=10.0, DitherModel="INTERNAL")
    elif kind == "gauss-dither":
        extra.update(SourceType="GAUSS", **{"S-GaussSigma": 45.0},
                     SpectrumType="FLAT", MinEnergy=0.3, MaxEnergy=9.0, DitherModel="INTERNAL")
    elif kind == "beta-offaxis":
        extra = dict(SourceRA=ON_AXIS["SourceRA"] + 0.04, SourceDEC=ON_AXIS["SourceDEC"] - 0.02,
                     SourceType="BETA", **{"S-BetaCoreRadius": 8.0, "S-BetaBeta": 0.67},
                     SpectrumType="FLAT", MinEnergy=0.5, MaxEnergy=7.0, DitherModel="NONE")
    elif kind == "line-file":
        f = write_power_law(str(tmp_path / "pl.dat"))
        extra = dict(SourceRA=ON_AXIS["SourceRA"] - 0.03, SourceDEC=ON_AXIS["SourceDEC"] + 0.01,
                     SourceType="LINE", **{"S-LineTheta": 900.0, "S-LinePhi": 35.0},
                     SpectrumType="FILE", SpectrumFile=f, DitherModel="NONE")
    elif kind == "image-offaxis-dither":
        rng = np.random.default_rng(3)
        img = rng.gamma(0.3, 1.0, (48, 64))
        f = write_image(str(tmp_path / "im.fits"), img, -32, cdelt=(-0.4 / 3600, 0.6 / 3600))
        extra = dict(SourceRA=ON_AXIS["SourceRA"] + 0.02, SourceDEC=ON_AXIS["SourceDEC"] - 0.01,
                     SourceType="IMAGE", **{"S-ImageFile": f}, SpectrumType="FLAT", MinEnergy=0.3,
                     MaxEnergy=8.0, DitherModel="INTERNAL")
    else:
        extra.update(SpectrumType="FLAT", MinEnergy=1.49, MaxEnergy=1.49, DitherModel="NONE")
        utimes = 1.0e4
    t = Tracer(seed=31, source=True, uniform_times=utimes, GratingType="NONE", DetectorType="NONE", **extra)
    pg, po = _both(t, 100000)
    a, b = pg.active, po.active
    compare(a, b, max_flips=0, check_pixels=False, label=f"source {kind}")
    assert np.array_equal(a["arrival_time"], b["arrival_time"])
    assert np.array_equal(a["dither"], b["dither"])


@pytest.mark.gpu
def test_source_exposure_cut_and_chunks(gpu):
    # 50 ks exposure cut, then a second chunk continuing start_time / tags
    t = Tracer(seed=32, source=True, GratingType="NONE", DetectorType="NONE",
               SpectrumType="FLAT", MinEnergy=1.0, MaxEnergy=2.0, DitherModel="INTERNAL", **ON_AXIS)
    num = 200000
    pg, po = Photons(num), Photons(num)
    o = Oracle(t)
    counts = []
    for exposure in (3000.0, 50000.0):
        ng = t.create_photons(pg, num, exposure)
        counts.append(ng)
        no = o.create_photons(po, num, exposure)
        assert ng == no
        assert pg.pt.start_time == po.pt.start_time
        compare(pg.active, po.active, max_flips=0, check_pixels=False, label=f"exposure {exposure}")
        assert pg.active["arrival_time"][-1] >= exposure or ng == num
    assert counts[0] < num      # the first chunk is cut by the exposure


@pytest.mark.gpu
@pytest.mark.parametrize("workload", ["hetg-acis-s", "letg-hrc-s"])
def test_full_chain_aos_and_device(gpu, tmp_path, workload):
    if workload == "hetg-acis-s":
        # HETG + ACIS-S with a power-law FILE spectrum and INTERNAL dither
        f = write_power_law(str(tmp_path / "pl.dat"))
        t = Tracer(seed=33, source=True, GratingType="HETG", DetectorType="ACIS-S",
                   SpectrumType="FILE", SpectrumFile=f, DitherModel="INTERNAL",
                   DitherAmp_RA=16, DitherAmp_Dec=16, DitherPeriod_RA=1000, DitherPeriod_Dec=707,
                   **ON_AXIS)
    else:
        # LETG + HRC-S, 2 arcmin uniform disk, flat 0.07-10 keV, INTERNAL dither
        t = Tracer(seed=34, source=True, GratingType="LETG", DetectorType="HRC-S", SourceType="DISK",
                   SpectrumType="FLAT", MinEnergy=0.07, MaxEnergy=10.0, DitherModel="INTERNAL",
                   **{"S-DiskTheta0": 0.0, "S-DiskTheta1": 120.0}, **ON_AXIS)
    num = 200000
    pg, po = Photons(num), Photons(num)
    t.create_photons(pg, num)
    t.reflect(pg)
    t.diffract(pg)
    t.detect(pg)
    o = Oracle(t)
    o.create_photons(po, num)
    o.reflect(po)
    o.diffract(po)
    o.detect(po)
    compare(pg.active, po.active, max_flips=2, label=f"chain AoS {workload}")

    # the same chunk traced without leaving HBM
    ncol, ndet, ms = t.trace_device(num, tag_start=0, start_time=0.0, timing=True)
    assert ncol == num
    pd = Photons(num)
    t.download(pd)
    compare(pd.active, po.active, max_flips=2, label=f"chain device {workload}")
    assert ndet == int(((po.active["flags"] & 0xFF) == 0).sum())
    assert all(m >= 0 for m in ms)


@pytest.mark.gpu
@pytest.mark.parametrize("args,dither,exposure", [("times stop=9.999", "INTERNAL", 50.0), ("", "NONE", None),
                                                  ("stop=9.999", "INTERNAL", 300.0)])
def test_user_source_parity(gpu, tmp_path, args, dither, exposure):
    """A USER plug-in (marx_create_photons over s-user.c): energies drawn on
    the device, rays from the plug-in on the host, then cut, tags, dither
    and sort, against the oracle calling the same plug-in.  The plug-in's
    own times count from 0 while the cut is at the observation start plus
    the exposure (source.c:295-309): TStart = 2100 reads as 2100 s from
    MJDref, within reach of the plug-in's clock, so the cut is exercised."""
    import ctypes
    from tests.user_plugin import build
    path, plug = build(str(tmp_path))
    t = Tracer(seed=5, GratingType="NONE", DetectorType="NONE", SourceType="USER", SpectrumType="FLAT",
               MinEnergy=1.0, MaxEnergy=10.0, DitherModel=dither, UserSourceFile=path, UserSourceArgs=args,
               SourceRA=ON_AXIS["SourceRA"] + 0.01, SourceDEC=ON_AXIS["SourceDEC"], TStart=2100.0)
    t.init_source(SourceFlux=20.0)
    tstart = ctypes.c_double.in_dll(t.lib, "_Marx_TStart_MJDsecs").value
    assert tstart == 2100.0
    n = 20000
    pg, po = Photons(n), Photons(n)
    ng = t.create_photons(pg, n, exposure)
    no = Oracle(t).create_user(po, n, plug, exposure_time=exposure, tstart=tstart)
    if "times" in args:
        assert no < n                          # the cut ends the chunk
    assert ng == no > 50
    assert np.array_equal(pg.sorted_index(), po.sorted_index())
    a, b = pg.active, po.active
    compare(a, b, max_flips=0, check_pixels=False, label=f"user source {args}")
    assert np.array_equal(a["arrival_time"], b["arrival_time"])
    assert np.array_equal(a["dither"], b["dither"])
    assert pg.pt.total_time == po.pt.total_time
    # a device-created trace refuses a plug-in source
    with pytest.raises(RuntimeError, match="USER"):
        t.trace_device(1000)
