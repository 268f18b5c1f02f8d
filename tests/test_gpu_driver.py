"""The drop-in route from C on the GPU: tests/driver/marx_driver.c (the
reference driver's chunk sequence, linked against libmarx_gpu.so with its
own parameter-file reader) traces one chunk; its records and pruned
sorted_index equal, record for record, the same chunk traced through the
Python binding of the same entry points with the same parameter file."""
import numpy as np
import pytest

from marx_amd import Photons, Tracer
from marx_amd._lib import ATTR_DTYPE
from marx_amd.sim import ParamFile
from tests.test_cpu_driver import build_driver, run_driver, write_par
from tests.test_gpu_batch import assert_same_records


@pytest.mark.gpu
@pytest.mark.parametrize("grating,det", [("HETG", "ACIS-S"), ("LETG", "HRC-S")])
def test_c_driver_equals_the_python_route(gpu, tmp_path, grating, det):
    n = 50_000
    exe = build_driver(tmp_path)
    par = write_par(str(tmp_path / "marx.par"), GratingType=grating, DetectorType=det, RandomSeed=5,
                    SourceFlux=0.01)
    out = str(tmp_path / "out.bin")
    print(f"driver case {grating}/{det} n={n}", flush=True)
    rc, lines, err = run_driver(exe, par, n, out)
    assert rc == 0, (lines, err)
    raw = open(out, "rb").read()
    m = int(np.frombuffer(raw[:4], np.uint32)[0])
    rec = np.frombuffer(raw[4:4 + m * ATTR_DTYPE.itemsize], ATTR_DTYPE)
    rest = raw[4 + m * ATTR_DTYPE.itemsize:]
    k = int(np.frombuffer(rest[:4], np.uint32)[0])
    idx = np.frombuffer(rest[4:4 + 4 * k], np.uint32)

    t = Tracer(par=ParamFile(par), seed=5, source=True)
    p = Photons(n)
    assert t.create_photons(p, n) == m
    t.reflect(p)
    t.diffract(p)
    t.detect(p)
    assert_same_records(rec, p.active.copy(), f"C driver vs Python, {grating}/{det}")
    assert np.array_equal(idx, p.sorted_index())
    assert k == int(((rec["flags"] & 0xFF) == 0).sum()) > 100
