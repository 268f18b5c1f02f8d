"""Batched point sources (mxg_trace_batch): one device pass over many sources
with uneven photon counts, against the CPU oracle; the same photons as a
per-source loop; independent of the launch schedule, staged or persistent;
HIP-graph replays equal to eager passes; the autotuner installs a valid
schedule."""
import numpy as np
import pytest

from marx_amd import Photons, Tracer, offaxis_directions
from marx_amd.spectra import write_power_law
from oracle.oracle import Oracle
from tests.parity import compare

ON_AXIS = dict(SourceRA=250.2134679741175, SourceDEC=-53.75743813458669)
KERNELS = ("source", "enter", "reflect", "grating", "detect")


@pytest.fixture(autouse=True)
def keep_schedule(gpu):
    """Every test here leaves the library's launch schedule as it found it
    (the default, for the tests that follow)."""
    import ctypes
    from marx_amd._lib import Schedule
    s = Schedule()
    assert gpu.mxg_get_schedule(ctypes.byref(s)) == 0
    yield
    assert gpu.mxg_set_schedule(ctypes.byref(s)) == 0


def _batch(tmp_path, nsrc, lo=100, hi=5000, seed=41, dither="NONE", grating="NONE", det="ACIS-I",
           shape="POINT", **shape_par):
    spec = write_power_law(str(tmp_path / "pl.dat"), gamma=2.0, emin=0.3, emax=10.0)
    t = Tracer(seed=seed, GratingType=grating, DetectorType=det, SourceType=shape, SpectrumType="FILE",
               SpectrumFile=spec, DitherModel=dither, **shape_par, **ON_AXIS)
    t.init_source()
    rng = np.random.default_rng(nsrc)
    dirs = offaxis_directions(np.radians(rng.uniform(0, 15, nsrc) / 60), rng.uniform(0, 2 * np.pi, nsrc))
    counts = np.round(10 ** rng.uniform(np.log10(lo), np.log10(hi), nsrc)).astype(np.uint32)
    return t, dirs, counts


def _download(t, n):
    p = Photons(n)
    t.download(p)
    return p.active.copy()


def assert_same_records(a, b, what):
    """Bit-identical photon records, field by field (numpy copies of this
    dtype do not carry its padding bytes); on failure names the fields and
    the first differing photon."""
    assert len(a) == len(b), what
    bad = {}
    for f in a.dtype.names:
        x = np.ascontiguousarray(a[f]).view(np.uint8).reshape(len(a), -1)
        y = np.ascontiguousarray(b[f]).view(np.uint8).reshape(len(b), -1)
        i = np.nonzero(np.any(x != y, axis=1))[0]
        if len(i):
            bad[f] = (len(i), int(i[0]))
    assert not bad, f"{what}: records differ (field: (photons, first index)) {bad}"


def _oracle(t, dirs, counts, tag_start, start_time, exposure, grating):
    po = Photons(int(counts.sum()))
    o = Oracle(t)
    o.create_batch(po, dirs, counts, tag_start, start_time, exposure)
    o.reflect(po)
    if grating != "NONE":
        o.diffract(po)
    o.detect(po)
    return po.active.copy()         # the view would dangle once po is freed


@pytest.mark.gpu
@pytest.mark.parametrize("grating,det,dither", [("NONE", "ACIS-I", "NONE"), ("HETG", "ACIS-S", "INTERNAL")])
def test_batch_parity(gpu, tmp_path, grating, det, dither):
    t, dirs, counts = _batch(tmp_path, 48, grating=grating, det=det, dither=dither)
    n = int(counts.sum())
    ndet, ms = t.trace_batch(dirs, counts, tag_start=1000, start_time=500.0, exposure=3.0e4,
                             use_graph=False)
    g = _download(t, n)
    o = _oracle(t, dirs, counts, 1000, 500.0, 3.0e4, grating)
    compare(g, o, max_flips=2, label=f"batch {grating}/{det}")
    assert ndet == int(((g["flags"] & 0xFF) == 0).sum())
    assert t.stage_counts()[0] == n


@pytest.mark.gpu
@pytest.mark.parametrize("shape,par", [("DISK", {"S-DiskTheta0": 5.0, "S-DiskTheta1": 40.0}),
                                       ("GAUSS", {"S-GaussSigma": 20.0}),
                                       ("BETA", {"S-BetaCoreRadius": 6.0, "S-BetaBeta": 0.7}),
                                       ("LINE", {"S-LineTheta": 300.0, "S-LinePhi": 60.0})])
def test_batch_of_extended_sources(gpu, tmp_path, shape, par):
    """Every source of the batch with the configured shape about its own
    direction (s-disk.c, s-gauss.c, s-beta.c, s-line.c)."""
    t, dirs, counts = _batch(tmp_path, 40, shape=shape, **par)
    n = int(counts.sum())
    t.trace_batch(dirs, counts, tag_start=9, start_time=0.0, exposure=2.0e4, use_graph=False)
    g = _download(t, n)
    o = _oracle(t, dirs, counts, 9, 0.0, 2.0e4, "NONE")
    compare(g, o, max_flips=2, label=f"batch {shape}")


@pytest.mark.gpu
def test_per_source_loop_gives_the_same_photons(gpu, tmp_path):
    t, dirs, counts = _batch(tmp_path, 12)
    n = int(counts.sum())
    t.trace_batch(dirs, counts, tag_start=77, exposure=1.0e4, use_graph=False)
    whole = _download(t, n)
    parts, off = [], 77
    for k in range(len(counts)):
        t.trace_batch(dirs[k:k + 1], counts[k:k + 1], tag_start=off, exposure=1.0e4, use_graph=False)
        parts.append(_download(t, int(counts[k])))
        off += int(counts[k])
    assert_same_records(np.concatenate(parts), whole, "per-source loop vs batch")


@pytest.mark.gpu
def test_batch_is_independent_of_the_schedule(gpu, tmp_path):
    t, dirs, counts = _batch(tmp_path, 64)
    n = int(counts.sum())
    outs = []
    for bpc, waves in ((2, 2), (16, 4), (5, 3)):
        t.set_schedule(blocks_per_cu=bpc, reflect_waves=waves)
        t.trace_batch(dirs, counts, tag_start=5, exposure=1.0e4, use_graph=False)
        outs.append(_download(t, n))
    assert_same_records(outs[0], outs[1], "schedule (2, 2) vs (16, 4)")
    assert_same_records(outs[0], outs[2], "schedule (2, 2) vs (5, 3)")


@pytest.mark.gpu
def test_graph_replays_equal_eager_passes(gpu, tmp_path):
    t, dirs, counts = _batch(tmp_path, 40)
    n = int(counts.sum())
    for rep, tag in enumerate((0, 123456, 0, 999)):
        ndet_g, _ = t.trace_batch(dirs, counts, tag_start=tag, exposure=2.0e4, use_graph=True)
        g = _download(t, n)
        ndet_e, _ = t.trace_batch(dirs, counts, tag_start=tag, exposure=2.0e4, use_graph=False)
        e = _download(t, n)
        assert ndet_g == ndet_e, f"replay {rep}"
        assert_same_records(g, e, f"replay {rep} (tag_start {tag})")
        assert t.stage_counts()[3] == ndet_e


@pytest.mark.gpu
@pytest.mark.parametrize("grating,det", [("NONE", "ACIS-I"), ("HETG", "ACIS-S")])
def test_autotuner_installs_a_valid_schedule(gpu, tmp_path, grating, det):
    t, dirs, counts = _batch(tmp_path, 32, grating=grating, det=det)
    n = int(counts.sum())
    t.trace_batch(dirs, counts, use_graph=False)
    ref = _download(t, n)
    sched, ms = t.autotune_batch(dirs, counts, repeats=1)
    assert sched == t.get_schedule()
    assert all(1 <= sched[k] <= 32 for k in ("source", "enter", "reflect", "detect"))
    assert 0 <= sched["grating"] <= 32
    assert sched["reflect_waves"] in (2, 3, 4)
    assert sched["persistent_tile"] in (0, 256, 512, 1024, 2048)
    assert ms > 0
    t.trace_batch(dirs, counts, use_graph=True)
    assert_same_records(_download(t, n), ref, "autotuned graph pass")


@pytest.mark.gpu
def test_graph_is_recaptured_after_the_store_grows(gpu, tmp_path):
    """A captured batch graph holds the addresses of the photon store and of
    the batch's source arrays.  A device trace larger than the store (while
    the store can still grow: one launch takes at most 2^26 photons), and a
    batch of more sources than any before, reallocate them; the next graph
    pass with the first sources must not replay kernels into the freed
    buffers but capture again, and give the eager pass's records."""
    import ctypes
    from marx_amd._lib import SoaView
    t, dirs, counts = _batch(tmp_path, 24)
    n = int(counts.sum())
    t.trace_batch(dirs, counts, tag_start=3, exposure=2.0e4, use_graph=True)
    first = _download(t, n)
    view = SoaView()
    assert t.lib.mxg_get_soa(ctypes.byref(view)) == 0
    big = int(view.capacity) + 1
    if big <= 1 << 26:
        t.trace_device(big, tag_start=0)
        assert t.lib.mxg_get_soa(ctypes.byref(view)) == 0
        assert int(view.capacity) >= big
    many = 20000
    rng = np.random.default_rng(5)
    many_dirs = offaxis_directions(np.radians(rng.uniform(0, 5, many) / 60), rng.uniform(0, 2 * np.pi, many))
    t.trace_batch(many_dirs, np.ones(many, dtype=np.uint32), tag_start=0, exposure=2.0e4, use_graph=False)
    t.trace_batch(dirs, counts, tag_start=3, exposure=2.0e4, use_graph=True)
    again = _download(t, n)
    t.trace_batch(dirs, counts, tag_start=3, exposure=2.0e4, use_graph=False)
    eager = _download(t, n)
    assert_same_records(again, eager, "graph pass after the reallocations")
    assert_same_records(first, eager, "graph pass before the reallocations")


@pytest.mark.gpu
@pytest.mark.parametrize("grating,det,dither", [("NONE", "ACIS-I", "NONE"), ("HETG", "ACIS-S", "INTERNAL"),
                                                ("LETG", "HRC-S", "INTERNAL"), ("NONE", "HRC-I", "NONE"),
                                                ("HETG", "NONE", "NONE"), ("LETG", "ACIS-S", "NONE"),
                                                ("HETG", "HRC-I", "INTERNAL")])
def test_persistent_pass_equals_the_staged_pass(gpu, tmp_path, grating, det, dither):
    """The batched pass as one persistent launch (blocks taking tiles of
    photons from a work counter, every stage on a tile back to back) gives
    the staged pass's records, detected list and stage counts, for tiles of
    256 to 4096 photons at 1 to 8 blocks per CU, eager and graph-launched.
    The staged pass hands survivors on through the dense relays (grating to
    detector; mirror to an LETG grating) while the persistent pass reads
    the store, so the combinations also check every detector family's and
    both gratings' relay instances against the store path."""
    t, dirs, counts = _batch(tmp_path, 48, grating=grating, det=det, dither=dither)
    n = int(counts.sum())
    t.set_schedule(blocks_per_cu=8, reflect_waves=4, persistent_tile=0)
    ndet, _ = t.trace_batch(dirs, counts, tag_start=7, exposure=2.0e4, use_graph=False)
    ref = _download(t, n)
    ref_counts = t.stage_counts()
    assert ref_counts[3] == ndet
    try:
        for tile, bpc in ((256, 1), (1024, 2), (4096, 4), (512, 8)):
            t.set_schedule(persistent_tile=tile, persistent_blocks_per_cu=bpc)
            for graph in (False, True, True):
                print(f"persistent {grating}/{det}: tile {tile} x {bpc}, graph {graph}", flush=True)
                nd, _ = t.trace_batch(dirs, counts, tag_start=7, exposure=2.0e4, use_graph=graph)
                assert nd == ndet
                assert t.stage_counts() == ref_counts
                assert_same_records(_download(t, n), ref, f"persistent tile {tile} x {bpc}, graph {graph}")
    finally:
        t.set_schedule(persistent_tile=0)
