"""marx's output files from AoS photon lists (marx_write_photons, the
dump-file reader and text dump, ray files) against the numpy restatement of
the formats in oracle/dumpfile.py (marx/libsrc/marxio.c, s-rayfile.c,
marx/doc/dump.txt): the same files, byte for byte."""
import ctypes
import os

import numpy as np
import pytest

from marx_amd import Tracer
from marx_amd._lib import ATTR_DTYPE, WRITE_ALL, load
from marx_amd.sim import Photons
from oracle import dumpfile as D

ON_AXIS = dict(SourceRA=250.2134679741175, SourceDEC=-53.75743813458669)
ALL_HISTORY = 0x01FFFFFF


def _records(n, seed, tag_start=0):
    rng = np.random.default_rng(seed)
    r = np.zeros(n, dtype=ATTR_DTYPE)
    r["energy"] = rng.uniform(0.1, 10.0, n)
    r["x"] = rng.normal(0, 30.0, (n, 3))
    r["p"] = rng.normal(0, 1.0, (n, 3))
    r["arrival_time"] = np.sort(rng.uniform(0, 1e4, n))
    # some photons absorbed (low byte), some only marked (0x100, 0x200)
    r["flags"] = rng.choice([0, 0, 0, 0x01, 0x08, 0x40, 0x100, 0x200], n)
    for k in ("y_pixel", "z_pixel", "u_pixel", "v_pixel", "pi"):
        r[k] = rng.uniform(-100, 1100, n).astype(np.float32)
    r["dither"] = rng.normal(0, 1e-3, (n, 6)).astype(np.float32)
    r["pulse_height"] = rng.integers(-5, 4096, n)
    r["mirror_shell"] = rng.choice([0, 1, 2, 3], n)
    r["ccd_num"] = rng.integers(-1, 10, n)
    r["detector_region"] = rng.integers(-1, 3, n)
    r["order"] = rng.integers(-10, 11, n)
    r["support_orders"] = rng.integers(-60, 61, (n, 4))
    r["tag"] = np.arange(tag_start, tag_start + n)
    return r


def _photons(records, history):
    p = Photons(max(len(records), 1))
    p.load(records)
    p.pt.history = history
    return p


@pytest.fixture(params=["NONE", "INTERNAL"])
def dither(request):
    Tracer(seed=3, source=True, GratingType="NONE", DetectorType="NONE", SourceType="POINT",
           SpectrumType="FLAT", MinEnergy=1.49, MaxEnergy=1.49, DitherModel=request.param, **ON_AXIS)
    return request.param != "NONE"


def _read_dir(d):
    return {f: open(os.path.join(d, f), "rb").read() for f in sorted(os.listdir(d))}


@pytest.mark.parametrize("history,mask", [
    (ALL_HISTORY, WRITE_ALL),
    (0x1F | 0x800, WRITE_ALL),                               # source + mirror stages only
    (ALL_HISTORY, D.ENERGY_OK | D.TIME_OK | D.X_OK | D.ORDER_OK | D.PHA_OK),
])
def test_columns_match_the_oracle(tmp_path, dither, history, mask):
    r = _records(3000, 1)
    p = _photons(r, history)
    p.write(str(tmp_path), write_mask=mask, open_mode=1, total_time=250.5)
    got = _read_dir(tmp_path)
    want = D.column_files([(r, history, 250.5)], mask, dither)
    assert sorted(got) == sorted(want)
    for f in want:
        assert got[f] == want[f], f


def test_append_updates_the_row_count(tmp_path, dither):
    r1, r2 = _records(1200, 2), _records(900, 3, tag_start=1200)
    p1, p2 = _photons(r1, ALL_HISTORY), _photons(r2, ALL_HISTORY)
    p1.write(str(tmp_path), open_mode=1, total_time=0.0)
    p2.write(str(tmp_path), open_mode=0, total_time=1234.0)
    got = _read_dir(tmp_path)
    want = D.column_files([(r1, ALL_HISTORY, 0.0), (r2, ALL_HISTORY, 1234.0)], WRITE_ALL, dither)
    assert got == want
    _, name, rows, cols, _ = D.read_column(got["energy.dat"])
    kept = ((r1["flags"] & 0xFF) == 0).sum() + ((r2["flags"] & 0xFF) == 0).sum()
    assert (name, rows, cols) == ("ENERGY", kept, 0)


def test_empty_list_writes_headers_only(tmp_path, dither):
    r = _records(50, 4)
    r["flags"] = 0x08
    _photons(r, ALL_HISTORY).write(str(tmp_path))
    for f, b in _read_dir(tmp_path).items():
        assert len(b) == 32 and D.read_column(b)[2] == 0, f


def test_read_dump_file_header(tmp_path):
    lib = load()
    r = _records(77, 5)
    _photons(r, ALL_HISTORY).write(str(tmp_path))
    d = lib.marx_open_read_dump_file(os.fsencode(tmp_path / "ocoarse3.dat"))
    assert d
    c = d.contents
    assert (c.type, c.colname, c.num_rows, c.num_cols) == (b"A", b"COARSE3", int(((r["flags"] & 0xFF) == 0).sum()), 0)
    lib.marx_close_read_dump_file(d)
    bad = tmp_path / "bad.dat"
    bad.write_bytes(b"\0" * 40)
    assert not lib.marx_open_read_dump_file(os.fsencode(bad))
    assert not lib.marx_open_read_dump_file(os.fsencode(tmp_path / "missing.dat"))


def test_created_header_matches_the_format(tmp_path):
    """marx_create_write_dump_file / marx_close_write_dump_file on their own:
    the header of marx/doc/dump.txt with the row count written at close."""
    lib = load()
    from marx_amd._lib import DumpFile
    dft = DumpFile()
    dft.type = b"D"
    dft.colname = b"SOME_LONG_NAME_X"[:15]
    dft.num_cols = 3
    path = os.fsencode(tmp_path / "c.dat")
    fp = lib.marx_create_write_dump_file(path, ctypes.byref(dft))
    assert fp and dft.fp == fp
    assert lib.marx_close_write_dump_file(ctypes.c_void_p(fp), 41) == 0
    assert open(path, "rb").read() == D.header("D", "SOME_LONG_NAME_", 41, 3)


def _drain_stdio(capfd):
    """Empties the C library's stdout buffer (messages of earlier calls)
    and the captured output, so that what follows is the call's own."""
    ctypes.CDLL(None).fflush(None)
    capfd.readouterr()


def test_dump_text(tmp_path, capfd):
    lib = load()
    r = _records(40, 6)
    _photons(r, ALL_HISTORY).write(str(tmp_path))
    names = ["energy.dat", "tag.dat", "order.dat", "pha.dat", "mirror.dat"]
    argv = (ctypes.c_char_p * (len(names) + 2))(b"marx", b"--dump", *[os.fsencode(tmp_path / f) for f in names])
    _drain_stdio(capfd)
    assert lib.marx_dump(len(names) + 2, argv) == 0
    out = capfd.readouterr().out
    files = [(f, open(tmp_path / f, "rb").read()) for f in names]
    assert out == D.dump_text(files)


def test_rayfile(tmp_path, capfd):
    r1, r2 = _records(500, 7), _records(300, 8, tag_start=500)
    p1, p2 = _photons(r1, 0x1F), _photons(r2, 0x1F)
    path = str(tmp_path / "rays.dat")
    p1.dump_to_rayfile(path, new_file=1, total_time=0.0)
    p2.dump_to_rayfile(path, new_file=0, total_time=5000.0)
    head, rows, mutated = D.rayfile([(r1, 0x1F, 0.0), (r2, 0x1F, 5000.0)])
    data = open(path, "rb").read()
    assert data[:16] == head and len(data) == 16 + len(rows) * ATTR_DTYPE.itemsize
    got = np.frombuffer(data[16:], dtype=ATTR_DTYPE)
    for k in ATTR_DTYPE.names:
        assert np.array_equal(got[k], rows[k]), k
    # the reference adds total_time to the written photons in the caller's list
    assert np.array_equal(p2.active["arrival_time"], mutated[1]["arrival_time"])
    lib = load()
    _drain_stdio(capfd)
    assert lib.marx_dump_rayfile(os.fsencode(path)) == 0
    lines = capfd.readouterr().out.splitlines()
    kept = ((r1["flags"] & 0xFF) == 0).sum() + ((r2["flags"] & 0xFF) == 0).sum()
    assert lines[0].startswith("#Energy") and len(lines) == 1 + kept
    assert lines[1].split("\t")[0] == "%f" % r1[(r1["flags"] & 0xFF) == 0][0]["energy"]


def test_append_without_files_fails(tmp_path):
    r = _records(10, 9)
    with pytest.raises(RuntimeError, match="Unable to open"):
        _photons(r, ALL_HISTORY).write(str(tmp_path), open_mode=0)
