"""marx's output-file formats restated with numpy.  TEST INFRASTRUCTURE ONLY:
tests compare the library's files with the bytes made here; nothing in the
product imports this module.

Column ("dump") files, marx/libsrc/marxio.c:66-200, 230-453 and
marx/doc/dump.txt: per column one file of a 32-byte header (magic
0x83 0x13 0x89 0x8D, the type character, the column name in 15 bytes padded
with NUL, the number of rows and of columns as big-endian int32, 4 reserved
bytes) and the values big-endian; the columns and their order are the table
of marxio.c:259-289; a photon is written when flags & 0xFF is zero
(marxio.c:425-435); a column when its history bit is in the write mask and
the list's history, the two dither bit groups also when dither is on
(marxio.c:405-412).  The byte order follows jdmath's writers, which swap to
big-endian on little-endian hosts (jdmath/src/binaryio.c:375-420,
jdmath/src/sysconf.c:73-78).

Ray files, marx/libsrc/s-rayfile.c:43-145: the magic 0xF8161961 and the
history as native unsigned longs, then the native 136-byte records of the
photons whose flags & 0xFF is zero, their arrival_time plus total_time.
"""
import numpy as np

MAGIC = bytes([0x83, 0x13, 0x89, 0x8D])
RAYFILE_MAGIC = 0xF8161961

PI_OK, ENERGY_OK, TIME_OK, X_OK, P_OK, TAG_OK = 0x40, 0x1, 0x2, 0x4, 0x8, 0x10
PHA_OK, DET_PIXEL_OK, DET_NUM_OK, DET_REGION_OK, UV_OK = 0x20, 0x80, 0x100, 0x200, 0x400
SHELL_OK, SKY_DITHER_OK, DET_DITHER_OK = 0x800, 0x1000, 0x2000
ORDER_OK, ORDER1_OK, ORDER2_OK, ORDER3_OK, ORDER4_OK = 0x100000, 0x200000, 0x400000, 0x800000, 0x1000000

DTYPES = {"E": ">f4", "J": ">i4", "I": ">i2", "A": "i1"}


def _f(name, k=None):
    return lambda r, t: (r[name] if k is None else r[name][:, k]).astype(np.float32)


# (file, column name, type, history bit, value of records r at total time t)
COLUMNS = [
    ("b_energy.dat", "B_ENERGY", "E", PI_OK, _f("pi")),
    ("energy.dat", "ENERGY", "E", ENERGY_OK, _f("energy")),
    ("time.dat", "TIME", "E", TIME_OK, lambda r, t: (r["arrival_time"] + t).astype(np.float32)),
    ("tag.dat", "TAG", "J", TAG_OK, lambda r, t: r["tag"].astype(np.int64).astype(np.int32)),
    ("xpos.dat", "XPOS", "E", X_OK, _f("x", 0)),
    ("ypos.dat", "YPOS", "E", X_OK, _f("x", 1)),
    ("zpos.dat", "ZPOS", "E", X_OK, _f("x", 2)),
    ("xcos.dat", "COSX", "E", P_OK, _f("p", 0)),
    ("ycos.dat", "COSY", "E", P_OK, _f("p", 1)),
    ("zcos.dat", "COSZ", "E", P_OK, _f("p", 2)),
    ("pha.dat", "PHA", "I", PHA_OK, lambda r, t: r["pulse_height"]),
    ("detector.dat", "CCDID", "A", DET_NUM_OK, lambda r, t: r["ccd_num"]),
    ("xpixel.dat", "CHIPX", "E", DET_PIXEL_OK, _f("y_pixel")),
    ("ypixel.dat", "CHIPY", "E", DET_PIXEL_OK, _f("z_pixel")),
    ("hrc_u.dat", "U", "E", UV_OK, _f("u_pixel")),
    ("hrc_v.dat", "V", "E", UV_OK, _f("v_pixel")),
    ("mirror.dat", "MIRROR", "I", SHELL_OK, lambda r, t: r["mirror_shell"].astype(np.int64).astype(np.int16)),
    ("hrcregion.dat", "HRCREGION", "A", DET_REGION_OK, lambda r, t: r["detector_region"]),
    ("order.dat", "ORDER", "A", ORDER_OK, lambda r, t: r["order"]),
    ("ofine.dat", "FINE", "A", ORDER1_OK, lambda r, t: r["support_orders"][:, 0]),
    ("ocoarse1.dat", "COARSE1", "A", ORDER2_OK, lambda r, t: r["support_orders"][:, 1]),
    ("ocoarse2.dat", "COARSE2", "A", ORDER3_OK, lambda r, t: r["support_orders"][:, 2]),
    ("ocoarse3.dat", "COARSE3", "A", ORDER4_OK, lambda r, t: r["support_orders"][:, 3]),
    ("sky_ra.dat", "RA", "E", SKY_DITHER_OK, _f("dither", 0)),
    ("sky_dec.dat", "DEC", "E", SKY_DITHER_OK, _f("dither", 1)),
    ("sky_roll.dat", "ROLL", "E", SKY_DITHER_OK, _f("dither", 2)),
    ("det_dy.dat", "DET_DY", "E", DET_DITHER_OK, _f("dither", 3)),
    ("det_dz.dat", "DET_DZ", "E", DET_DITHER_OK, _f("dither", 4)),
    ("det_theta.dat", "DET_THETA", "E", DET_DITHER_OK, _f("dither", 5)),
]


def header(type_char, name, rows, cols=0):
    h = bytearray(32)
    h[0:4] = MAGIC
    h[4] = ord(type_char)
    nb = name.encode()[:15]
    h[5:5 + len(nb)] = nb
    h[20:24] = int(rows).to_bytes(4, "big", signed=True)
    h[24:28] = int(cols).to_bytes(4, "big", signed=True)
    return bytes(h)


def effective_mask(write_mask, history, dither_on):
    if dither_on:
        return write_mask & (history | SKY_DITHER_OK | DET_DITHER_OK)
    return write_mask & history


def written(records):
    return records[(records["flags"] & 0xFF) == 0]


def column_files(chunks, write_mask, dither_on):
    """{file name: bytes} of writing the chunks [(records, history,
    total_time), ...] into one directory, the first creating the files and
    the others appending (marx.c:587 with open_mode 1, then 0)."""
    out = {}
    mask0 = effective_mask(write_mask, chunks[0][1], dither_on)
    for fname, name, typ, bit, value in COLUMNS:
        if not (bit & mask0):
            continue
        body, rows = b"", 0
        for records, history, total_time in chunks:
            if not (bit & effective_mask(write_mask, history, dither_on)):
                continue
            r = written(records)
            body += np.ascontiguousarray(value(r, total_time)).astype(DTYPES[typ]).tobytes()
            rows += len(r)
        out[fname] = header(typ, name, rows) + body
    return out


def read_column(data):
    """(type, name, rows, cols, values) of a column file's bytes."""
    assert data[:4] == MAGIC
    typ = chr(data[4])
    name = data[5:20].split(b"\0")[0].decode()
    rows = int.from_bytes(data[20:24], "big", signed=True)
    cols = int.from_bytes(data[24:28], "big", signed=True)
    return typ, name, rows, cols, np.frombuffer(data[32:], dtype=DTYPES[typ])


def dump_text(files):
    """The text of marxio.c:506-620 ('marx --dump') for column files given as
    [(name, bytes), ...] of vectors."""
    cols = [read_column(b) for _, b in files]
    lines = ["#" + "".join("%16s" % c[1] for c in cols)]
    n = min(len(c[4]) for c in cols) if cols else 0
    fmt = {"A": "%16d", "I": "%16d", "J": "%16d", "E": "%16e", "D": "%16e"}
    for i in range(n):
        lines.append("".join(fmt[c[0]] % (c[4][i].item()) for c in cols))
    # the row whose first read fails still ends with a newline
    return "\n".join(lines) + "\n" + "\n"


def rayfile(chunks):
    """A ray file written from chunks [(records, history, total_time), ...],
    the first creating it: (its 16 header bytes, the records it holds, the
    records as left in the callers' lists, arrival_time += total_time for
    those written).  The records' padding bytes are not part of the format."""
    head = np.array([RAYFILE_MAGIC, chunks[0][1]], dtype="<u8").tobytes()
    rows, mutated = [], []
    for records, history, total_time in chunks:
        r = records.copy()
        keep = (r["flags"] & 0xFF) == 0
        r["arrival_time"][keep] += total_time
        rows.append(r[keep])
        mutated.append(r)
    return head, np.concatenate(rows), mutated
