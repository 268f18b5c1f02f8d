"""The ACIS subpixel (EDSER) tables of a CALDB ACISSUBPIX file, read with a
minimal FITS binary-table reader of the oracle's own.

TEST INFRASTRUCTURE ONLY (see oracle.py).  Follows acis_subpix.c:121-268:
extensions MARX_ACIS_SUBPIX_FI (table 0) and MARX_ACIS_SUBPIX_BI (table 1)
with rows (fltgrade, npoints, energy[], chipx_offset[], chipy_offset[]);
npoints <= 0 leaves the grade without offsets.  The flat layout is the one
orc_event_coords takes: for table t and grade g, sub_n[256 t + g] values
at sub_off[256 t + g] in subtab, then as many chip x and chip y offsets."""
import numpy as np

BLOCK = 2880


def _cards(buf, pos):
    """(header dict, data start) of the HDU at pos."""
    h = {}
    while True:
        blk = buf[pos:pos + BLOCK]
        pos += BLOCK
        done = False
        for j in range(0, BLOCK, 80):
            card = blk[j:j + 80].decode("ascii")
            key = card[:8].strip()
            if key == "END":
                done = True
                break
            if card[8:10] == "= ":
                v = card[10:].split("/")[0].strip() if not card[10:].strip().startswith("'") \
                    else card[10:].strip()[1:].split("'")[0].strip()
                h[key] = v
        if done:
            return h, pos


def read_bintables(path):
    """{HDUNAME (else EXTNAME): {column name (upper case): array (rows,) or
    (rows, repeat)}}."""
    buf = open(path, "rb").read()
    h, pos = _cards(buf, 0)
    out = {}
    while pos < len(buf):
        h, pos = _cards(buf, pos)
        rows, width = int(h["NAXIS2"]), int(h["NAXIS1"])
        size = rows * width
        if h.get("XTENSION") == "BINTABLE":
            fields, off = [], 0
            for k in range(1, int(h["TFIELDS"]) + 1):
                form = h[f"TFORM{k}"]
                rep = int(form[:-1]) if len(form) > 1 else 1
                code = {"J": ">i4", "E": ">f4", "D": ">f8", "I": ">i2", "B": "u1"}[form[-1]]
                fields.append((h[f"TTYPE{k}"].upper(), code, (rep,) if rep > 1 else ()))
                off += rep * np.dtype(code).itemsize
            assert off == width
            dt = np.dtype([(n, c, s) for n, c, s in fields])
            rec = np.frombuffer(buf[pos:pos + size], dtype=dt)
            key = h.get("HDUNAME", h.get("EXTNAME", ""))
            out[key] = {n: rec[n].astype(np.dtype(c).newbyteorder("=")) for n, c, _ in fields}
        pos += (size + BLOCK - 1) // BLOCK * BLOCK
    return out


def subpix_tables(path):
    """(subtab float32, sub_off uint32[512], sub_n uint32[512])."""
    tabs = read_bintables(path)
    sub_off = np.zeros(512, np.uint32)
    sub_n = np.zeros(512, np.uint32)
    parts, length = [], 0
    for t, ext in enumerate(("MARX_ACIS_SUBPIX_FI", "MARX_ACIS_SUBPIX_BI")):
        c = tabs[ext]
        for row in range(len(c["FLTGRADE"])):
            g, n = int(c["FLTGRADE"][row]), int(c["NPOINTS"][row])
            if not 0 <= g < 256:
                raise ValueError("Illegal flight grade in subpix column")
            if n <= 0:
                sub_n[256 * t + g] = 0
                continue
            if n == 1 or n > c["ENERGY"].shape[1]:
                raise ValueError(f"bad NPOINTS {n} in row {row + 1} of {ext}")
            v = np.concatenate([c["ENERGY"][row, :n], c["CHIPX_OFFSET"][row, :n],
                                c["CHIPY_OFFSET"][row, :n]]).astype(np.float32)
            sub_off[256 * t + g], sub_n[256 * t + g] = length, n
            parts.append(v)
            length += len(v)
    return np.concatenate(parts) if parts else np.zeros(0, np.float32), sub_off, sub_n
